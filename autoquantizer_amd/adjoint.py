"""Reverse mode through MPM substeps: checkpointed recomputation around the
hand-written adjoint kernels (aq_adj_substep), a torch.autograd.Function,
and the reference's error-contribution sums.

The reference runs its AD pass with Taichi's Tape and an O(n log n)
checkpoint schedule (mpm_exps/engine/mpm_solver.py:918-931, 1121-1188) and
accumulates, per quantized component, the squared adjoints of the state over
every intermediate substep (accu_grad :1232-1271, get_gradients :1273-1302).
Here the forward keeps a checkpoint every `checkpoint_every` substeps; the
backward re-runs each segment forward from its checkpoint, keeps the
segment's states, and walks them in reverse with aq_adj_substep.  The
adjoint runs on plain storage (f32 or f64, as the reference's AD runs on
unquantized fields); the sums are reduced in f64.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .mpm import MPMSolver, _ptr, _stream


class SubstepAdjoint:
    """The adjoint of one substep of a plain-storage MPMSolver."""

    def __init__(self, solver: MPMSolver):
        if solver.storage == _lib.AQ_STORE_QUANT:
            raise ValueError("the adjoint runs on plain (f32/f64) storage")
        self.solver = solver
        nbytes = _lib.lib().aq_adj_workspace_bytes(solver._h)
        if nbytes < 0:
            _lib.check(-1)
        self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device=solver.device)

    def step(self, state_t, adj_next, adj_out, gsum=None, step=0):
        """adj_out = dL/ds_t from adj_next = dL/ds_{t+1}; all [ncomp, capacity]
        tensors of the solver's dtype.  gsum (float64 [ncomp] on the device):
        sum_p adj_out^2 is added to it.  step: the substep s_t -> s_{t+1}
        ran at (the collider's schedule)."""
        s = self.solver
        _lib.check(_lib.lib().aq_adj_substep(s._h, int(step), _ptr(state_t), s.current_idx, _ptr(adj_next),
                                             _ptr(adj_out), _ptr(gsum) if gsum is not None else None,
                                             _ptr(self.ws), self.ws.numel(), _stream()))


def energy_loss(solver, state):
    """compute_ek (mpm_solver.py:1043-1052) of a [ncomp, >= n] state tensor,
    differentiable in torch: 3D sum(1/2 m |v|^2 + g m y), 2D sum |v|^2."""
    n, d = solver.current_idx, solver.dim
    v = state[d:2 * d, :n].double()
    if d == 3:
        return (0.5 * solver.p_mass * (v * v).sum() + solver.param.g * solver.p_mass * state[1, :n].double().sum())
    return (v * v).sum()


def _advance_plain(solver, state, nsteps, step0=0):
    """nsteps plain substeps of `state` (written in place; LDS-tiled when the
    solver's last sort sorted this buffer, else untiled), the first at
    substep step0 (the collider's schedule)."""
    if nsteps <= 0:
        return
    _lib.check(_lib.lib().aq_advance(solver._h, _ptr(state), solver.current_idx, int(nsteps), int(step0), 0, None,
                                     _stream()))


class Noise:
    """Modelled rounding noise: after substep t (1-based) the state gets
    sigma[c] * xi(seed, t, c, id), xi uniform in [-1/2, 1/2), keyed by the
    particle id (aq_noise_add), so that a sorted copy draws the same noise
    per particle as the unsorted state."""

    def __init__(self, sigma, seed=0):
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        self.seed = int(seed)

    def add(self, solver, state, step, ids=None):
        _lib.check(_lib.lib().aq_noise_add(solver._h, _ptr(state), solver.current_idx,
                                           _ptr(ids) if ids is not None else None,
                                           self.sigma.ctypes.data_as(ctypes.c_void_p), self.seed, int(step),
                                           _stream()))

    def dot(self, solver, adj, step, out, ids=None):
        _lib.check(_lib.lib().aq_noise_dot(solver._h, _ptr(adj), solver.current_idx,
                                           _ptr(ids) if ids is not None else None, self.seed, int(step), _ptr(out),
                                           _stream()))


def _sorted_copy(solver, state):
    """A copy of `state` sorted by grid block (aq_sort), so that the tiled
    kernels run on it; returns (copy, original index of each sorted slot
    (int64 [n]), particle ids of the slots (int32 [capacity]))."""
    n = solver.current_idx
    w = state.clone()
    ids = torch.arange(w.shape[1], dtype=torch.int32, device=w.device)
    _lib.check(_lib.lib().aq_sort(solver._h, _ptr(w), n, _ptr(ids), _stream()))
    return w, ids[:n].long(), ids


class _Schedule:
    """Checkpointed forward and reverse passes.  Each segment runs on a
    sorted copy of its checkpoint (LDS-tiled kernels, like the forward
    solver); states, checkpoints and adjoints at segment boundaries stay in
    the original particle order, and the adjoint is carried through the
    segment's permutation.  Noise is keyed by particle id, so it does not
    depend on the order.

    adjoint_solver (optional): a solver of the same scene in another
    precision that runs the reverse pass (e.g. the fp32 adjoint kernels
    around an f64 forward): each of its segments sorts a copy of the
    checkpoint with its own sort, and the forward's states are cast and
    permuted into that order.

    Substep t of the schedule (s_t -> s_{t+1}) runs at the solver's
    step_count + t, as advance() would run it (the collider's schedule).

    grad_clip (use_grad_clip, accu_grad at mpm_solver.py:1234-1252): the
    squared x adjoint of state t+1, particle p, axis k enters the sums only
    where it differs from state t's by at most grad_clip."""

    def __init__(self, solver, nsteps, checkpoint_every, noise=None, adjoint_solver=None, grad_clip=None):
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        if noise is not None and len(noise.sigma) != solver.ncomp:
            raise ValueError("sigma needs one entry per component")
        if adjoint_solver is not None and (adjoint_solver.ncomp != solver.ncomp
                                           or adjoint_solver.current_idx != solver.current_idx
                                           or adjoint_solver.capacity != solver.capacity):
            raise ValueError("the adjoint solver must hold the same scene")
        self.solver, self.nsteps, self.k = solver, int(nsteps), int(checkpoint_every)
        self.noise = noise
        self.adj_solver = adjoint_solver if adjoint_solver is not None else solver
        self.checkpoints = []
        self.step0 = int(solver.step_count)
        self.grad_clip = None if grad_clip is None else float(grad_clip)

    def _advance(self, s, ids, t0, nsteps):
        """nsteps substeps of s (sorted, particle ids `ids`), the first
        producing state t0 + 1."""
        if self.noise is None:
            _advance_plain(self.solver, s, nsteps, self.step0 + t0)
            return
        for t in range(t0 + 1, t0 + nsteps + 1):
            _advance_plain(self.solver, s, 1, self.step0 + t - 1)
            self.noise.add(self.solver, s, t, ids)

    def forward(self, state0):
        """Runs the forward from state0 ([ncomp, capacity]); returns s_T."""
        s = state0.clone()
        n = self.solver.current_idx
        for t0 in range(0, self.nsteps, self.k):
            self.checkpoints.append(s.clone())
            steps = min(self.k, self.nsteps - t0)
            w, idx, ids = _sorted_copy(self.solver, s)
            self._advance(w, ids, t0, steps)
            s[:, idx] = w[:, :n]
        return s

    def backward(self, adj_T, gsum=None, adj=None, noise_dot=None):
        """dL/ds_0 from dL/ds_T (tensors of the adjoint solver's dtype),
        adding the squared adjoints of states 1..T-1 into gsum and, with
        noise, sum_p dL/ds_t xi_t (t = 1..T) into noise_dot (dL/dsigma)."""
        a = adj_T.clone()
        b = torch.empty_like(a)
        n = self.solver.current_idx
        d = self.solver.dim
        mixed = self.adj_solver is not self.solver
        clip = self.grad_clip is not None and gsum is not None
        xsum = torch.zeros(d, dtype=torch.float64, device=a.device) if clip else None
        for j in reversed(range(len(self.checkpoints))):
            t0 = j * self.k
            t1 = min(t0 + self.k, self.nsteps)
            # the segment's states, in the order of the forward's sorted copy
            w, idx, ids = _sorted_copy(self.solver, self.checkpoints[j])
            states = [w.clone()]
            for t in range(t0 + 1, t1):
                self._advance(w, ids, t - 1, 1)
                states.append(w.clone())
            if mixed:
                # the adjoint solver's own sorted copy, and where each of its slots sits in the forward's order
                wa, idx_a, ids_a = _sorted_copy(self.adj_solver, self.checkpoints[j].to(a.dtype))
                slot = torch.empty_like(idx)
                slot[idx] = torch.arange(n, device=idx.device)
                gather = slot[idx_a]
            else:
                wa, idx_a, ids_a = w, idx, ids
            a[:, :n] = a[:, idx_a]
            for t in reversed(range(t0, t1)):
                if noise_dot is not None:
                    self.noise.dot(self.adj_solver, a, t + 1, noise_dot, ids_a)
                st = states[t - t0]
                if mixed:
                    wa[:, :n] = st[:, gather].to(a.dtype)
                else:
                    wa.copy_(st)
                adj.step(wa, a, b, gsum if t >= 1 else None, self.step0 + t)
                if clip and t + 1 <= self.nsteps - 1:
                    # state t+1's x adjoint (a) against state t's (b), same order
                    ga, gb = a[:d, :n].double(), b[:d, :n].double()
                    keep = (ga - gb).abs() <= self.grad_clip
                    xsum += torch.where(keep, ga * ga, torch.zeros_like(ga)).sum(dim=1)
                a, b = b, a
            b[:, idx_a] = a[:, :n]
            a[:, :n] = b[:, :n]
        if clip:
            gsum[:d] = xsum
        return a


class Rollout(torch.autograd.Function):
    """state_T = Rollout.apply(state_0, solver, nsteps, checkpoint_every, noise,
    adjoint_solver): nsteps plain substeps (with a Noise after each when
    noise is given) as a differentiable torch op.  backward runs the
    hand-written adjoint (of adjoint_solver when given: e.g. the fp32 adjoint
    kernels around an f64 forward); the error-contribution sums of the
    backward pass are left in solver.adjoint_gsum and, with noise, dL/dsigma
    in solver.adjoint_dsigma (float64 [ncomp] on the device)."""

    @staticmethod
    def forward(ctx, state0, solver, nsteps, checkpoint_every=16, noise=None, adjoint_solver=None, grad_clip=None):
        sched = _Schedule(solver, nsteps, checkpoint_every, noise, adjoint_solver, grad_clip)
        out = sched.forward(state0.detach())
        ctx.sched = sched
        return out

    @staticmethod
    def backward(ctx, grad_T):
        sched = ctx.sched
        solver = sched.solver
        adj = SubstepAdjoint(sched.adj_solver)
        gsum = torch.zeros(solver.ncomp, dtype=torch.float64, device=solver.device)
        dsig = None if sched.noise is None else torch.zeros_like(gsum)
        adt = sched.adj_solver.state.dtype
        grad0 = sched.backward(grad_T.contiguous().to(adt), gsum, adj, dsig)
        solver.adjoint_gsum = gsum
        solver.adjoint_dsigma = dsig
        return grad0.to(grad_T.dtype), None, None, None, None, None, None


def error_contributions(solver, nsteps, loss=energy_loss, checkpoint_every=16, grad_clip=None):
    """The reference's AD run (run_forward_with_ad_nlogn + get_gradients):
    nsteps substeps from the solver's current state (substep t at
    step_count + t), loss(solver, s_T), and per component g_c =
    sum_{t=1}^{T-1} sum_p (dL/ds_{t,p,c})^2 (f64); grad_clip: the
    reference's use_grad_clip threshold for the x sums (0.1 in its 3D fluid
    AD run).  Returns (loss value, dL/ds_0 as [ncomp, n] numpy, g as numpy
    [ncomp]).  The solver's state is left unchanged."""
    n = solver.current_idx
    s0 = solver.state.clone().requires_grad_(True)
    sT = Rollout.apply(s0, solver, nsteps, checkpoint_every, None, None, grad_clip)
    L = loss(solver, sT)
    L.backward()
    torch.cuda.current_stream().synchronize()
    g = solver.adjoint_gsum.cpu().numpy()
    return float(L.detach().cpu()), s0.grad[:, :n].double().cpu().numpy(), g


def final_position_error(solver, state, xref):
    """sum_p |x_T,p - xref_p|^2 (f64) of a [ncomp, >= n] state against the
    [dim, n] positions of an unquantized FP64 run."""
    n, d = solver.current_idx, solver.dim
    e = state[:d, :n].double() - xref
    return (e * e).sum()


def noise_scale_gradients(solver, nsteps, sigma, xref, seed=0, checkpoint_every=16, adjoint_solver=None):
    """The adjoint error-gradient run: nsteps substeps from the solver's
    current state with modelled rounding noise sigma[c] * xi after every
    substep, L = |x_T - xref|^2, and dL/dsigma[c] = sum_t sum_p
    dL/ds_t[c,p] xi_t[c,p] from the hand-written adjoint (products and sums
    in f64).  adjoint_solver (optional, same scene): runs the reverse pass
    in its precision (the fp32 adjoint around an f64 forward, whose
    x_T - xref resolves noise far below fp32's rounding of x).  Returns (L,
    dL/dsigma as numpy [ncomp]).  The solver's state is left unchanged."""
    xr = torch.as_tensor(np.asarray(xref), dtype=torch.float64, device=solver.device)
    s0 = solver.state.clone().requires_grad_(True)
    sT = Rollout.apply(s0, solver, nsteps, checkpoint_every, Noise(sigma, seed), adjoint_solver)
    L = final_position_error(solver, sT, xr)
    L.backward()
    torch.cuda.current_stream().synchronize()
    return float(L.detach().cpu()), solver.adjoint_dsigma.cpu().numpy()


def field_gradients(dsigma, sigma, fields):
    """Per-field derivatives dL/ds_f at s_f = 1 when every component c of
    field f has noise width s_f * sigma[c]: sum_{c in f} sigma[c] dL/dsigma[c].
    fields: sequence of (name, slice)."""
    ds = np.asarray(dsigma, dtype=np.float64) * np.asarray(sigma, dtype=np.float64)
    return {name: float(ds[sl].sum()) for name, sl in fields}


def noise_widths(bits, ranges):
    """sigma[c] = the quantization step of a signed fixed-point component:
    range / 2^(bits - 1) (sign bit included in bits)."""
    return np.asarray(ranges, dtype=np.float64) * np.exp2(1.0 - np.asarray(bits, dtype=np.float64))
