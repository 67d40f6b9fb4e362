"""Batched scheme search: many quantized simulations of uneven sizes, each
with its own candidate bit allocation, advanced together on one GPU.

The reference evaluates candidate schemes one simulation at a time, each a
separate process (mpm_exps/run_2d_exp.py:46-70 launches
run_mpm_optimality_check.py once per candidate).  Here a batch of simulations (2D and 3D, fluid and elastic,
16k-500k particles, grids 64^2-128^3, seeded random blocks) is captured once
into a HIP graph: every simulation's full run (sorts, fused G2P/P2G, grid
updates) is a branch of the graph, branches are spread over a few streams
so that small simulations run side by side, and the whole search is one
graph launch.  Each simulation's sort interval comes from a small autotuner
that times the candidates on the simulation itself.  Every quantized run is
judged by the final-position error against an unquantized FP64 run of the
same scene.
"""
import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import MPMConfig, MPMSolver, _lib
from .mpm import component_names
from .quant import quant


@dataclass
class SimSpec:
    """One simulation of the batch and its candidate scheme (fraction bits
    per field, sign included for fixed point; C/F digit bits for the
    shared-exponent floats)."""
    dim: int
    fluid: bool
    n: int
    n_grid: int
    bits: dict
    seed: int
    lo: tuple
    size: float
    velocity: tuple
    sort_every: int = 8
    extra: dict = field(default_factory=dict)


def random_specs(nsims=32, seed=0, n_min=16384, n_max=500000):
    """Seeded batch: dimension, material, particle count (log-uniform in
    [n_min, n_max]), grid (2D 64^2..256^2, 3D 64^3 or 128^3: a random
    starting size, grown until the block of 8 particles per cell fits in
    0.6 of the domain), block position, initial velocity and a candidate
    allocation per simulation."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nsims):
        dim = 3 if rng.uniform() < 0.5 else 2
        fluid = True if dim == 3 else bool(rng.uniform() < 0.5)
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        grids = [64, 128] if dim == 3 else [64, 128, 256]
        g = int(rng.integers(len(grids)))
        while g + 1 < len(grids) and (n / 8.0) ** (1.0 / dim) / grids[g] > 0.6:
            g += 1
        n_grid = grids[g]
        size = min(0.6, (n / 8.0) ** (1.0 / dim) / n_grid)
        lo = tuple(float(v) for v in rng.uniform(0.1, 0.9 - size, dim))
        vel = tuple(float(v) for v in rng.uniform(-1.0, 1.0, dim))
        bits = {"x": int(rng.integers(14, 23)), "v": int(rng.integers(10, 17)),
                "C": int(rng.integers(7, 14)), "J": int(rng.integers(10, 17)), "F": int(rng.integers(9, 16))}
        out.append(SimSpec(dim, fluid, n, n_grid, bits, int(rng.integers(1 << 30)), lo, size, vel))
    return out


def spec_layout(spec):
    """(types, groups, shared) of a spec's candidate scheme: x, v (and J)
    fixed point with dithering, C (and F) one shared-exponent float group per
    field, all in one bit_pack."""
    d = spec.dim
    names = component_names(d, spec.fluid)
    b = spec.bits
    t = {}
    for a in range(d):
        t[f"x{a}"] = quant.fixed(b["x"], signed=True, range=1.0, use_dithering=True)
        t[f"v{a}"] = quant.fixed(b["v"], signed=True, range=16.0, use_dithering=True)
        for c in range(d):
            t[f"C{a}{c}"] = quant.float(exp=8, frac=b["C"])
            if not spec.fluid:
                t[f"F{a}{c}"] = quant.float(exp=6, frac=b["F"])
    shared = ("C",)
    if spec.fluid:
        t["J"] = quant.fixed(b["J"], signed=True, range=0.5, offset=1.0, use_dithering=True)
    else:
        shared = ("C", "F")
    return t, [("pack", 0, names)], shared


def spec_config(spec, dt=None):
    cfg = MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid = spec.dim, spec.fluid, spec.n_grid
    cfg.p_vol = (1 / spec.n_grid * 0.5) ** spec.dim
    cfg.n_particles = spec.n
    cfg.la = 400.0
    cfg.E, cfg.nu = 1e4, 0.2
    cfg.sticky = not spec.fluid
    # CFL 0.3 of the faster of the elastic wave (sqrt E) and the fluid's sound speed (sqrt la)
    c = np.sqrt(cfg.E if not spec.fluid else cfg.la)
    cfg.dt = dt or float(0.3 / spec.n_grid / max(c, 20.0))
    cfg.compute_fp = "f32"
    cfg.seed = spec.seed
    return cfg


def spec_state(spec):
    """Initial state [ncomp, n]: uniform in the block, uniform velocity,
    C = 0, F = I / J = 1."""
    rng = np.random.default_rng(spec.seed)
    d, n = spec.dim, spec.n
    names = component_names(d, spec.fluid)
    st = np.zeros((len(names), n))
    st[:d] = np.asarray(spec.lo)[:, None] + spec.size * rng.uniform(size=(d, n))
    st[d:2 * d] = np.asarray(spec.velocity)[:, None]
    if spec.fluid:
        st[-1] = 1.0
    else:
        st[-d * d:] = np.eye(d).reshape(-1, 1)
    return st


class SimGroup:
    """Quantized simulations advanced together through one aq_group: every
    substep's fused G2P->P2G is one launch per class of simulations
    (dimension, material, record-word bucket) whose workgroups walk the
    items of all of them, and the gather + grid update one launch per
    dimension.  States, particle counts and ids are bound at creation; all
    simulations must be at the same step."""

    def __init__(self, sims):
        self.sims = list(sims)
        n = len(self.sims)
        L = _lib.lib()
        nbytes = L.aq_group_workspace_bytes(n)
        if nbytes < 0:
            _lib.check(nbytes)
        dev = self.sims[0].state.device
        self.workspace = torch.empty(int(nbytes) + 256, dtype=torch.uint8, device=dev)
        base = self.workspace.data_ptr()
        ws = ctypes.c_void_p((base + 255) // 256 * 256)
        V = ctypes.c_void_p * n
        handles = V(*[sim._h.value for sim in self.sims])
        states = V(*[sim.state.data_ptr() for sim in self.sims])
        ids = V(*[sim.ids.data_ptr() for sim in self.sims])
        ns = (ctypes.c_int64 * n)(*[sim.current_idx for sim in self.sims])
        h = ctypes.c_void_p()
        _lib.check(L.aq_group_create(handles, states, ns, ids, n, ws, int(nbytes), ctypes.byref(h)))
        self._h = h

    def advance(self, nsteps, sort_every, workgroups=0):
        """nsteps substeps of every simulation, sorted every `sort_every`;
        `workgroups`: grid of the fused launches (0: from the particle count)."""
        step0 = self.sims[0].step_count
        if any(sim.step_count != step0 for sim in self.sims):
            raise ValueError("the simulations of a group must be at the same step")
        _lib.check(_lib.lib().aq_group_advance(self._h, int(nsteps), int(step0), int(sort_every),
                                               int(workgroups), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        for sim in self.sims:
            sim.step_count += int(nsteps)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().aq_group_destroy(h)
            except Exception:
                pass


class Batch:
    """The quantized simulations of a batch (and, on request, their FP64
    references), set to their initial states."""

    def __init__(self, specs, device="cuda"):
        self.specs = list(specs)
        self.sims, self.init_words = [], []
        for sp in self.specs:
            t, g, sh = spec_layout(sp)
            sim = MPMSolver(spec_config(sp), types=t, groups=g, shared=sh, device=device)
            sim.set_state(spec_state(sp))
            self.sims.append(sim)
            self.init_words.append(sim.state.clone())
        self.graph = None

    def reset(self):
        for sim, w in zip(self.sims, self.init_words):
            sim.state.copy_(w)
            sim.ids[:sim.current_idx] = torch.arange(sim.current_idx, dtype=torch.int32, device=sim.state.device)
            sim.step_count = 0

    def autotune(self, candidates=(8, 16, 32), steps=32):
        """Sort interval per simulation: the candidate with the least time
        for `steps` substeps from the initial state (HIP events)."""
        for sp, sim in zip(self.specs, self.sims):
            best = None
            for k in candidates:
                self.reset()
                sim.advance(steps, sort_every=k)  # warm-up of this schedule
                self.reset()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                sim.advance(steps, sort_every=k)
                e1.record()
                e1.synchronize()
                ms = e0.elapsed_time(e1)
                if best is None or ms < best[1]:
                    best = (k, ms)
            sp.sort_every = best[0]
            sp.extra["tuned_ms_per_substep"] = best[1] / steps
        self.reset()

    def run_sequential(self, nsteps):
        """Every simulation's run, one after the other, on the current stream
        (one library call each); returns wall seconds."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for sp, sim in zip(self.specs, self.sims):
            sim.advance(nsteps, sort_every=sp.sort_every)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def capture(self, nsteps, nstreams=4):
        """Captures all runs into one HIP graph: simulations are dealt to
        `nstreams` streams by decreasing particle count (least-loaded first),
        forked from and joined to the capturing stream."""
        order = sorted(range(len(self.sims)), key=lambda i: -self.specs[i].n)
        load = [0] * nstreams
        lanes = [[] for _ in range(nstreams)]
        for i in order:
            j = int(np.argmin(load))
            lanes[j].append(i)
            load[j] += self.specs[i].n
        self.reset()
        torch.cuda.synchronize()
        streams = [torch.cuda.Stream() for _ in range(nstreams)]
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            main = torch.cuda.current_stream()
            for s, lane in zip(streams, lanes):
                s.wait_stream(main)
                with torch.cuda.stream(s):
                    for i in lane:
                        self.sims[i].advance(nsteps, sort_every=self.specs[i].sort_every)
            for s in streams:
                main.wait_stream(s)
        torch.cuda.synchronize()
        self.graph, self.lanes = g, lanes
        return g

    def group(self):
        """The batch's simulations as one SimGroup (created once)."""
        if getattr(self, "_group", None) is None:
            self._group = SimGroup(self.sims)
        return self._group

    def autotune_group(self, sorts=(8, 16, 32), workgroups=(0, 1024, 2048, 4096), steps=32):
        """Launch options of the grouped path: the sort interval and the
        fused launches' grid with the least time for `steps` substeps of the
        whole batch from the initial states (HIP events, best of two)."""
        grp = self.group()
        best = None
        for k in sorts:
            for wg in workgroups:
                ms = []
                for _ in range(2):
                    self.reset()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    grp.advance(steps, k, wg)
                    e1.record()
                    e1.synchronize()
                    ms.append(e0.elapsed_time(e1))
                if best is None or min(ms) < best[2]:
                    best = (k, wg, min(ms))
        self.group_sort_every, self.group_workgroups = best[0], best[1]
        self.reset()
        return {"sort_every": best[0], "workgroups": best[1], "ms_per_substep": best[2] / steps}

    def run_group(self, nsteps, sort_every=None, workgroups=None):
        """All simulations through the grouped launches; returns wall seconds."""
        grp = self.group()
        k = sort_every or getattr(self, "group_sort_every", 16)
        wg = workgroups if workgroups is not None else getattr(self, "group_workgroups", 0)
        self.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        grp.advance(nsteps, k, wg)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def capture_group(self, nsteps):
        """The grouped run captured into one HIP graph (its launches replayed
        without host work)."""
        grp = self.group()
        k = getattr(self, "group_sort_every", 16)
        wg = getattr(self, "group_workgroups", 0)
        self.reset()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            grp.advance(nsteps, k, wg)
        torch.cuda.synchronize()
        self.group_graph, self.group_graph_steps = g, int(nsteps)
        return g

    def run_group_graph(self):
        """One replay of the grouped graph from the initial states; wall seconds."""
        self.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.group_graph.replay()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        for sim in self.sims:
            sim.step_count += self.group_graph_steps
        return el

    def run_graph(self):
        """One launch of the captured graph from the initial states; returns
        wall seconds."""
        self.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.graph.replay()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def positions(self):
        return [sim.get_state(original_order=True)[:sp.dim] for sp, sim in zip(self.specs, self.sims)]


def fp64_positions(specs, nsteps):
    """Final positions of the unquantized FP64 runs of the specs' scenes."""
    out = []
    for sp in specs:
        cfg = spec_config(sp)
        cfg.compute_fp = "f64"
        sim = MPMSolver(cfg)
        sim.set_state(spec_state(sp))
        sim.advance(nsteps, sort_every=8)
        out.append(sim.get_state(original_order=True)[:sp.dim])
        del sim
    return out


def position_errors(quant_x, ref_x):
    """Per simulation: RMS and max |x - x_FP64| over particles and axes."""
    return [{"rms": float(np.sqrt(np.mean((a - b) ** 2))), "max": float(np.abs(a - b).max())}
            for a, b in zip(quant_x, ref_x)]
