"""MLS-MPM solver over quantized particle records: the drop-in for the
reference's MPMConfig / MPMSolver operator API on the substep path.

Mirrors mpm_exps/mpm_config.py:1-53 and mpm_exps/engine/mpm_solver.py
(MPMSolver.__init__ :19-122, init_quantaichi :364-455, reshape_params
:522-535, initialize :564-576, build_pid :578-588, p2g_bls :590-622,
substep :624-631, grid_op :633-665, g2p_bls :667-758, detect_bound* and
get_limits :944-997, compute_ek :1043-1052, add_cube :1328-1342).
Every phase runs in libaqmpm on the GPU; the state lives in torch tensors.
"""
import ctypes
import math
import warnings

import numpy as np
import torch

from . import _lib
from .fields import comps_array, plan
from .quant import PlainF32, place_record, quant


class MPMConfig:
    """Defaults of mpm_exps/mpm_config.py:1-53 (fields the substep path reads)."""

    def __init__(self):
        self.n_grid = 128
        self.n_particles = 9000
        self.dim = 2
        self.dt = 1e-4
        self.p_rho = 1
        self.p_vol = (1 / self.n_grid * 0.5) ** self.dim
        self.E, self.nu = 0.1e4, 0.2
        self.la = 100000
        self.quantized_properties = None
        self.quant_bits = None
        self.quant_ranges = None
        self.use_quantaichi = False
        self.use_bitpack = False
        self.use_fluid = True
        self.use_dithering = True
        self.bound = None
        self.lb = None
        self.ub = None
        self.g = 9.8
        self.compute_fp = 'f64'
        self.size = 1.0
        self.softer = False
        # additions: wall kind and dithering seed
        self.sticky = False
        self.seed = 0


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def component_names(dim, fluid):
    names = [f"x{a}" for a in range(dim)] + [f"v{a}" for a in range(dim)]
    names += [f"C{a}{b}" for a in range(dim) for b in range(dim)]
    names += ["J"] if fluid else [f"F{a}{b}" for a in range(dim) for b in range(dim)]
    return names


def reference_layout(cfg):
    """The record init_quantaichi builds (mpm_solver.py:364-455): every
    component a signed fixed-point type of quant_bits[c] bits and range
    quant_ranges[c] (J with offset 1.0), in one bit_pack when use_bitpack,
    else the bit_struct words of :411-444.  Returns (containers, places)."""
    dim, fluid = cfg.dim, cfg.use_fluid
    names = component_names(dim, fluid)
    bits = list(np.asarray(cfg.quant_bits).reshape(-1))
    ranges = list(np.asarray(cfg.quant_ranges, dtype=np.float64).reshape(-1))
    if len(bits) != len(names) or len(ranges) != len(names):
        raise ValueError(f"need {len(names)} bits and ranges")
    types = {}
    for k, nm in enumerate(names):
        types[nm] = quant.fixed(frac=int(bits[k]), range=float(ranges[k]), signed=True,
                                offset=1.0 if nm == "J" else 0.0, use_dithering=cfg.use_dithering)
    if cfg.use_bitpack:
        groups = [("pack", 0, names)]
    elif fluid and dim == 3:
        groups = [("struct", 64, ["x0", "x1", "x2"]), ("struct", 64, ["v0", "v1", "v2"]),
                  ("struct", 64, ["C00", "C01", "C02", "C10"]), ("struct", 64, ["C11", "C12", "C20", "C21"]),
                  ("struct", 32, ["C22", "J"])]
    elif not fluid and dim == 2:
        groups = [("struct", 64, ["x0", "x1", "v0"]), ("struct", 64, ["v1", "C00", "C01", "C10", "C11"]),
                  ("struct", 64, ["F00", "F01", "F10"]), ("struct", 32, ["F11"])]
    else:
        groups = [("pack", 0, names)]
    return types, groups


def plan_state(names, types, groups, shared=()):
    """Place the state components: groups = [(kind, num_bits, [names])];
    names listed in `shared` prefixes form one shared-exponent group per field.
    Returns canonical-order components and the record words."""
    containers, places, owner = [], [], []
    gid = 0
    for ci, (kind, nb, members) in enumerate(groups):
        containers.append((_lib.AQ_BIT_PACK if kind == "pack" else _lib.AQ_BIT_STRUCT, nb))
        cur_field, cur_gid = None, -1
        for nm in members:
            rec = place_record(types[nm])
            fieldname = nm[0]
            sg = -1
            if fieldname in shared and rec["kind"] == _lib.AQ_FLOAT:
                if cur_field != fieldname:
                    cur_field, cur_gid = fieldname, gid
                    gid += 1
                sg = cur_gid
            else:
                cur_field = None
            rec.update(container=ci, shared_group=sg)
            places.append(rec)
            owner.append(nm)
    comps, nw = plan(containers, places)
    by = dict(zip(owner, comps))
    missing = [nm for nm in names if nm not in by]
    if missing:
        raise ValueError(f"components not placed: {missing}")
    return [by[nm] for nm in names], nw


class MPMSolver:
    """MLS-MPM on one GPU with quantized (or plain fp32/fp64) particle state.

    Quantized storage: `types` maps component names (x0, v1, C01, J, F10, ...)
    to quant types, `groups` lists containers [(kind, bits, names)], and
    `shared` names the fields whose members share one exponent.  Without
    `types`, cfg.use_quantaichi selects the reference's fixed-point scheme
    (cfg.quant_bits / cfg.quant_ranges) and otherwise plain storage in
    cfg.compute_fp precision."""

    def __init__(self, cfg, types=None, groups=None, shared=(), capacity=None, device="cuda"):
        self.device = device
        self._setup(cfg, types, groups, shared, capacity)
        if self.storage == _lib.AQ_STORE_QUANT:
            self.state = torch.zeros((self.nwords, self.capacity), dtype=torch.int64, device=device)
        else:
            self.state = torch.zeros((self.ncomp, self.capacity),
                                     dtype=torch.float64 if self.storage == _lib.AQ_STORE_F64 else torch.float32,
                                     device=device)
        self.ids = torch.arange(self.capacity, dtype=torch.int32, device=device)
        self.current_idx = 0
        self.step_count = 0
        self._make_sim(cfg)
        self.init_limits()

    @classmethod
    def describe(cls, cfg, types=None, groups=None, shared=(), capacity=None):
        """The aq_sim_desc of a solver for cfg, without device work (what
        aq_jit_build takes)."""
        self = cls.__new__(cls)
        self.device = "cpu"
        self._setup(cfg, types, groups, shared, capacity)
        return self._build_desc(cfg)

    def _setup(self, cfg, types, groups, shared, capacity):
        """Physical parameters and the record layout (host only)."""
        self.param = cfg
        self.dim = cfg.dim
        self.fluid = bool(cfg.use_fluid)
        self.n_particles = int(cfg.n_particles)
        self.capacity = int(capacity or self.n_particles)
        # dx = 1 / n_grid; the grid spans `size` (MPMConfig.size) in every
        # axis, n_grid * size cells
        self.dx = 1 / int(cfg.n_grid)
        self.inv_dx = float(int(cfg.n_grid))
        self.n_grid = int(round(int(cfg.n_grid) * max(float(getattr(cfg, "size", 1.0)), 1.0)))
        self.dt = cfg.dt
        self.p_vol = cfg.p_vol
        self.p_mass = cfg.p_vol * cfg.p_rho
        self.E, self.nu = cfg.E, cfg.nu
        self.mu_0 = self.E / (2 * (1 + self.nu))
        self.la_0 = self.E * self.nu / ((1 + self.nu) * (1 - 2 * self.nu))
        if cfg.softer:
            self.mu_0 *= 0.3
            self.la_0 *= 0.3
        self.names = component_names(self.dim, self.fluid)
        self.ncomp = len(self.names)
        if types is None and cfg.use_quantaichi:
            types, groups = reference_layout(cfg)
        self.compute_f64 = False
        if types is not None:
            if str(cfg.compute_fp) in ("f64", "float64"):
                # init_quantaichi (mpm_solver.py:372-374) computes in cfg.compute_fp, f64 by default (the
                # reference's 2D runs); quantized records compute in f64 here in 2D
                if self.dim == 2:
                    self.compute_f64 = True
                else:
                    warnings.warn("3D quantized storage computes in f32; cfg.compute_fp='f64' applies to 2D "
                                  "quantized and to plain storage", stacklevel=3)
            self.storage = _lib.AQ_STORE_QUANT
            if groups is None:
                groups = [("pack", 0, self.names)]
            self.comps, self.nwords = plan_state(self.names, types, groups, tuple(shared))
        else:
            self.comps, self.nwords = [], 0
            f64 = str(cfg.compute_fp) in ("f64", "float64")
            self.storage = _lib.AQ_STORE_F64 if f64 else _lib.AQ_STORE_F32
        bound = cfg.bound if cfg.bound is not None else [3] * self.dim
        self.bound = int(bound[0])
        # per-axis walls (mpm_solver.py:97-105): lb = bound, ub = cells - bound unless given
        lb = cfg.lb if getattr(cfg, "lb", None) is not None else [int(b) for b in bound]
        ub = cfg.ub if getattr(cfg, "ub", None) is not None else [self.n_grid - int(b) for b in bound]
        self.lb, self.ub = [int(v) for v in lb][:self.dim], [int(v) for v in ub][:self.dim]
        self.bound_clip = bool(getattr(cfg, "use_bound_clip", False))

    # ---- handle -----------------------------------------------------------
    def _build_desc(self, cfg):
        d = _lib.SimDesc()
        d.dim, d.material = self.dim, _lib.AQ_FLUID if self.fluid else _lib.AQ_ELASTIC
        d.n_grid = self.n_grid
        d.boundary = _lib.AQ_STICKY if getattr(cfg, "sticky", False) else _lib.AQ_SEPARATING
        d.bound = self.bound
        d.storage = self.storage
        d.ncomp = self.ncomp
        d.nwords = self.nwords
        d.capacity = self.capacity
        d.dx, d.inv_dx, d.dt, d.p_vol = self.dx, self.inv_dx, self.dt, self.p_vol
        d.p_mass, d.gravity = self.p_mass, cfg.g
        d.la, d.mu_0, d.la_0 = cfg.la, self.mu_0, self.la_0
        d.seed = int(getattr(cfg, "seed", 0))
        if self.lb != [self.bound] * self.dim or self.ub != [self.n_grid - self.bound] * self.dim:
            d.axis_bounds = 1
            for a in range(self.dim):
                d.lb[a], d.ub[a] = self.lb[a], self.ub[a]
        d.bound_clip = 1 if self.bound_clip else 0
        d.compute_f64 = 1 if self.compute_f64 else 0
        for k, c in enumerate(self.comps):
            for key, val in c.items():
                setattr(d.comps[k], key, val)
        return d

    def _make_sim(self, cfg):
        d = self._build_desc(cfg)
        self._desc = d
        L = _lib.lib()
        nbytes = L.aq_sim_workspace_bytes(ctypes.byref(d))
        if nbytes < 0:
            _lib.check(-1)
        self.workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        h = ctypes.c_void_p()
        _lib.check(L.aq_sim_create(ctypes.byref(d), _ptr(self.workspace), nbytes, ctypes.byref(h)))
        self._h = h
        self._carr = comps_array(self.comps) if self.comps else None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().aq_sim_destroy(h)
            except Exception:
                pass

    # ---- state in and out -------------------------------------------------
    def specialize(self, cache_dir=None):
        """Compile the fused substep kernel for this solver's record layout
        (hipRTC, aq_sim_specialize) and use it for every later advance();
        code objects are cached in cache_dir (default: AQ_JIT_CACHE or a
        directory under the system temp dir).  Quantized storage only; with
        f64 compute the generic kernel stays (the specialised one is f32):
        returns whether the kernel was specialised."""
        if self.storage != _lib.AQ_STORE_QUANT:
            raise ValueError("specialisation applies to quantized storage")
        if self.compute_f64:
            return False
        cache = _lib.jit_cache_dir(cache_dir)
        _lib.check(_lib.lib().aq_sim_specialize(self._h, _lib.CSRC_DIR.encode(), _lib.INCLUDE_DIR.encode(),
                                                cache.encode()))
        return True

    def unspecialize(self):
        """Back to the generic (descriptor-driven) fused kernel."""
        _lib.check(_lib.lib().aq_sim_specialize(self._h, None, None, None))

    def set_state(self, vals, n=None, dither_step=0xFFFFFFFF):
        """vals: [ncomp, n] float array (canonical component order).  With
        quantized storage the values are encoded with the dithering noise of
        step `dither_step`."""
        if not (isinstance(vals, np.ndarray) and vals.dtype == np.float32 and self.storage != _lib.AQ_STORE_F64):
            vals = np.asarray(vals, dtype=np.float64)  # float32 input goes to the f32 paths unconverted
        n = vals.shape[1] if n is None else n
        if n > self.capacity:
            raise ValueError("more particles than capacity")
        self.current_idx = n
        if self.storage == _lib.AQ_STORE_QUANT and self.compute_f64:
            # encoded in double, as the kernels of an f64-compute simulation store
            v = torch.as_tensor(np.ascontiguousarray(vals[:, :n], dtype=np.float64), device=self.device)
            _lib.check(_lib.lib().aq_sim_pack(self._h, _ptr(v), n, int(dither_step) & 0xFFFFFFFF, _ptr(self.state),
                                              _stream()))
        elif self.storage == _lib.AQ_STORE_QUANT:
            v = torch.as_tensor(np.ascontiguousarray(vals[:, :n], dtype=np.float32), device=self.device)
            tmp = torch.zeros((self.nwords, n), dtype=torch.int64, device=self.device)
            _lib.check(_lib.lib().aq_pack(self._carr, self.ncomp, self.nwords, _ptr(v), None,
                                          self._desc.seed, int(dither_step) & 0xFFFFFFFF, _ptr(tmp), n,
                                          _stream()))
            self.state[:, :n] = tmp
        else:
            self.state[:, :n] = torch.as_tensor(vals[:, :n], dtype=self.state.dtype, device=self.device)
        self.ids[:n] = torch.arange(n, dtype=torch.int32, device=self.device)
        torch.cuda.current_stream().synchronize()

    def set_words(self, words):
        """Load raw records [nwords, n] (quantized storage)."""
        w = torch.as_tensor(np.asarray(words).view(np.int64), device=self.device)
        n = w.shape[1]
        self.state[:, :n] = w
        self.current_idx = n
        self.ids[:n] = torch.arange(n, dtype=torch.int32, device=self.device)

    def get_words(self, original_order=False):
        n = self.current_idx
        w = self.state[:, :n]
        if original_order:
            w = w[:, self._inverse_perm()]
        return w.contiguous().cpu().numpy().view(np.uint64)

    def _inverse_perm(self):
        n = self.current_idx
        ids = self.ids[:n].long()
        inv = torch.empty_like(ids)
        inv[ids] = torch.arange(n, device=ids.device)
        return inv

    def get_state(self, original_order=True):
        """Decoded state [ncomp, n] as float64 numpy (particle order of set_state)."""
        n = self.current_idx
        if self.storage == _lib.AQ_STORE_QUANT and self.compute_f64:
            vals = torch.empty((self.ncomp, n), dtype=torch.float64, device=self.device)
            _lib.check(_lib.lib().aq_sim_unpack(self._h, _ptr(self.state), n, _ptr(vals), _stream()))
        elif self.storage == _lib.AQ_STORE_QUANT:
            w = self.state[:, :n].contiguous()
            vals = torch.empty((self.ncomp, n), dtype=torch.float32, device=self.device)
            _lib.check(_lib.lib().aq_unpack(self._carr, self.ncomp, self.nwords, _ptr(w), _ptr(vals), n,
                                            _stream()))
        else:
            vals = self.state[:, :n]
        if original_order:
            vals = vals[:, self._inverse_perm()]
        return vals.double().cpu().numpy()

    # ---- scene setup (mpm_solver.py:564-576, 1328-1342, 1561-1589) --------
    def initialize(self):
        self.current_idx = 0
        self._pending = []

    def add_cube(self, low_corner, box_size, num, velocity=None, seed=0, lattice=False):
        """Particles in a box: uniform random (seeded numpy generator) or, with
        lattice=True, the regular grid of fill_in_box_2duniform (:1561-1571)."""
        d = self.dim
        if lattice:
            m = int(round(num ** (1.0 / d)))
            g = np.stack(np.meshgrid(*[np.arange(m)] * d, indexing="ij"), -1).reshape(-1, d)[:, ::-1]
            x = np.asarray(low_corner) + g / m * np.asarray(box_size)
        else:
            rng = np.random.default_rng(seed)
            x = np.asarray(low_corner) + rng.random((num, d)) * np.asarray(box_size)
        vals = np.zeros((self.ncomp, x.shape[0]))
        vals[:d] = x.T
        if velocity is not None:
            vals[d:2 * d] = np.asarray(velocity, dtype=np.float64).reshape(d, 1)
        if self.fluid:
            vals[-1] = 1.0
        else:
            vals[-d * d:] = np.eye(d).reshape(-1, 1)
        self._pending = getattr(self, "_pending", []) + [vals]
        allv = np.concatenate(self._pending, axis=1)
        self.set_state(allv)

    def emit(self, vals, dither_step=None):
        """Append particles ([ncomp, k] values, canonical component order) to
        the running state, as the reference's emitters write new particles
        after current_idx (fill_in_ball_with_vel / emit_water,
        mpm_solver.py:1598-1632; emit_letters_one_frame, :131-134).  Particles
        already there keep their records and order; the new ones get ids
        n .. n + k - 1 and, with quantized storage, are encoded with the
        dithering noise of `dither_step` (default: the current step with the
        top bit set, apart from the substeps' own noise)."""
        vals = np.asarray(vals, dtype=np.float64)
        k = vals.shape[1]
        n0 = self.current_idx
        if n0 + k > self.capacity:
            raise ValueError("more particles than capacity")
        if self.storage == _lib.AQ_STORE_QUANT:
            step = (self.step_count | 0x80000000) if dither_step is None else int(dither_step)
            v = torch.as_tensor(np.ascontiguousarray(vals, dtype=np.float32), device=self.device)
            tmp = torch.zeros((self.nwords, k), dtype=torch.int64, device=self.device)
            _lib.check(_lib.lib().aq_pack(self._carr, self.ncomp, self.nwords, _ptr(v), None,
                                          self._desc.seed, step & 0xFFFFFFFF, _ptr(tmp), k, _stream()))
            self.state[:, n0:n0 + k] = tmp
        else:
            self.state[:, n0:n0 + k] = torch.as_tensor(vals, dtype=self.state.dtype, device=self.device)
        self.ids[n0:n0 + k] = torch.arange(n0, n0 + k, dtype=torch.int32, device=self.device)
        self.current_idx = n0 + k
        torch.cuda.current_stream().synchronize()

    def add_mesh(self, triangles, sample_density=None, velocity=None, translation=None, seed=0):
        """add_mesh (mpm_solver.py:1356-1380): the mesh voxelized on the
        solver's voxelizer (Voxelizer(res = (n_grid,)^3, dx, super_sample 2),
        :115-116), sample_density + 1 particles seeded per inside voxel in
        solver cells (seed_from_voxels, :1414-1441), appended with
        `velocity`, C = 0 and F = I (J = 1).  Returns the particle count."""
        from .voxelizer import Voxelizer, seed_from_voxels
        if self.dim != 3:
            raise ValueError("add_mesh is 3D")
        if sample_density is None:
            sample_density = 2 ** self.dim
        if getattr(self, "_voxelizer", None) is None:
            self._voxelizer = Voxelizer((int(self.param.n_grid),) * 3, self.dx, super_sample=2)
        vox = self._voxelizer.voxelize(triangles)
        x, v = seed_from_voxels(vox, sample_density, self.dx, translation if translation is not None else (0, 0, 0),
                                velocity if velocity is not None else (0, 0, 0), np.random.default_rng(seed))
        d = self.dim
        vals = np.zeros((self.ncomp, x.shape[0]))
        vals[:d] = x.T
        vals[d:2 * d] = v.T
        if self.fluid:
            vals[-1] = 1.0
        else:
            vals[-d * d:] = np.eye(d).reshape(-1, 1)
        self.emit(vals)
        return x.shape[0]

    def emit_water(self, center, radius, velocity, num, seed=0):
        """emit_water (mpm_solver.py:1627-1630): num particles at center + r
        (cos t, sin t, 0) * radius with t and r uniform in [0, 2 pi) and
        [0, 1) (random_point_in_unit_sphere_in_xy, :1606-1616), velocity
        `velocity`, J = 1 (fluid) or F = I; seeded numpy draws."""
        d = self.dim
        rng = np.random.default_rng(seed)
        t = rng.random(num) * np.pi * 2.0
        r = rng.random(num)
        vals = np.zeros((self.ncomp, num))
        off = np.zeros((d, num))
        off[0], off[1] = np.cos(t) * r, np.sin(t) * r
        vals[:d] = np.asarray(center, dtype=np.float64).reshape(d, 1) + off * radius
        vals[d:2 * d] = np.asarray(velocity, dtype=np.float64).reshape(d, 1)
        if self.fluid:
            vals[-1] = 1.0
        else:
            vals[-d * d:] = np.eye(d).reshape(-1, 1)
        self.emit(vals)

    def set_collider(self, sdf, normal, offset=(0, 0, 0), center=None, omega=5.0, height_rate=0.15, lag=1,
                     velocity_from_rotated=False, first_step=None):
        """The SDF collider of the 3D fluid scenes (init_sdf_collider /
        init_sdf_value, mpm_solver.py:284-319; sdf_collision(_no_arg),
        rot_scene, step_collider(_difftaichi), :1444-1559): sdf [sx, sy, sz]
        and its gradient normal [sx, sy, sz, 3] at node offset `offset`,
        rotating about the vertical line through `center` (default
        get_center_idx: (n_grid, n_grid // 2, n_grid) of MPMConfig.n_grid)
        at omega, rising at height_rate; lag 1 / rate 0.15 is the range and
        AD runs' schedule, lag 0 / rate 0.1 the forward runs'.  The
        collider's substep 0 is the current step unless first_step is given.
        None removes it."""
        if sdf is None:
            self._collider = None
            _lib.check(_lib.lib().aq_sim_set_collider(self._h, None))
            return
        sdf = np.ascontiguousarray(sdf, dtype=np.float32)
        normal = np.ascontiguousarray(normal, dtype=np.float32).reshape(sdf.shape + (3,))
        ng = int(self.param.n_grid)
        if center is None:
            center = (ng * 2 // 2, ng // 2, ng * 2 // 2)
        sdf_t = torch.as_tensor(sdf, device=self.device)
        nrm_t = torch.as_tensor(normal, device=self.device)
        c = _lib.Collider()
        c.sdf, c.normal = sdf_t.data_ptr(), nrm_t.data_ptr()
        for a in range(3):
            c.size[a], c.offset[a], c.center[a] = sdf.shape[a], int(offset[a]), int(center[a])
        c.lag, c.velocity_from_rotated = int(lag), 1 if velocity_from_rotated else 0
        c.first_step = int(self.step_count if first_step is None else first_step) & 0xFFFFFFFF
        c.omega, c.height_rate = float(omega), float(height_rate)
        _lib.check(_lib.lib().aq_sim_set_collider(self._h, ctypes.byref(c)))
        self._collider = (sdf_t, nrm_t, c)  # the device arrays stay alive while set

    # ---- the substep (mpm_solver.py:624-631) ------------------------------
    def substep(self, s=0, t=None):
        _lib.check(_lib.lib().aq_substep(self._h, _ptr(self.state), self.current_idx, _ptr(self.ids),
                                         self.step_count, _stream()))
        self.step_count += 1

    def build_pid(self, s=0):
        _lib.check(_lib.lib().aq_sort(self._h, _ptr(self.state), self.current_idx, _ptr(self.ids),
                                      _stream()))

    def p2g_bls(self, s=0, t=None):
        _lib.check(_lib.lib().aq_p2g(self._h, _ptr(self.state), self.current_idx, _stream()))

    def grid_op(self):
        _lib.check(_lib.lib().aq_grid_op(self._h, self.step_count, _stream()))

    def g2p_bls(self, s=0, t=None):
        _lib.check(_lib.lib().aq_g2p(self._h, _ptr(self.state), self.current_idx, _ptr(self.ids),
                                     self.step_count, _stream()))
        self.step_count += 1

    def advance(self, nsteps, sort_every=0):
        """nsteps substeps with G2P fused into the next P2G."""
        _lib.check(_lib.lib().aq_advance(self._h, _ptr(self.state), self.current_idx, int(nsteps),
                                         self.step_count, int(sort_every), _ptr(self.ids), _stream()))
        self.step_count += int(nsteps)

    def _grid_view(self):
        g = ctypes.c_void_p()
        nodes = ctypes.c_int64()
        _lib.check(_lib.lib().aq_sim_grid(self._h, ctypes.byref(g), ctypes.byref(nodes)))
        dt = torch.float64 if self.storage == _lib.AQ_STORE_F64 or self.compute_f64 else torch.float32
        esz = 8 if dt == torch.float64 else 4
        off = g.value - self.workspace.data_ptr()
        nbytes = nodes.value * 4 * esz
        return self.workspace[off:off + nbytes].view(dt).reshape(nodes.value, 4)

    def grid_velocity(self):
        """Last grid [nodes, 4] (v_x, v_y, v_z, m) in the blocked node order."""
        torch.cuda.current_stream().synchronize()
        return self._grid_view().cpu().numpy()

    def load_grid(self, g):
        """Overwrite the grid with g ([nodes, 4], blocked node order): the
        grid the next G2P reads."""
        v = self._grid_view()
        v.copy_(torch.as_tensor(np.asarray(g), dtype=v.dtype).reshape(v.shape))
        torch.cuda.current_stream().synchronize()

    def grid_dense(self):
        """Grid [d+1, n^d] in dense (i, j[, k]) order: velocity components then mass."""
        g = self.grid_velocity()
        n, d = self.n_grid, self.dim
        e = 4 if d == 3 else 8
        nb = n // e
        idx = np.indices((n,) * d).reshape(d, -1)
        if d == 3:
            blk = ((idx[0] // 4) * nb + idx[1] // 4) * nb + idx[2] // 4
            loc = (idx[0] % 4) * 16 + (idx[1] % 4) * 4 + idx[2] % 4
        else:
            blk = (idx[0] // 8) * nb + idx[1] // 8
            loc = (idx[0] % 8) * 8 + idx[1] % 8
        node = blk * 64 + loc
        out = np.zeros((d + 1, n ** d))
        for a in range(d):
            out[a] = g[node, a]
        out[d] = g[node, 3]
        return out

    # ---- on-disk formats (mpm_solver.py:1761-1895) -------------------------
    def dump_states(self, folder, frame):
        """data_{frame:06d}.ply: packed point cloud of the positions (3D)."""
        from .io import dump_states
        return dump_states(self, folder, frame)

    def checkpoint_states(self, folder):
        """v.npy, C.npy, J.npy / F.npy of the decoded state (float32)."""
        from .io import checkpoint_states
        checkpoint_states(self, folder)

    def load_states(self, folder, frame):
        """The state from dump_states' PLY and checkpoint_states' npy files."""
        from .io import load_states
        load_states(self, folder, frame)

    # ---- loss and ranges -------------------------------------------------
    def compute_ek(self):
        out = torch.zeros(1, dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().aq_energy(self._h, _ptr(self.state), self.current_idx, _ptr(out),
                                        _stream()))
        return float(out.item())

    compute_loss = compute_ek

    def init_limits(self):
        self._mxmn = torch.empty(2 * self.ncomp, dtype=torch.float32, device=self.device)
        self._mxmn[:self.ncomp] = -math.inf
        self._mxmn[self.ncomp:] = math.inf

    def detect_bound_kernel(self, s=0):
        _lib.check(_lib.lib().aq_update_limits(self._h, _ptr(self.state), self.current_idx,
                                               _ptr(self._mxmn), _stream()))

    def count_wraps(self, enable=True):
        """Range-violation counter (quantized storage): while enabled, every
        record encode adds, per component, the stores whose fixed-point code
        wrapped (value outside the member's range).  Read with wrap_counts()."""
        if enable:
            self._wraps = torch.zeros(self.ncomp, dtype=torch.int64, device=self.device)
            _lib.check(_lib.lib().aq_sim_count_wraps(self._h, _ptr(self._wraps)))
        else:
            _lib.check(_lib.lib().aq_sim_count_wraps(self._h, None))
            self._wraps = None

    def wrap_counts(self):
        """Per-component wrapped stores since count_wraps() (numpy int64)."""
        return self._wraps.cpu().numpy()

    def advance_detect(self, nsteps, sort_every=0):
        """nsteps substeps (as advance) with detect_bound_kernel after each."""
        _lib.check(_lib.lib().aq_advance_limits(self._h, _ptr(self.state), self.current_idx, int(nsteps),
                                                self.step_count, int(sort_every), _ptr(self.ids), _ptr(self._mxmn),
                                                _stream()))
        self.step_count += int(nsteps)

    def detect_bound(self):
        m = self._mxmn.cpu().numpy().astype(np.float64)
        self.limits = np.maximum(np.abs(m[:self.ncomp]), np.abs(m[self.ncomp:]))
        return self.limits

    def get_limits(self):
        """get_limits (mpm_solver.py:987-997): x placeholder 2, then max|v|, max|C|, max|J| or max|F|."""
        lim = self.detect_bound()
        d = self.dim
        return np.concatenate([[2.0] * d, lim[d:]])
