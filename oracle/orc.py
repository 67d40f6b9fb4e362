"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the autoquantizer_amd package.
The C sources restate the reference (see oracle/orc.h for citations).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

FIXED, INT, FLOAT, F32 = 0, 1, 2, 3
FLUID, ELASTIC = 0, 1
SEPARATING, STICKY = 0, 1


class Comp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nbits", ctypes.c_int32),
                ("is_signed", ctypes.c_int32), ("bit_offset", ctypes.c_int32),
                ("exp_bits", ctypes.c_int32), ("exp_offset", ctypes.c_int32),
                ("shared", ctypes.c_int32), ("dither", ctypes.c_int32),
                ("scale", ctypes.c_double), ("offset", ctypes.c_double)]


class MPM(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("material", ctypes.c_int32),
                ("n_grid", ctypes.c_int32), ("boundary", ctypes.c_int32),
                ("bound", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("dx", ctypes.c_double), ("inv_dx", ctypes.c_double),
                ("dt", ctypes.c_double), ("p_vol", ctypes.c_double),
                ("p_mass", ctypes.c_double), ("gravity", ctypes.c_double),
                ("la", ctypes.c_double), ("mu_0", ctypes.c_double),
                ("la_0", ctypes.c_double),
                ("axis_bounds", ctypes.c_int32), ("lb", ctypes.c_int32 * 3), ("ub", ctypes.c_int32 * 3),
                ("bound_clip", ctypes.c_int32)]


class Collider(ctypes.Structure):
    _fields_ = [("sdf", ctypes.c_void_p), ("normal", ctypes.c_void_p),
                ("size", ctypes.c_int32 * 3), ("offset", ctypes.c_int32 * 3), ("center", ctypes.c_int32 * 3),
                ("from_rot", ctypes.c_int32), ("yshift", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("cos_t", ctypes.c_double), ("sin_t", ctypes.c_double), ("omega", ctypes.c_double)]


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32, i32, f32, f64 = ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_double
        P = ctypes.c_void_p
        L.orc_fixed_encode_f32.restype = u32
        L.orc_fixed_encode_f32.argtypes = [f32, f64, f64, i32, i32, i32, f32]
        L.orc_fixed_encode_f64.restype = u32
        L.orc_fixed_encode_f64.argtypes = [f64, f64, f64, i32, i32, i32, f64]
        L.orc_fixed_decode_f32.restype = f32
        L.orc_fixed_decode_f32.argtypes = [u32, i32, i32, f64, f64]
        L.orc_fixed_decode_f64.restype = f64
        L.orc_fixed_decode_f64.argtypes = [u32, i32, i32, f64, f64]
        L.orc_float_encode.restype = None
        L.orc_float_encode.argtypes = [f32, i32, i32, i32, P, P]
        L.orc_float_decode.restype = f32
        L.orc_float_decode.argtypes = [u32, u32, i32, i32, i32]
        L.orc_shared_encode.restype = None
        L.orc_shared_encode.argtypes = [P, i32, P, P, i32, P, P]
        L.orc_shared_decode.restype = f32
        L.orc_shared_decode.argtypes = [u32, u32, i32, i32, i32]
        L.orc_get_bits.restype = u32
        L.orc_get_bits.argtypes = [P, i32, i32]
        L.orc_put_bits.restype = None
        L.orc_put_bits.argtypes = [P, i32, i32, u32]
        L.orc_sext.restype = ctypes.c_int32
        L.orc_sext.argtypes = [u32, i32]
        L.orc_hash32.restype = u32
        L.orc_hash32.argtypes = [u32]
        L.orc_noise_key.restype = u32
        L.orc_noise_key.argtypes = [ctypes.c_uint64, u32, u32]
        L.orc_noise_u.restype = f32
        L.orc_noise_u.argtypes = [ctypes.c_uint64, u32, u32, u32]
        L.orc_dither_u.restype = f32
        L.orc_dither_u.argtypes = [ctypes.c_uint64, u32, u32, u32]
        L.orc_noise_fill.restype = None
        L.orc_noise_fill.argtypes = [ctypes.c_uint64, u32, ctypes.c_int, ctypes.c_int64, P]
        for name in ("orc_pack_f32", "orc_pack_f64"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [P, i32, i32, P, P, ctypes.c_uint64, u32, P, ctypes.c_int64]
        for name in ("orc_unpack_f32", "orc_unpack_f64"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [P, i32, i32, P, P, ctypes.c_int64]
        L.orc_store_one_f32.restype = None
        L.orc_store_one_f32.argtypes = [P, i32, P, i32, f32, f32]
        L.orc_num_comps.restype = i32
        L.orc_num_comps.argtypes = [P]
        for sfx in ("f32", "f64"):
            getattr(L, "orc_p2g_" + sfx).argtypes = [P, P, ctypes.c_int64, P, P]
            getattr(L, "orc_grid_op_" + sfx).argtypes = [P, P, P]
            getattr(L, "orc_collide_" + sfx).argtypes = [P, P, P]
            getattr(L, "orc_g2p_" + sfx).argtypes = [P, P, P, P, ctypes.c_int64]
            getattr(L, "orc_run_plain_" + sfx).argtypes = [P, P, ctypes.c_int64, i32, P, P, i32]
            getattr(L, "orc_energy_" + sfx).restype = f64
            getattr(L, "orc_energy_" + sfx).argtypes = [P, P, ctypes.c_int64]
            for nm in ("orc_p2g_", "orc_grid_op_", "orc_g2p_", "orc_run_plain_"):
                getattr(L, nm + sfx).restype = None
        L.orc_run_quant_f32.restype = None
        L.orc_run_quant_f32.argtypes = [P, P, i32, P, ctypes.c_int64, i32, ctypes.c_uint64, u32,
                                        P, P, P, i32]
        L.orc_run_quant_f64.argtypes = [P, P, i32, P, ctypes.c_int64, i32, ctypes.c_uint64, u32,
                                        P, P, P, i32]
        L.orc_limits_f32.restype = None
        L.orc_limits_f32.argtypes = [P, P, ctypes.c_int64, P, P]
        L.orc_adj_work_doubles.restype = ctypes.c_int64
        L.orc_adj_work_doubles.argtypes = [P]
        L.orc_adj_substep_f64.argtypes = [P, P, ctypes.c_int64, P, P, P]
        L.orc_adj_substep_col_f64.argtypes = [P, P, P, ctypes.c_int64, P, P, P]
        L.orc_adjoint_run_f64.restype = ctypes.c_double
        L.orc_adjoint_run_f64.argtypes = [P, P, ctypes.c_int64, i32, i32, P, P, P]
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def comps_array(comps):
    """comps: list of dicts with Comp field names -> ctypes array."""
    arr = (Comp * len(comps))()
    for i, c in enumerate(comps):
        for k, v in c.items():
            setattr(arr[i], k, v)
    return arr


# ---- scalar codec ---------------------------------------------------------

def fixed_scale(nbits, rng, signed):
    """ti.quant.fixed scale (TS/python/taichi/lang/quant_impl.py:42-45)."""
    return rng / 2 ** (nbits - 1) if signed else rng / 2 ** nbits


def fixed_encode(v, nbits, rng=1.0, signed=True, offset=0.0, dither=False, u=0.0, f64=False):
    s = fixed_scale(nbits, rng, signed)
    if f64:
        return lib().orc_fixed_encode_f64(v, s, offset, nbits, int(signed), int(dither), u)
    return lib().orc_fixed_encode_f32(v, s, offset, nbits, int(signed), int(dither), u)


def fixed_decode(code, nbits, rng=1.0, signed=True, offset=0.0, f64=False):
    s = fixed_scale(nbits, rng, signed)
    if f64:
        return lib().orc_fixed_decode_f64(code, nbits, int(signed), s, offset)
    return lib().orc_fixed_decode_f32(code, nbits, int(signed), s, offset)


def float_encode(f, nbits, signed, exp_bits):
    e = ctypes.c_uint32()
    d = ctypes.c_uint32()
    lib().orc_float_encode(f, nbits, int(signed), exp_bits, ctypes.byref(e), ctypes.byref(d))
    return e.value, d.value


def float_decode(e, d, nbits, signed, exp_bits):
    return lib().orc_float_decode(e, d, nbits, int(signed), exp_bits)


def shared_encode(values, nbits, signed, exp_bits):
    n = len(values)
    f = np.asarray(values, dtype=np.float32)
    nb = np.asarray(nbits, dtype=np.int32)
    sg = np.asarray([int(s) for s in signed], dtype=np.int32)
    d = np.zeros(n, dtype=np.uint32)
    e = ctypes.c_uint32()
    lib().orc_shared_encode(_ptr(f), n, _ptr(nb), _ptr(sg), exp_bits, ctypes.byref(e), _ptr(d))
    return e.value, d


def shared_decode(e, d, nbits, signed, exp_bits):
    return lib().orc_shared_decode(e, int(d), nbits, int(signed), exp_bits)


def noise_u(seed, step, slot, comp):
    return lib().orc_noise_u(seed, step, slot, comp)


def dither_u(seed, step, slot, comp):
    """The codec's dithering noise of component comp of record `slot`."""
    return lib().orc_dither_u(seed, step, slot, comp)


def noise_xi(seed, step, nc, n):
    """[nc, n] float64 modelled rounding noise xi = u - 1/2 of (seed, step)."""
    out = np.zeros((nc, n), dtype=np.float64)
    lib().orc_noise_fill(seed, step, nc, n, _ptr(out))
    return out


# ---- record level ----------------------------------------------------------

def pack(comps, nwords, vals, noise=None, seed=0, step=0, f64=False):
    """vals: [nc, n] -> words [nwords, n] uint64 (word planes)."""
    ca = comps_array(comps)
    dt = np.float64 if f64 else np.float32
    vals = np.ascontiguousarray(vals, dtype=dt)
    nc, n = vals.shape
    words = np.zeros((nwords, n), dtype=np.uint64)
    nz = None if noise is None else np.ascontiguousarray(noise, dtype=dt)
    fn = lib().orc_pack_f64 if f64 else lib().orc_pack_f32
    fn(ca, nc, nwords, _ptr(vals), None if nz is None else _ptr(nz), seed, step, _ptr(words), n)
    return words


def unpack(comps, nwords, words, f64=False):
    ca = comps_array(comps)
    words = np.ascontiguousarray(words, dtype=np.uint64)
    n = words.shape[1]
    dt = np.float64 if f64 else np.float32
    vals = np.zeros((len(comps), n), dtype=dt)
    fn = lib().orc_unpack_f64 if f64 else lib().orc_unpack_f32
    fn(ca, len(comps), nwords, _ptr(words), _ptr(vals), n)
    return vals


def store_one(comps, rec, c, v, u=0.0):
    """Single-member store into one record (uint64 array of nwords), in place."""
    ca = comps_array(comps)
    lib().orc_store_one_f32(ca, len(comps), _ptr(rec), c, v, u)


# ---- MPM -------------------------------------------------------------------

def mpm_params(dim, material, n_grid, dt, p_vol=None, p_rho=1.0, gravity=9.8,
               la=400.0, E=1e4, nu=0.2, boundary=SEPARATING, bound=3):
    """Parameters as MPMSolver.__init__ derives them (mpm_exps/engine/mpm_solver.py:19-40)."""
    s = MPM()
    dx = 1.0 / n_grid
    if p_vol is None:
        p_vol = (dx * 0.5) ** dim
    s.dim, s.material, s.n_grid, s.boundary, s.bound = dim, material, n_grid, boundary, bound
    s.dx, s.inv_dx, s.dt, s.p_vol = dx, 1.0 / dx, dt, p_vol
    s.p_mass = p_vol * p_rho
    s.gravity = gravity
    s.la = la
    s.mu_0 = E / (2 * (1 + nu))
    s.la_0 = E * nu / ((1 + nu) * (1 - 2 * nu))
    return s


def num_comps(s):
    return lib().orc_num_comps(ctypes.byref(s))


def num_nodes(s):
    return s.n_grid ** s.dim


def grids(s, f64=False):
    dt = np.float64 if f64 else np.float32
    nodes = num_nodes(s)
    return np.zeros(nodes, dtype=dt), np.zeros((s.dim, nodes), dtype=dt)


def p2g(s, state, gm, gv):
    f64 = state.dtype == np.float64
    fn = lib().orc_p2g_f64 if f64 else lib().orc_p2g_f32
    fn(ctypes.byref(s), _ptr(state), state.shape[1], _ptr(gm), _ptr(gv))


def grid_op(s, gm, gv):
    f64 = gm.dtype == np.float64
    fn = lib().orc_grid_op_f64 if f64 else lib().orc_grid_op_f32
    fn(ctypes.byref(s), _ptr(gm), _ptr(gv))


def set_walls(s, lb, ub, clip=False):
    """MPMConfig.lb / ub per axis (mpm_solver.py:97-105) and use_bound_clip (:696-698)."""
    s.axis_bounds = 1
    for a in range(s.dim):
        s.lb[a], s.ub[a] = int(lb[a]), int(ub[a])
    s.bound_clip = 1 if clip else 0


def collider_state(dt, dx, step, omega=5.0, height_rate=0.15, lag=1):
    """(theta, height shift in nodes, omega) of the collider at substep
    `step`: step_collider_difftaichi (mpm_solver.py:1555-1559, lag 1: substep
    s runs with the state set after substep s - 1) or step_collider
    (:1548-1552, lag 0, rate 0.1); collider_omega is 0 until the first step
    sets it."""
    k = max(step - lag, 0)
    theta = omega * dt * k
    height = height_rate * dt * k
    return theta, int(height / dx + 0.5), (omega if step >= 1 else 0.0)


def collider(sdf, normal, theta, yshift, omega, center, offset=(0, 0, 0), from_rot=False):
    """The collider of one substep as the C struct; the arrays it points to
    (float32 copies of sdf [sx, sy, sz] and normal [sx, sy, sz, 3]) are kept
    on it."""
    sdf = np.ascontiguousarray(sdf, dtype=np.float32)
    normal = np.ascontiguousarray(normal, dtype=np.float32)
    c = Collider()
    c.sdf, c.normal = sdf.ctypes.data, normal.ctypes.data
    c._keep = (sdf, normal)
    for a in range(3):
        c.size[a], c.offset[a], c.center[a] = sdf.shape[a], int(offset[a]), int(center[a])
    c.from_rot, c.yshift = 1 if from_rot else 0, int(yshift)
    c.cos_t, c.sin_t, c.omega = float(np.cos(theta)), float(np.sin(theta)), float(omega)
    return c


def collide(s, gv, sdf, normal, theta, yshift, omega, center, offset=(0, 0, 0), from_rot=False):
    """sdf_collision on gv [3, nodes] in place (after grid_op).  sdf
    [sx, sy, sz] and normal [sx, sy, sz, 3] float32."""
    c = collider(sdf, normal, theta, yshift, omega, center, offset, from_rot)
    fn = lib().orc_collide_f64 if gv.dtype == np.float64 else lib().orc_collide_f32
    fn(ctypes.byref(s), ctypes.byref(c), _ptr(gv))


def g2p(s, state, gv):
    f64 = state.dtype == np.float64
    out = np.zeros_like(state)
    fn = lib().orc_g2p_f64 if f64 else lib().orc_g2p_f32
    fn(ctypes.byref(s), _ptr(state), _ptr(gv), _ptr(out), state.shape[1])
    return out


def run_plain(s, state, nsteps, nthreads=1):
    """In place: state [nc, n] float32/float64."""
    f64 = state.dtype == np.float64
    gm, gv = grids(s, f64)
    fn = lib().orc_run_plain_f64 if f64 else lib().orc_run_plain_f32
    fn(ctypes.byref(s), _ptr(state), state.shape[1], nsteps, _ptr(gm), _ptr(gv), nthreads)
    return gm, gv


def run_quant(s, comps, nwords, words, nsteps, seed=0, step0=0, nthreads=1, f64=False):
    """In place on words [nwords, n]; f64: compute in double."""
    ca = comps_array(comps)
    n = words.shape[1]
    gm, gv = grids(s, f64=f64)
    scratch = np.zeros(2 * len(comps) * n, dtype=np.float64 if f64 else np.float32)
    fn = lib().orc_run_quant_f64 if f64 else lib().orc_run_quant_f32
    fn(ctypes.byref(s), ca, nwords, _ptr(words), n, nsteps, seed, step0, _ptr(gm), _ptr(gv), _ptr(scratch), nthreads)
    return gm, gv


def energy(s, state):
    f64 = state.dtype == np.float64
    fn = lib().orc_energy_f64 if f64 else lib().orc_energy_f32
    return fn(ctypes.byref(s), _ptr(state), state.shape[1])


def limits(s, state):
    nc = state.shape[0]
    mx = np.zeros(nc, dtype=np.float32)
    mn = np.zeros(nc, dtype=np.float32)
    lib().orc_limits_f32(ctypes.byref(s), _ptr(np.ascontiguousarray(state, dtype=np.float32)),
                         state.shape[1], _ptr(mx), _ptr(mn))
    return mx, mn


def adj_substep(s, state, adj_next, col=None):
    """dL/ds_t from dL/ds_{t+1} for one plain f64 substep from `state` [nc, n];
    col: the substep's collider (collider(...)) or None."""
    st = np.ascontiguousarray(state, dtype=np.float64)
    an = np.ascontiguousarray(adj_next, dtype=np.float64)
    out = np.zeros_like(st)
    work = np.zeros(lib().orc_adj_work_doubles(ctypes.byref(s)), dtype=np.float64)
    if col is None:
        lib().orc_adj_substep_f64(ctypes.byref(s), _ptr(st), st.shape[1], _ptr(an), _ptr(out), _ptr(work))
    else:
        lib().orc_adj_substep_col_f64(ctypes.byref(s), ctypes.byref(col), _ptr(st), st.shape[1], _ptr(an),
                                      _ptr(out), _ptr(work))
    return out


def adjoint_run(s, state0, nsteps, loss="energy", direction=None):
    """(loss, dL/ds_0, error-contribution sums per component) of nsteps plain
    f64 substeps; loss "energy" (compute_ek) or "linear" (sum(direction * s_T))."""
    st = np.ascontiguousarray(state0, dtype=np.float64)
    nc = st.shape[0]
    adj0 = np.zeros_like(st)
    gsum = np.zeros(nc, dtype=np.float64)
    kind = 0 if loss == "energy" else 1
    d = np.ascontiguousarray(direction if direction is not None else np.zeros_like(st), dtype=np.float64)
    L = lib().orc_adjoint_run_f64(ctypes.byref(s), _ptr(st), st.shape[1], int(nsteps), kind, _ptr(d),
                                  _ptr(adj0), _ptr(gsum))
    return L, adj0, gsum


def noise_scale_run(s, state0, nsteps, sigma, seed, xref):
    """The adjoint error-gradient functional on plain f64 substeps:
    s_t = substep(s_{t-1}) + sigma * xi_t (t = 1..nsteps, xi_t = noise_xi(seed, t)),
    L = sum_p |x_T - xref|^2.  Returns (L, dL/dsigma [nc]) with
    dL/dsigma_c = sum_t sum_p lambda_t[c, p] xi_t[c, p], lambda_t = dL/ds_t,
    from the adjoint of adj_substep (reverse mode over stored states)."""
    st = np.array(state0, dtype=np.float64)
    nc, n = st.shape
    dim = 3 if nc in (16, 24) else 2
    sig = np.asarray(sigma, dtype=np.float64)[:, None]
    states = [st.copy()]
    for t in range(1, nsteps + 1):
        run_plain(s, st, 1)
        st += sig * noise_xi(seed, t, nc, n)
        states.append(st.copy())
    d = st[:dim] - np.asarray(xref, dtype=np.float64)
    L = float((d * d).sum())
    lam = np.zeros_like(st)
    lam[:dim] = 2.0 * d
    dsig = np.zeros(nc)
    for t in range(nsteps, 0, -1):
        dsig += (lam * noise_xi(seed, t, nc, n)).sum(axis=1)
        lam = adj_substep(s, states[t - 1], lam)
    return L, dsig


def noise_scale_loss(s, state0, nsteps, sigma, seed, xref):
    """L of noise_scale_run without the adjoint (for finite differences)."""
    st = np.array(state0, dtype=np.float64)
    nc, n = st.shape
    dim = 3 if nc in (16, 24) else 2
    sig = np.asarray(sigma, dtype=np.float64)[:, None]
    for t in range(1, nsteps + 1):
        run_plain(s, st, 1)
        st += sig * noise_xi(seed, t, nc, n)
    d = st[:dim] - np.asarray(xref, dtype=np.float64)
    return float((d * d).sum())
