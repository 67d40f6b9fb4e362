"""Benchmark of the quantized MLS-MPM substep on one MI355X.

Workload: 3D weakly compressible fluid (J-based EOS, bulk modulus 400),
1,048,576 particles uniformly at random (seed 0) in [0.1, 0.5]^3, 128^3 grid,
dt = 2e-4, gravity 9.8, separating walls 3 cells thick; state stored as the
32-byte quantized record of autoquantizer_amd.layouts.fluid_3d (x 3 x fixed21,
v 3 x fixed16, C 9 x shared-exponent float(8, 10), J fixed16, dithered),
f32 compute.  A step is one substep of all particles (P2G, grid update, G2P;
G2P fused with the next P2G); the fused kernel is specialised to the record
layout at run time (--generic-kernel for the descriptor-driven one); the
same scene with fp32 storage is timed beside it.  Printed: one JSON line,
with the roofline entry of the fused kernel, the error of both runs against
an FP64 run, the CPU baseline (the C oracle on the host's cores), the scheme
search on the adjoint scene (adjoint path and learned surrogate, each
allocation run as dithered quantized runs; headline `scheme_search_s` is the
surrogate's time plus the adjoint search's times its held-out fallback
rate), the reference's own search experiment on its elastic scene, the
adjoint check against finite differences, and the batched scheme search.  Timing is steady state: the
warm-up substeps run first, and the timed substeps continue from there.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s + scheme-search time, quantized vs fp32 MPM, 1 MI355X"
HBM_PEAK = 8.0e12
# HBM bytes per launch of the kernels on this workload are measured by
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy's
# readings) and summarised in this file, whose `bench-traffic:` line the
# benchmark reads; a profiler run, not this script, measures them.
TRAFFIC_PROFILE = os.path.join("profiles", "k_tiled_fused_fluid3d_1M.md")


def profiled_traffic():
    """{kernel: HBM bytes per launch} from the profile summary's
    `bench-traffic: name=bytes ...` line (empty if absent)."""
    out = {}
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            for line in f:
                if line.startswith("bench-traffic:"):
                    for tok in line.split(":", 1)[1].split():
                        k, v = tok.split("=")
                        out[k] = float(v)
    except OSError:
        pass
    return out


def make_solver(storage, n, n_grid, seed=0):
    import autoquantizer_amd as aq
    from autoquantizer_amd import layouts
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la, cfg.g = 3, True, n_grid, 2e-4, 400.0, 9.8
    cfg.p_vol = (1 / n_grid * 0.5) ** 3
    cfg.n_particles = n
    cfg.seed = 1234
    if storage == "quant":
        cfg.compute_fp = "f32"
        names, t, g, sh = layouts.fluid_3d()
        return aq.MPMSolver(cfg, types=t, groups=g, shared=sh)
    cfg.compute_fp = storage
    return aq.MPMSolver(cfg)


def initial_state(n, seed=0):
    rng = np.random.default_rng(seed)
    vals = np.zeros((16, n))
    vals[:3] = rng.uniform(0.1, 0.5, (3, n))
    vals[15] = 1.0
    return vals


def timed_run(sim, steps, sort_every):
    """Wall time of `steps` substeps (synchronised both sides)."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.advance(steps, sort_every=sort_every)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


EVENT_STEPS = 40  # substeps of the event-timed repeat


def kernel_times(sim, steps, sort_every):
    """Per-kernel event timing over `steps` substeps (aq_advance_timed)."""
    import torch
    from autoquantizer_amd import _lib
    ms = (ctypes.c_float * 4)()
    launches = (ctypes.c_int32 * 4)()
    act = ctypes.c_int64()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().aq_advance_timed(sim._h, ctypes.c_void_p(sim.state.data_ptr()), sim.current_idx,
                                           steps, sim.step_count, sort_every,
                                           ctypes.c_void_p(sim.ids.data_ptr()), ms, launches,
                                           ctypes.byref(act), stream))
    sim.step_count += steps
    return list(ms), list(launches), act.value


def field_errors(a, b):
    """Per-field RMS and max |a - b| over particles; a, b: [16, n]."""
    out = {}
    for name, sl in (("x", slice(0, 3)), ("v", slice(3, 6)), ("C", slice(6, 15)), ("J", slice(15, 16))):
        d = a[sl] - b[sl]
        out[name] = {"rms": float(np.sqrt(np.mean(d * d))), "max": float(np.abs(d).max()),
                     "ref_rms": float(np.sqrt(np.mean(b[sl] ** 2)))}
    return out


def host_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set (the GPU
    box's CPU share), else the cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(n, n_grid, target_s=10.0):
    """The C oracle (the port of the reference substep) on the host's cores, on
    the same quantized workload, for a bounded number of substeps: with all
    the threads of the host's share and with one thread."""
    from oracle import orc
    from autoquantizer_amd import layouts
    from autoquantizer_amd.mpm import plan_state
    names, t, g, sh = layouts.fluid_3d()
    comps, nw = plan_state(names, t, g, sh)
    oc = [dict(kind=c["kind"], nbits=c["num_bits"], is_signed=c["is_signed"], bit_offset=c["bit_offset"],
               exp_bits=c["exp_bits"], exp_offset=c["exp_offset"], shared=c["shared"],
               dither=c["use_dithering"], scale=c["scale"], offset=c["offset"]) for c in comps]
    s = orc.mpm_params(3, orc.FLUID, n_grid, 2e-4, la=400.0)
    words = orc.pack(oc, nw, initial_state(n).astype(np.float32), seed=1234, step=0)

    def rate(threads, budget):
        t0 = time.perf_counter()
        orc.run_quant(s, oc, nw, words, 1, seed=1234, step0=0, nthreads=threads)
        t1 = time.perf_counter() - t0
        k = int(max(1, min(20, budget / max(t1, 1e-3))))
        t0 = time.perf_counter()
        orc.run_quant(s, oc, nw, words, k, seed=1234, step0=1, nthreads=threads)
        tk = time.perf_counter() - t0
        return n * k / tk, k, tk

    threads = host_threads()
    v_all, k_all, t_all = rate(threads, target_s)
    v_one, k_one, t_one = rate(1, target_s)
    return {"value": v_all, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{n} particles, {k_all} quantized substeps after 1 untimed (128^3 grid), C oracle "
                      f"with OpenMP on {threads} threads, {t_all:.1f} s",
            "single_thread": {"value": v_one, "cores": 1,
                              "sample": f"{k_one} substeps after 1 untimed, {t_one:.1f} s"},
            "nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)),
            "build": "gcc -O3 -ffp-contract=off -fopenmp (oracle/Makefile)"}


def _adjoint_scene(n, n_grid, dt, seed=0):
    """The adjoint scene: 2D elastic jelly (fixed-corotated, E=1e4, nu=0.2),
    n particles uniform in [0.3,0.5]^2 with v=(0,-1), n_grid^2 grid, sticky
    walls.  Returns (make_solver(storage), initial state [12, n])."""
    import autoquantizer_amd as aq

    def make(storage="f32"):
        cfg = aq.MPMConfig()
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu, cfg.sticky = 2, False, n_grid, dt, 1e4, 0.2, True
        cfg.p_vol = (1 / n_grid * 0.5) ** 2
        cfg.n_particles = n
        cfg.compute_fp = storage
        return aq.MPMSolver(cfg)

    rng = np.random.default_rng(seed)
    init = np.zeros((12, n))
    init[:2] = rng.uniform(0.3, 0.5, (2, n))
    init[3] = -1.0
    init[8] = init[11] = 1.0
    return make, init


def scheme_search(n=65536, n_grid=256, nsteps=256, checkpoint_every=16, dt=2.5e-5, eps=1e-4, runs=20):
    """The scheme search on the adjoint scene, timed for the adjoint path and
    the learned surrogate, each allocation then run as the reference's
    experiment does (run_mpm_elastic_2d.py:186-208, 227-242).

    Adjoint search: range run + checkpointed fp32 adjoint (f64 sums) of the
    energy loss (compute_ek, sum |v|^2) + FP64 LagrangianEB for the error
    bound eps * loss.  Surrogate search: features + MLP + range run + solve.
    The adjoint judges the surrogate: on the 64 held-out scenes of the
    shipped dataset, the fraction of surrogate allocations that miss the
    bound under the adjoint's sums is the fallback rate, and the expected
    search time is the surrogate's plus that rate times the adjoint's (each
    miss falls back to an adjoint search).  Each allocation is then clamped
    to [1, 32] bits (set_bits_and_ranges), instantiated as the reference's
    quantized solver (fixed-point components in one bit_pack, dithered) and
    run `runs` times; the success fraction is the reference's
    |loss - f64| < 3 eps f64 against an FP64 run of the scene.  On this
    scene (a uniformly moving block: the energy is first-order insensitive
    to x and C) the linear error model gives x and C a few bits, far outside
    its small-noise regime, and the runs miss the bound; reference_search
    runs the same loop on the reference's own scene.  dt = 2.5e-5: with
    E = 1e4 the explicit update diverges at dt = 1e-4 and 5e-5 on this grid
    (wave speed 100 per unit time, CFL number dt * 100 * n_grid)."""
    import torch
    from autoquantizer_amd.search import reference_run, search_scheme, validate_scheme

    make, init = _adjoint_scene(n, n_grid, dt)
    plain, _, L64 = reference_run(make, init, nsteps)
    search_scheme(make, init, 8, tol_rel=eps, checkpoint_every=4)  # warm-up (kernel loading)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = search_scheme(make, init, nsteps, tol_rel=eps, checkpoint_every=checkpoint_every)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    val = validate_scheme(plain, init, nsteps, res["bits"], res["ranges"], L64, eps, runs=runs)
    out = {"adjoint": {"seconds": t_adj, "phases": res["seconds"], "bits": [int(b) for b in res["bits"]],
                       "bits_clamped": [int(b) for b in val["bits"]], "record_bits": val["record_bits"],
                       "grads": [float(g) for g in res["grads"]], "success": val["success"],
                       "rel_error_median": float(np.median(val["rel_error"])),
                       "rel_error_max": float(np.max(val["rel_error"]))},
           "f64_loss": L64, "eps": eps, "runs": runs,
           "scene": f"2D elastic, {n} particles, {n_grid}^2, dt {dt}, {nsteps} substeps, checkpoints every "
                    f"{checkpoint_every}; error bound eps * loss on the energy loss sum |v|^2; success: "
                    f"|loss - f64| < 3 eps f64 over {runs} dithered quantized runs (f32 compute)"}
    data = os.path.join(ROOT, "autoquantizer_amd", "data")
    wpath = os.path.join(data, "surrogate_elastic2d.npz")
    if os.path.exists(wpath):
        from autoquantizer_amd.surrogate import Surrogate, search_scheme_surrogate
        from autoquantizer_amd.train_surrogate import allocations
        model = Surrogate.load(wpath)
        d = np.load(os.path.join(data, "surrogate_elastic2d_dataset.npz"), allow_pickle=False)
        met = [m for _, _, m in allocations(model.predict(d["X"][512:]), d["G"][512:], d["R"][512:], eps,
                                             model.margin)]
        miss = 1.0 - float(np.mean(met))
        search_scheme_surrogate(model, init, 1e4, dt, n_grid, 8, tol_rel=eps)  # warm-up
        torch.cuda.synchronize()
        r = search_scheme_surrogate(model, init, 1e4, dt, n_grid, nsteps, tol_rel=eps, judge=res)
        sval = validate_scheme(plain, init, nsteps, r["bits"], r["ranges"], L64, eps, runs=runs)
        expected = r["seconds"] + miss * t_adj
        out["surrogate"] = {
            "seconds": r["seconds"], "heldout_scenes": len(met), "heldout_fallback_rate": miss,
            "expected_seconds": expected, "speedup_expected": t_adj / expected,
            "speedup": t_adj / r["seconds"], "seconds_with_adjoint_judge": r["seconds"] + t_adj,
            "meets_bound_under_adjoint_sums": r["met"],
            "bits": [int(b) for b in r["bits"]], "bits_clamped": [int(b) for b in sval["bits"]],
            "record_bits": sval["record_bits"], "success": sval["success"],
            "rel_error_median": float(np.median(sval["rel_error"])), "rel_error_max": float(np.max(sval["rel_error"]))}
    return out


def reference_search(eps_list=(0.1, 0.01), runs=20, checkpoint_every=64):
    """run_mpm_elastic_2d.py's experiment on its own scene (eight jelly
    blocks, 80,000 particles, 128^2, dt 2e-4, 8192 substeps; FP64 loss
    6893.409): ranges (-r) and error-contribution sums (-a) from FP64 runs,
    LagrangianEB at each eps (the script's default 0.1, and 0.01) with
    bits + 1, then `runs` dithered quantized runs judged by
    |loss - f64| < 3 eps f64: computing in f64 as the reference's 2D runs
    do (init_quantaichi's compute_fp), and in f32 beside it."""
    import torch
    from autoquantizer_amd.scenes import ELASTIC_JELLY_SUBSTEPS, elastic_jelly_2d
    from autoquantizer_amd.search import search_scheme, validate_scheme
    from autoquantizer_amd.solvers import LagrangianEB

    make, init = elastic_jelly_2d()
    nsteps = ELASTIC_JELLY_SUBSTEPS
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = search_scheme(make, init, nsteps, tol_rel=eps_list[0], checkpoint_every=checkpoint_every)
    torch.cuda.synchronize()
    t_search = time.perf_counter() - t0
    L64 = res["loss"]
    plain = make("f64")
    out = {"scene": f"run_mpm_elastic_2d.py: 8 jelly blocks, {init.shape[1]} particles, 128^2, dt 2e-4, "
                    f"{nsteps} substeps, FP64 sums, checkpoints every {checkpoint_every}",
           "f64_loss": L64, "search_seconds": t_search, "phases": res["seconds"], "runs": runs, "eps": {}}
    for eps in eps_list:
        bits, _ = LagrangianEB(np.asarray(res["grads"]) * np.asarray(res["ranges"]) ** 2, ref=L64, tol_rel=eps)
        bits = bits + 1
        val = validate_scheme(plain, init, nsteps, bits, res["ranges"], L64, eps, runs=runs, compute_fp="f64")
        v32 = validate_scheme(plain, init, nsteps, bits, res["ranges"], L64, eps, runs=runs, compute_fp="f32")
        out["eps"][str(eps)] = {"bits_clamped": [int(b) for b in val["bits"]], "record_bits": val["record_bits"],
                                "compression_vs_fp32": val["record_bits"] / (32.0 * len(bits)),
                                "compute": "f64",
                                "success": val["success"],
                                "rel_error_median": float(np.median(val["rel_error"])),
                                "rel_error_max": float(np.max(val["rel_error"])),
                                "runs_seconds": val["seconds"],
                                "f32_compute": {"success": v32["success"],
                                                "rel_error_median": float(np.median(v32["rel_error"])),
                                                "rel_error_max": float(np.max(v32["rel_error"])),
                                                "runs_seconds": v32["seconds"]}}
    return out


def adjoint_check(n=65536, n_grid=256, nsteps=256, checkpoint_every=16, dt=2.5e-5):
    """The adjoint error-gradient run on the adjoint scene: noise per field
    uniform of width 2^-bits times the field's range (x, v 16 bits, C 4 bits,
    F 13 bits), L = |x_T - x_T^FP64|^2, and dL/d(noise scale of the field)
    from the fp32 adjoint kernels (f64 sums) run around the f64 forward,
    against central differences (h = 1e-3) of that same f64 forward (sorted,
    LDS-tiled segments, noise keyed by particle id).  Each field is checked
    in a run with that field's noise alone: with all fields noisy, x_T is
    moved far more by the position noise than by the C noise, and a
    difference quotient of the C scale would measure rounding, not the
    derivative.  C gets 4 bits because its noise at more bits moves x_T by
    less than 1e-12, where f64 rounding of the forward limits the
    differences.  The f64 adjoint is reported beside it, and the fp32
    adjoint of the fp32 forward, whose own rounding moves x_T more than the
    v and C noise do (so its loss is a different function there)."""
    import torch
    from autoquantizer_amd.adjoint import _Schedule, Noise, _advance_plain, field_gradients, final_position_error, \
        noise_scale_gradients
    from autoquantizer_amd.search import detect_ranges, elastic_2d_ranges

    make, init = _adjoint_scene(n, n_grid, dt)
    s64 = make("f64")
    s64.set_state(init)
    ranges = elastic_2d_ranges(detect_ranges(s64, nsteps), 2)
    s64.set_state(init)
    base = s64.state.clone()
    x = base.clone()
    _advance_plain(s64, x, nsteps)
    xref = x[:2, :n].clone()
    xr = xref.cpu().numpy()
    bits = np.array([16] * 4 + [4] * 4 + [13] * 4, dtype=np.float64)
    sigma = ranges * np.exp2(-bits)
    fields = [("x", slice(0, 2)), ("v", slice(2, 4)), ("C", slice(4, 8)), ("F", slice(8, 12))]
    s32 = make("f32")
    s32.set_state(init)
    noise_scale_gradients(s64, 8, sigma, xr, seed=1, checkpoint_every=4, adjoint_solver=s32)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L, dsig = noise_scale_gradients(s64, nsteps, sigma, xr, seed=1, checkpoint_every=checkpoint_every,
                                    adjoint_solver=s32)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    all_fields = field_gradients(dsig, sigma, fields)

    def loss(sig):
        sched = _Schedule(s64, nsteps, checkpoint_every, Noise(sig, seed=1))
        return float(final_position_error(s64, sched.forward(base), xref))

    h = 1e-3
    per_field = {}
    for name, sl in fields:
        sf = np.zeros(12)
        sf[sl] = sigma[sl]
        L32, d32 = noise_scale_gradients(s64, nsteps, sf, xr, seed=1, checkpoint_every=checkpoint_every,
                                         adjoint_solver=s32)
        L64, d64 = noise_scale_gradients(s64, nsteps, sf, xr, seed=1, checkpoint_every=checkpoint_every)
        Lp, dp = noise_scale_gradients(s32, nsteps, sf, xr, seed=1, checkpoint_every=checkpoint_every)
        ad, ad64, adp = float(sf @ d32), float(sf @ d64), float(sf @ dp)
        fd = (loss(sf * (1 + h)) - loss(sf * (1 - h))) / (2 * h)
        per_field[name] = {"loss_f64": L64, "adjoint_fp32": ad, "adjoint_f64": ad64, "finite_difference": fd,
                           "relative_difference_fp32": abs(ad - fd) / max(abs(fd), 1e-300),
                           "relative_difference_f64": abs(ad64 - fd) / max(abs(fd), 1e-300),
                           "fp32_forward": {"loss": Lp, "adjoint": adp,
                                            "relative_difference": abs(adp - fd) / max(abs(fd), 1e-300)}}
    return {"seconds": t_adj, "loss": L, "bits": [int(b) for b in bits],
            "field_gradients_all_noise": all_fields, "per_field": per_field,
            "max_relative_difference_fp32": max(v["relative_difference_fp32"] for v in per_field.values()),
            "max_relative_difference_f64": max(v["relative_difference_f64"] for v in per_field.values()),
            "scene": f"2D elastic, {n} particles, {n_grid}^2, dt {dt}, {nsteps} substeps, checkpoints every "
                     f"{checkpoint_every}; fp32 adjoint kernels (f64 sums) around the f64 forward, central "
                     f"differences of the f64 forward"}


def batched_search(nsims=32, nsteps=300, seed=0):
    """The batched scheme search: nsims quantized simulations (2D/3D,
    fluid/elastic, 16k-500k particles, grids 64^2-128^3, each with its own
    candidate scheme) advanced nsteps substeps: one-by-one library calls
    and one captured HIP graph over 4 streams (sort intervals per
    simulation from the autotuner), and the grouped launches (aq_group)
    eagerly and from a captured graph (sort interval and fused grid from
    the group autotuner); final-position error of each against its FP64
    run.  particle_substeps_per_s: the fastest of the three batched runs."""
    import torch
    from autoquantizer_amd.batch import Batch, fp64_positions, position_errors, random_specs

    specs = random_specs(nsims, seed=seed)
    b = Batch(specs)
    t0 = time.perf_counter()
    b.autotune()
    t_tune = time.perf_counter() - t0
    b.reset()
    b.run_sequential(8)  # warm-up
    b.reset()
    t_seq = b.run_sequential(nsteps)
    t0 = time.perf_counter()
    b.capture(nsteps, nstreams=4)
    t_cap = time.perf_counter() - t0
    t_graph = b.run_graph()
    got = b.positions()
    # grouped launches: each substep's fused G2P->P2G one launch per class of
    # simulations, the gather one launch per dimension; sort interval and
    # fused grid from the group autotuner
    t0 = time.perf_counter()
    tune = b.autotune_group()
    t_gtune = time.perf_counter() - t0
    b.run_group(8)  # warm-up
    t_group = b.run_group(nsteps)
    t0 = time.perf_counter()
    b.capture_group(nsteps)
    t_gcap = time.perf_counter() - t0
    t_ggraph = b.run_group_graph()
    got_g = b.positions()
    diff = [float(np.sqrt(np.mean((a - r) ** 2))) for a, r in zip(got_g, got)]
    del b
    torch.cuda.empty_cache()
    ref = fp64_positions(specs, nsteps)
    errs = position_errors(got, ref)
    errs_g = position_errors(got_g, ref)
    total = sum(sp.n for sp in specs)
    rms = [e["rms"] for e in errs]
    rms_g = [e["rms"] for e in errs_g]
    best = min(t_graph, t_group, t_ggraph)
    return {"sims": nsims, "substeps": nsteps, "particles": total,
            "particle_substeps_per_s": total * nsteps / best,
            "sequential_s": t_seq, "graph_s": t_graph, "graph_speedup": t_seq / t_graph,
            "graph_particle_substeps_per_s": total * nsteps / t_graph, "capture_s": t_cap,
            "group": {"eager_s": t_group, "graph_s": t_ggraph, "capture_s": t_gcap, "autotune_s": t_gtune,
                      "tuned": tune, "speedup_vs_graph": t_graph / min(t_group, t_ggraph),
                      "particle_substeps_per_s": total * nsteps / min(t_group, t_ggraph),
                      "error_vs_fp64_rms": {"median": float(np.median(rms_g)), "max": float(np.max(rms_g))},
                      "rms_vs_graph_path": {"median": float(np.median(diff)), "max": float(np.max(diff))}},
            "autotune_s": t_tune, "sort_every": [sp.sort_every for sp in specs],
            "error_vs_fp64_rms": {"median": float(np.median(rms)), "max": float(np.max(rms))},
            "per_sim": [[sp.dim, "fluid" if sp.fluid else "elastic", sp.n, sp.n_grid, sp.bits["x"], sp.bits["v"],
                         sp.bits["C"], sp.bits["J"] if sp.fluid else sp.bits["F"], round(e["rms"], 9),
                         round(e["max"], 9)] for sp, e in zip(specs, errs)],
            "per_sim_fields": ["dim", "material", "particles", "grid", "x bits", "v bits", "C digits",
                               "J bits | F digits", "x error rms", "x error max"]}


def _progress(what):
    """A line on stderr as each section starts (the JSON result goes to stdout)."""
    print(f"[bench] {what}", file=sys.stderr, flush=True)


def elastic_3d_benchmark(n=30000000, steps=20, warmup=5, sort_every=16):
    """The reference's 3D elastic performance configuration
    (run_mpm_elastic_3d.py:352-384: 3e7 particles of voxelized letters, the
    24-component benchmark record, with and without bit_pack and dithering)
    on scenes.elastic_3d_letters (256 cells per unit, 2 x 2 x 2 box): the
    quantized variants and fp32 storage, particle-substeps/s over `steps`
    substeps after `warmup` (steady state, HIP-event free wall time)."""
    import torch
    from autoquantizer_amd.scenes import elastic_3d_letters
    t0 = time.perf_counter()
    make, st = elastic_3d_letters(n_target=n)
    setup = time.perf_counter() - t0
    out = {"particles": int(st.shape[1]), "grid": "512^3 (256 per unit, 2 x 2 x 2)", "steps": steps,
           "warmup": warmup, "sort_every": sort_every, "scene_setup_s": setup, "variants": {}}
    for name, storage, bitpack, dither in (("quant_bitpack_dither", "quant", True, True),
                                           ("quant_bitpack", "quant", True, False),
                                           ("quant_bitstruct_dither", "quant", False, True),
                                           ("quant_bitstruct", "quant", False, False),
                                           ("f32", "f32", None, None)):
        sim = make(storage, bitpack=bool(bitpack), dither=bool(dither))
        sim.set_state(st)
        if storage == "quant":
            sim.specialize()
        sim.advance(warmup, sort_every=sort_every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sim.advance(steps, sort_every=sort_every)
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        ek = sim.compute_ek()
        out["variants"][name] = {"particle_substeps_per_s": st.shape[1] * steps / el, "ms_per_substep": 1e3 * el / steps,
                                 "record_bytes": (sim.nwords * 8) if storage == "quant" else 24 * 4,
                                 "energy": float(ek), "finite": bool(np.isfinite(ek))}
        del sim
        torch.cuda.empty_cache()
    q, f = out["variants"]["quant_bitpack_dither"], out["variants"]["f32"]
    out["speedup_vs_fp32"] = q["particle_substeps_per_s"] / f["particle_substeps_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--sort-every", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-search", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--elastic3d", action="store_true",
                    help="also time the reference's 3D elastic performance configuration (3e7 particles, ~2 min)")
    ap.add_argument("--generic-kernel", action="store_true",
                    help="run the descriptor-driven fused kernel instead of the one specialised to the layout")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the substep has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch
    n, K, W = args.particles, args.steps, args.warmup
    init = initial_state(n)

    results = {}
    jit_s = None
    for storage in ("quant", "f32"):
        sim = make_solver(storage, n, args.grid)
        if storage == "quant" and not args.generic_kernel:
            t0 = time.perf_counter()
            sim.specialize()  # the fused kernel compiled for this record layout (cached)
            jit_s = time.perf_counter() - t0
        # every kernel of the path launched once, untimed: first launches in a process pay the
        # runtime's code-object loading, which is not the path's cost
        sim.set_state(init)
        sim.advance(2, sort_every=args.sort_every)
        # cold: K substeps from the unsorted initial state (the window opens with the cold sort of
        # uniformly random particles)
        sim.set_state(init)
        sim.step_count = 0
        el_cold = timed_run(sim, K, args.sort_every)
        # steady state (the headline): W warm-up substeps from the initial state, then K timed
        # substeps continuing from there; every advance call sorts first, so the window opens with a
        # sort of the nearly sorted state and sorts again every sort_every substeps
        sim.set_state(init)
        sim.step_count = 0
        if W:
            sim.advance(W, sort_every=args.sort_every)
        el = timed_run(sim, K, args.sort_every)
        final = sim.get_state()
        # event-timed repeat continuing from there, for the per-kernel roofline; at most
        # EVENT_STEPS substeps: over hundreds of event pairs the recorded durations grow well past
        # what the kernels take (500 substeps: fused kernel 104 us by events in an 89 us substep)
        ms, launches, act = kernel_times(sim, min(K, EVENT_STEPS), args.sort_every)
        results[storage] = dict(elapsed=el, elapsed_cold=el_cold, final=final, ms=ms, launches=launches,
                                act=act, nwords=sim.nwords, g2p_words=None)
        if storage == "quant":
            mask = 0
            for c in list(range(3)) + [15]:
                cc = sim.comps[c]
                for b in range(cc["bit_offset"], cc["bit_offset"] + cc["num_bits"]):
                    mask |= 1 << (b >> 6)
            results[storage]["g2p_words"] = bin(mask).count("1")
        del sim
        torch.cuda.empty_cache()

    q, f = results["quant"], results["f32"]
    value = n * K / q["elapsed"]
    fp32_value = n * K / f["elapsed"]

    # roofline of the dominant kernel, the fused G2P->P2G launch: algorithmic bytes per substep
    # B = N_p 2 S_p + N_act 32 B (each particle's record read once and written once, each active
    # grid node's (m v, m) written once and velocity read once); HBM traffic per launch is measured
    # separately with rocprofv3 FETCH_SIZE/WRITE_SIZE
    avg_s = q["ms"][0] / 1e3 / max(q["launches"][0], 1)
    per_particle = 2 * q["nwords"] * 8
    per_node = 32
    alg_bytes = n * per_particle + q["act"] * per_node
    achieved = alg_bytes / avg_s
    # what the fused kernel itself must move: the record words G2P reads (x and J) + the record written
    g2p_per_particle = q["g2p_words"] * 8 + q["nwords"] * 8
    f_avg = f["ms"][0] / 1e3 / max(f["launches"][0], 1)
    f_bytes = n * 2 * 16 * 4 + f["act"] * per_node
    traffic = profiled_traffic()
    default_workload = not args.generic_kernel and n == 1 << 20 and args.grid == 128
    sorts_in_window = 1 + (K - 1) // args.sort_every if args.sort_every > 0 and K > 0 else 0

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": q["elapsed"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D fluid MLS-MPM, {n} particles uniform in [0.1,0.5]^3 (seed 0), "
                               f"{args.grid}^3 grid, dt 2e-4, quantized 32 B/particle record, dithered",
                   "particles": n, "grid": args.grid, "sort_every": args.sort_every,
                   "fused_kernel": "generic" if args.generic_kernel else "specialised to the record layout",
                   "specialise_seconds": jit_s,
                   "bytes_per_particle": q["nwords"] * 8, "fp32_bytes_per_particle": 64},
        "timing": "steady state: after the warm-up substeps, K substeps continuing from there; the window "
                  f"holds {sorts_in_window} sort(s) of the nearly sorted state; per-kernel times from HIP "
                  f"events over the next {min(K, EVENT_STEPS)} substeps",
        "cold": {"value": n * K / q["elapsed_cold"], "fp32_value": n * K / f["elapsed_cold"],
                 "note": "K substeps from the unsorted initial state (the window opens with the cold sort)"},
        "fp32": {"value": fp32_value, "ms_per_step": f["elapsed"] * 1e3 / K,
                 "fused_kernel_us": f_avg * 1e6, "fused_kernel_GBps": f_bytes / f_avg / 1e9,
                 "traffic": traffic.get("k_tiled_f32") if default_workload else None},
        "speedup_vs_fp32": value / fp32_value,
        "kernel_ms": {"fused_g2p2g": q["ms"][0], "grid_update": q["ms"][1], "sort": q["ms"][2],
                      "first_p2g_last_g2p": q["ms"][3], "launches": q["launches"]},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": traffic.get("aq_jit_tiled") if default_workload else None,
                     "traffic_source": f"rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE ({TRAFFIC_PROFILE})",
                     "gather_traffic": traffic.get("k_gather") if default_workload else None,
                     "kernel": ("k_tiled<3,fluid,quant,8,float,G2P+P2G>" if args.generic_kernel else
                                "aq_jit_tiled (k_tiled<3,fluid,quant,8,float,G2P+P2G> specialised to the record)")
                     + " (fused substep kernel)", "kernel_us": avg_s * 1e6,
                     "algorithmic_bytes": alg_bytes, "active_nodes": q["act"],
                     "bytes_per_particle": per_particle, "bytes_per_node": per_node,
                     "fused_kernel_bytes_per_particle": g2p_per_particle,
                     "fused_kernel_frac": (n * g2p_per_particle + q["act"] * per_node) / avg_s / HBM_PEAK},
        "scheme_search_s": None,
    }
    if not args.no_search:
        try:
            _progress("scheme search")
            ss = scheme_search()
            # the search as shipped: the surrogate, falling back to the adjoint search at the rate its
            # allocations miss the bound under the adjoint's sums on the held-out scenes
            sur = ss.get("surrogate")
            out["scheme_search_s"] = sur["expected_seconds"] if sur is not None else ss["adjoint"]["seconds"]
            out["scheme_search_adjoint_s"] = ss["adjoint"]["seconds"]
            out["scheme_search"] = ss
        except Exception as e:  # reported, not hidden: the line still carries the throughput
            out["scheme_search"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            _progress("reference search")
            out["reference_search"] = reference_search()
        except Exception as e:
            out["reference_search"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            _progress("adjoint check")
            out["adjoint_check"] = adjoint_check()
        except Exception as e:
            out["adjoint_check"] = {"error": f"{type(e).__name__}: {e}"}
    if args.elastic3d:
        try:
            _progress("3D elastic benchmark")
            out["elastic_3d_benchmark"] = elastic_3d_benchmark()
        except Exception as e:
            out["elastic_3d_benchmark"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_batch:
        try:
            _progress("batched search")
            out["batched_search"] = batched_search()
        except Exception as e:
            out["batched_search"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_fp64:
        from autoquantizer_amd.layouts import fluid_3d_error_bounds
        _progress("fp64 reference run")
        sim = make_solver("f64", n, args.grid)
        sim.set_state(init)
        if W:
            sim.advance(W, sort_every=args.sort_every)
        sim.advance(K, sort_every=args.sort_every)
        ref = sim.get_state()
        del sim
        eq = field_errors(q["final"], ref)
        bounds = fluid_3d_error_bounds(W + K, args.grid)
        out["error_vs_fp64"] = {"quant": eq, "fp32": field_errors(f["final"], ref),
                                "quant_rms_bound": bounds,
                                "quant_within_bound": all(eq[k]["rms"] <= bounds[k] for k in bounds)}
    if not args.no_cpu:
        _progress("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(n, args.grid)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
