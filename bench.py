"""Benchmark of the projective-dynamics frame on one MI355X.

Workload: a 256x256 cloth (65,536 vertices, 130,050 triangles, 194,565 bend
hinges), 1 m x 1 m, whole top edge pinned, stretch 1e3, bend 1e-2, density
0.1 kg/m^2, dt 1/60, gravity 9.81 m/s^2, seeded +-1e-3 m jitter, uniform wind
0.5 N/m^2 (synthetic, seeded input). A "step" is one simulation frame of PD
(Jacobi + Chebyshev on FP64 displacement iterates with FP32 storage of the
frame data) iterated until the FP64 relative residual ||b - K x|| / ||b||
falls to the tolerance (default 1e-5). `value` is frames per second of that
loop with the state resident in HBM, each frame started from x_n (timed
with device events).

config.warm_start compares the plain start with the learned warm start
(trained here, on the GPU, by the library's own kernels, on frames the
batched solver generates): `cfg5_64x256` on 64 held-out 256x256 cloths with
random wind of 0-2 N/m^2 (2,000 converged training cloth-frames, the start
chosen on 8 validation cloths, against the best plain schedule), and
`at_64` / `at_32` on the workload's scene at 64x64 and 32x32 (beside the
second-order extrapolation), each frame to the same tolerance with the plain
start as arbiter and fallback.

The roofline object describes the kernel that dominates the run, the fused
iteration kernel at 256x256 in the launch shape the autotuner chose (one or
two iterations per launch): algorithmic bytes per launch, 56 N + 32 T + 40 E
for one iteration and 68 N + 32 T + 40 E for two, over its average launch
time, measured here with HIP events on the solver's stream, against the
MI355X HBM peak of 8 TB/s. `traffic` is the HBM
bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a copy
kernel), read from the committed profile summary for the launch shape the
autotuner chose (null when that shape was not profiled). The same figures for
the 1024x1024 irregular mesh and the 512-cloth batch are reported under
roofline.mesh_1024 and roofline.batched_512.

cpu_baseline times the C++ oracle (FP32 storage, FP64 iterates) on the
host's cores: frames of the workload to the tolerance (`value`), one thread
and all threads at 20 iterations per frame, the reference's CPU backend on
the 32x32 workload, and under `plan` the 32x32 workload for 200 frames in
both modes, the 1024x1024 mesh per frame, and one cloth of the 64-cloth
random-wind workload, each beside the GPU's figure for the same frames and
the GPU's error against the FP64 oracle.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "cloth-cuda_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)
os.environ.setdefault("CLOTH_QUIET", "1")

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_iter_256x256_pd_traffic.md")
TRAFFIC_FILE_1024 = os.path.join(ROOT, "profiles", "fused_iter_1024x1024_pd_traffic.md")
TRAFFIC_FILE_BATCH = os.path.join(ROOT, "profiles", "fused_iter_batch512_pd_traffic.md")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# hipGraph capture of the frames of every solver the benchmark creates
# (--no-graph clears it: chunks and fixed-length frames are then enqueued
# directly, with the same kernels and bits)
USE_GRAPH = 1


def solver_config(ch, **kw):
    """clh_config with the benchmark's graph setting unless kw names one."""
    kw.setdefault("use_graph", USE_GRAPH)
    return ch.config(**kw)


def schedule(args):
    """The frame schedule's damping and restart settings (clh_config)."""
    return dict(relax=args.relax, restart_rule=args.restart_rule,
                restart_factor=args.restart_factor)


def oracle_schedule(args, sol):
    """The same settings for the oracle, with the step scale the solver used."""
    return dict(relax=sol.stats().relax, restart_rule=args.restart_rule,
                restart_factor=args.restart_factor)


def algorithmic_bytes(N, T, E, iters_per_launch=1):
    """Bytes one launch of the iteration kernel must move at the least: per
    vertex x_k, x_{k-1} and y (12 B each), mass and Jacobi diagonal (4 B each)
    read once and 12 B written per iteration; per triangle 32 B and per hinge
    40 B of constants and indices read once. 56 N + 32 T + 40 E for one
    iteration per launch, 68 N + 32 T + 40 E for two."""
    return (44 + 12 * iters_per_launch) * N + 32 * T + 40 * E


def tune_line(cands, digits):
    """tile/threads, x iterations per launch, r element regions: us per iteration"""
    return ", ".join(f"{c['tile_verts']}/{c['block_threads']}x{c['iters_per_launch']}"
                     f"r{c['element_regions']}: {c['us_per_iteration']:.{digits}f}"
                     for c in cands) + " us/iteration"


def profiled_traffic(path, plan):
    """HBM bytes per launch of the plan's launch shape, as recorded in a
    committed rocprofv3 summary: the `bytes / launch` column of the row
    `| tile / threads | dispatches | FETCH KiB | WRITE KiB | bytes / launch |`
    of one-iteration launches (None for two iterations per launch, which the
    summaries do not cover, or a shape they lack)."""
    if plan.iters_per_launch != 1:
        return None
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"^\|\s*%d\s*/\s*%d\s*\|\s*\d+\s*\|\s*[0-9.]+\s*\|\s*[0-9.]+\s*\|"
                  r"\s*([0-9,]+)\s*\|" % (plan.tile_verts, plan.block_threads), text, re.M)
    return float(m.group(1).replace(",", "")) if m else None


def kernel_name(plan, mesh):
    if plan.iters_per_launch == 2:
        return f"k_fused_pair (two PD iterations per launch, {mesh})"
    return f"k_fused_iter (one PD iteration per launch, {mesh})"


def launch_ms(sol, args, launches, windows=3):
    """Average milliseconds per launch of the iteration kernel in the frame
    loop's steady phase, HIP events around `launches` back-to-back launches:
    the best of `windows` such windows (the first after a pause runs while
    the GPU's clocks ramp up)."""
    return min(sol.time_iteration_kernel(args, launches) for _ in range(windows))


def build_scene(spec):
    import clothhip as ch
    s = ch.Scene(**spec)
    s.upload()
    return s


def fp64_error(ch, s, pos0, vel0, common, iterations):
    """Max |x_gpu - x_fp64| / extent after one frame of `iterations`
    iterations from the same state (GPU PD vs the FP64 oracle)."""
    import numpy as np
    import pyoracle as po
    from helpers import CFG2, CFG2_WIND, extent, oracle_mesh, scene_arrays
    s.set_state(pos0, vel0)
    sol = ch.Solver(solver_config(ch, max_iterations=iterations, **common), s.N, s.T, s.E)
    sol.step(s.step_args())
    x_gpu, _ = s.get_state()
    m = oracle_mesh(scene_arrays(s))
    p = po.params(mode=po.PD, iterations=iterations, gravity=-9.81, dt=CFG2["dt"],
                  rho=common["rho"], cheb_delay=common["cheb_delay"], ext_accel=CFG2_WIND,
                  threads=host_cpu()["omp_threads"], relax=sol.stats().relax)
    x64, *_ = po.step(m, p, pos0, vel0, "f64")
    return float(np.abs(x_gpu.astype(np.float64) - x64).max() / extent(pos0))


def host_cpu():
    """nproc, the OpenMP thread count the environment grants, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "omp_threads": threads, "model": model}


def tolerance_frame_error(ch, s, sol, args_c, pos_w, vel_w, args):
    """The error of one benchmark frame (to the tolerance, from the state the
    timed frames start from) against the oracle: the same algorithm and
    settings in FP32 (its stopping iteration) and in FP64 (its positions), and
    the FP64 oracle run on to a relative residual of 1e-8 (the frame's PD
    fixed point, three orders of magnitude closer than the tolerance), all as
    max |x_gpu - x_oracle| / extent."""
    import numpy as np
    import pyoracle as po
    from helpers import CFG2, CFG2_WIND, extent, oracle_mesh, scene_arrays
    s.set_state(pos_w, vel_w)
    sol.step(args_c)
    st = sol.stats()
    x_gpu, _ = s.get_state()
    m = oracle_mesh(scene_arrays(s))
    threads = host_cpu()["omp_threads"]
    kw = dict(mode=po.PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho, ext_accel=CFG2_WIND,
              check_every=args.check_every, threads=threads, **oracle_schedule(args, sol))
    e = extent(pos_w)
    out = {"gpu_iterations": int(st.iterations)}
    for name, prec, tol, delay, cap in (("fp32", "f32", args.tol, args.cheb_delay, args.max_iters),
                                         ("fp64", "f64", args.tol, args.cheb_delay, args.max_iters),
                                         ("fp64_converged", "f64", 1e-8, args.cheb_delay,
                                          40000)):
        p = po.params(iterations=cap, tolerance=tol, cheb_delay=delay, **kw)
        t0 = time.time()
        x, _, _, _, it = po.step(m, p, pos_w, vel_w, prec)
        out[name] = {"error": float(np.abs(x_gpu - x).max() / e), "iterations": int(it),
                     "tolerance": tol, "seconds": time.time() - t0}
    return out


def cpu_plan_32(ch, threads):
    """The 32x32 workload (10 iterations/frame, Chebyshev rho 0.99, 200
    frames from rest) in reference mode and PD mode: the C++ oracle in FP32
    on one thread and on `threads`, the GPU solver on the same frames, and
    the GPU trajectory's error after 200 frames against the oracle's FP32
    and FP64 trajectories (max |x_gpu - x_oracle| / extent). In reference
    mode any two of these trajectories part by about 5e-2 of the extent
    within 20 frames, the FP32 and FP64 oracles too: the reference's
    rotation takes its V = I branch on a discriminant below 1e-10, which
    rounding decides for the nearly undeformed triangles, so reference mode
    is pinned one iteration at a time from identical inputs instead."""
    import numpy as np
    import pyoracle as po
    from helpers import CFG1, extent, oracle_mesh, scene_arrays
    F = 200
    out = {"workload": "32x32 workload, 10 iterations/frame, rho 0.99, 200 frames from rest"}
    for name, mode, omode in (("ref_cuda", ch.MODE_REF_CUDA, po.REF_CUDA),
                              ("pd", ch.MODE_PD, po.PD)):
        s = build_scene(CFG1)
        pos0, vel0 = s.get_state()
        m = oracle_mesh(scene_arrays(s))
        kw = dict(mode=omode, iterations=10, gravity=-9.81, dt=CFG1["dt"], rho=0.99, cheb_delay=1)
        r = {}
        for th in sorted({1, threads}):
            p = po.params(threads=th, **kw)
            x, v = pos0.copy(), vel0.copy()
            t0 = time.time()
            for _ in range(F):
                x, v, *_ = po.step(m, p, x, v, "f32")
            r[f"cpu_frames_per_s_{th}_threads"] = F / (time.time() - t0)
        p = po.params(threads=threads, **kw)
        x64, v64 = pos0.astype(np.float64), vel0.astype(np.float64)
        for _ in range(F):
            x64, v64, *_ = po.step(m, p, x64, v64, "f64")
        sol = ch.Solver(solver_config(ch, mode=mode, max_iterations=10, gravity=-9.81,
                                      dt=CFG1["dt"], rho=0.99, cheb_delay=1), s.N, s.T, s.E)
        a = s.step_args()
        sol.step(a)  # plan and graph, then the timed trajectory from rest
        s.set_state(pos0, vel0)
        ms = sol.run_frames(a, F)
        xg, _ = s.get_state()
        r["gpu_frames_per_s"] = F * 1e3 / ms
        r["fp32_oracle_error_200_frames"] = float(np.abs(xg - x).max() / extent(pos0))
        r["fp64_oracle_error_200_frames"] = float(np.abs(xg.astype(np.float64) - x64).max()
                                                  / extent(pos0))
        out[name] = r
        log(f"cpu plan 32x32 {name}: " + json.dumps(r))
        del sol, s
    return out


def cpu_plan_1024(ch, s, sol, a, threads, frames=2):
    """The 1024x1024 irregular workload at 30 iterations/frame: the C++
    oracle in FP32 on `threads`, seconds per frame over `frames` frames from
    rest, and the GPU's first frame from rest against the oracle's FP64
    frame (max |x_gpu - x_fp64| / extent)."""
    import numpy as np
    import pyoracle as po
    from helpers import CFG3, extent, oracle_mesh, scene_arrays
    pos0 = s.get("pos0").reshape(-1, 3).astype(np.float32)
    vel0 = np.zeros_like(pos0)
    t0 = time.time()
    m = oracle_mesh(scene_arrays(s))
    setup = time.time() - t0
    kw = dict(mode=po.PD, iterations=30, gravity=-9.81, dt=CFG3["dt"], rho=0.99, cheb_delay=1,
              threads=threads, relax=sol.stats().relax)
    p = po.params(**kw)
    x, v = pos0.copy(), vel0.copy()
    t0 = time.time()
    for _ in range(frames):
        x, v, *_ = po.step(m, p, x, v, "f32")
    per_frame = (time.time() - t0) / frames
    x64, *_ = po.step(m, p, pos0, vel0, "f64")
    s.set_state(pos0, vel0)
    sol.step(a)
    xg, _ = s.get_state()
    out = {"workload": "1024x1024 irregular mesh, 30 iterations/frame, rho 0.99, from rest",
           "cpu_seconds_per_frame": per_frame, "cpu_frames_per_s": 1.0 / per_frame,
           "cpu_frames": frames, "cores": threads, "oracle_setup_s": setup,
           "fp64_oracle_error_one_frame": float(np.abs(xg.astype(np.float64) - x64).max()
                                                / extent(pos0))}
    log("cpu plan 1024x1024: " + json.dumps(out))
    return out


def cpu_plan_cfg5(ch, threads, args, frames=2):
    """One cloth of the 64-cloth random-wind workload (held-out cloth 0),
    frames to the tolerance from rest at the benchmark's schedule: the GPU
    solving it alone and the C++ oracle (FP32) on `threads`, and the GPU's
    first frame against the FP64 oracle run for the GPU's iteration count
    with the same checks (two runs stopped by the tolerance stop at
    different checks when their arithmetic differs, and at a relative
    residual of 1e-5 the positions of such a stiff cloth are bounded only to
    about 1e-2 of the extent). The learned start is not deployed at 256x256
    (its validation picks the plain start), so plain against plain is the
    comparison the deployed solver makes."""
    import numpy as np
    import pyoracle as po
    from helpers import CFG2, cfg5_specs, extent, oracle_mesh, scene_arrays
    especs, eacc = cfg5_specs(64, seed=9, cloth_seed=9000)
    s = build_scene(especs[0])
    pos0, vel0 = s.get_state()
    acc = ch.DeviceArray.from_host(np.ascontiguousarray(eacc[:1], np.float32))
    a = s.step_args()
    a.d_cloth_accel = acc.ptr
    cfg = solver_config(ch, mode=ch.MODE_PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho,
                        tolerance=args.tol, check_every=args.check_every, stall_checks=40,
                        max_iterations=8000, cheb_delay=args.cheb_delay, **schedule(args))
    sol = ch.Solver(cfg, s.N, s.T, s.E)
    sol.step(a)  # plan, bound and graphs
    s.set_state(pos0, vel0)
    gi = []
    xg0 = None
    t0 = time.time()
    for f in range(frames):
        sol.step(a)
        gi.append(int(sol.stats().iterations))
        if f == 0:
            t_get = time.time()
            xg0, _ = s.get_state()
            t0 += time.time() - t_get
    gpu_s = time.time() - t0
    m = oracle_mesh(scene_arrays(s))
    kw = dict(mode=po.PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho, cheb_delay=args.cheb_delay,
              ext_accel=tuple(float(z) for z in eacc[0]), iterations=8000, tolerance=args.tol,
              check_every=args.check_every, threads=threads, **oracle_schedule(args, sol))
    p = po.params(**kw)
    x, v = pos0.copy(), vel0.copy()
    ci = []
    t0 = time.time()
    for _ in range(frames):
        x, v, _, _, n = po.step(m, p, x, v, "f32")
        ci.append(int(n))
    cpu_s = time.time() - t0
    kw.update(iterations=gi[0], tolerance=1e-30)
    x64, *_ = po.step(m, po.params(**kw), pos0, vel0, "f64")
    out = {"workload": f"held-out cloth 0 of the 64 x 256x256 random-wind workload, {frames} "
                       f"frames from rest to relative residual {args.tol}, plain start",
           "gpu_frames_per_s": frames / gpu_s, "gpu_iterations": gi,
           "cpu_frames_per_s": frames / cpu_s, "cpu_iterations": ci, "cores": threads,
           "fp64_oracle_error_frame_0": float(np.abs(xg0.astype(np.float64) - x64).max()
                                              / extent(pos0)),
           "fp64_oracle_iterations_frame_0": gi[0]}
    log("cpu plan 64 x 256x256, one cloth: " + json.dumps(out))
    return out


def cpu_baseline(ch, s, sol, pos_w, vel_w, args, plan_1024=None):
    """The oracle (C++, FP32 storage and FP64 iterates, as the GPU) timed on
    the host: frames of the benchmark's workload to the tolerance at the
    benchmark's settings, from the state the GPU's timed frames start from,
    for about args.cpu_seconds on the OpenMP threads the environment grants;
    fixed-length iterations on one thread; and the CPU reference backend
    (CpuStretchReferenceSolver) on the 32x32 workload, factored once and
    refactored every iteration as the reference does."""
    import pyoracle as po
    from helpers import CFG1, CFG2, CFG2_WIND, oracle_mesh, scene_arrays
    host = host_cpu()
    threads = host["omp_threads"]
    m = oracle_mesh(scene_arrays(s))
    kw = dict(mode=po.PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho, cheb_delay=args.cheb_delay,
              ext_accel=CFG2_WIND, **oracle_schedule(args, sol))
    p = po.params(iterations=args.max_iters, tolerance=args.tol, check_every=args.check_every,
                  threads=threads, **kw)
    pos, vel = pos_w.copy(), vel_w.copy()
    t0 = time.time()
    frames, its = 0, 0
    while time.time() - t0 < args.cpu_seconds:
        pos, vel, _, _, n = po.step(m, p, pos, vel, "f32")
        frames += 1
        its += n
    dt_cpu = time.time() - t0
    log(f"cpu baseline: {frames} frames to {args.tol} in {dt_cpu:.1f} s on {threads} threads "
        f"({its / frames:.0f} iterations/frame)")
    # one thread: iterations/s of fixed 20-iteration frames over a third of the time
    p1 = po.params(iterations=20, threads=1, **kw)
    pos, vel = pos_w.copy(), vel_w.copy()
    t1 = time.time()
    frames1 = 0
    while time.time() - t1 < args.cpu_seconds / 3:
        pos, vel, *_ = po.step(m, p1, pos, vel, "f32")
        frames1 += 1
    dt1 = time.time() - t1
    it1 = frames1 * 20 / dt1
    # all threads: frames/s at the workload's own 20 iterations per frame
    p20 = po.params(iterations=20, threads=threads, **kw)
    pos, vel = pos_w.copy(), vel_w.copy()
    t1 = time.time()
    frames20 = 0
    while time.time() - t1 < args.cpu_seconds / 3:
        pos, vel, *_ = po.step(m, p20, pos, vel, "f32")
        frames20 += 1
    fps20 = frames20 / (time.time() - t1)
    # the CPU reference backend on the 32x32 workload, 50 iterations a frame
    c1 = ch.Scene(**CFG1)
    ref_cpu = {"workload": "32x32 workload, CpuStretchReferenceSolver (stretch only, 50 "
                           "iterations of exact solves per frame)"}
    for refactor, frames_c in ((False, 200), (True, 20)):
        c1.set_cpu_state(c1.get("pos0"))
        r = ch.CpuRefSolver(c1, gravity=-9.81, refactor=refactor)
        t2 = time.time()
        r.step(frames_c)
        key = "refactor_every_iteration" if refactor else "factor_once"
        ref_cpu[key] = {"frames_per_s": frames_c / (time.time() - t2), "frames": frames_c,
                        "factorizations": r.factorizations()}
        r.close()
    log(f"cpu reference backend at 32x32: {ref_cpu['factor_once']['frames_per_s']:.1f} frames/s "
        f"factored once, {ref_cpu['refactor_every_iteration']['frames_per_s']:.1f} refactored")
    return {"value": frames / dt_cpu, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{frames} frames of the 256x256 workload to relative residual {args.tol} "
                      f"at the benchmark's settings (rho {args.rho}, delay {args.cheb_delay}, "
                      f"step scale {kw['relax']:.6f}, restart rule {args.restart_rule}, "
                      f"check every {args.check_every}), from the state the GPU's timed frames "
                      f"start from: C++ oracle, FP32 storage and FP64 iterates, OpenMP "
                      f"{threads} threads, {dt_cpu:.1f} s, {its / frames:.0f} iterations/frame",
            "host": host,
            "single_core": {"iterations_per_s": it1, "cores": 1,
                            "sample": f"{frames1} frames x 20 iterations in {dt1:.1f} s"},
            "fixed_20_iterations": {"frames_per_s": fps20, "cores": threads,
                                    "frames": frames20},
            "ref_cpu_32x32": ref_cpu,
            "plan": {"32x32": cpu_plan_32(ch, threads), "1024x1024": plan_1024,
                     "cfg5_one_cloth": cpu_plan_cfg5(ch, threads, args)}}


def persistent_frames(ch, s, args_c, pos_w, vel_w, common, args):
    """The same timed frames (from the same state, same schedule) in
    persistent cooperative launches, one per chunk of check_every iterations
    (tiles wait only for their neighbour tiles between iterations), at
    256-vertex tiles and 1024-thread workgroups, against launches per
    iteration at the same shape; both end on the same bits."""
    import numpy as np
    out = {}
    ends = {}
    for name, extra in (("launch", {}), ("persistent", dict(persistent=ch.PERSIST_ON))):
        s.set_state(pos_w, vel_w)
        sol = ch.Solver(solver_config(ch, max_iterations=args.max_iters, tolerance=args.tol,
                                      check_every=args.check_every, tile_verts=256,
                                      block_threads=1024, **extra, **common), s.N, s.T, s.E)
        # the plan, the step-scale bound and the chunk graphs are built in
        # warm-up frames; the timed frames start again from the same state
        for _ in range(2):
            sol.step(args_c)
        s.set_state(pos_w, vel_w)
        ms_total, its = 0.0, []
        for _ in range(args.steps):
            ms_total += sol.run_frames(args_c, 1)
            its.append(sol.stats().iterations)
        info = sol.plan_info()
        ends[name] = s.get_state()
        out[name] = {"frames_per_s": args.steps / (ms_total / 1e3),
                     "iterations_per_frame": float(np.mean(its)),
                     "us_per_iteration": 1e3 * ms_total / sum(its),
                     "persistent": int(info.persistent)}
    out["same_bits"] = bool(np.array_equal(ends["launch"][0], ends["persistent"][0]) and
                            np.array_equal(ends["launch"][1], ends["persistent"][1]))
    out["shape"] = "256-vertex tiles, 1024-thread workgroups"
    log(f"persistent: {out['persistent']['frames_per_s']:.2f} frames/s "
        f"({out['persistent']['us_per_iteration']:.2f} us/iteration) against "
        f"{out['launch']['frames_per_s']:.2f} launched at the same shape; same bits: "
        f"{out['same_bits']}")
    return out


def warm_start_study(ch, n, rho, args, pos0=None, vel0=None, scene=None, gen_tol=None,
                     gen_mode=None):
    """Trains the warm start on frames of args.ws_cloths batched n x n cloths
    with random wind and evaluates plain / extrapolation / learned frames to
    the tolerance on the n x n workload (the 256x256 cloth's material, pins
    and wind), each from the same state with the same warm-up."""
    import numpy as np
    from helpers import CFG2, CFG2_WIND, cfg5_specs
    from warmstart import FrameLoop, select_gain
    dt = CFG2["dt"]
    t0 = time.time()
    specs, acc = cfg5_specs(args.ws_cloths, n=n)
    b = build_scene(dict(specs=specs))
    # training frames run to a tighter tolerance than the evaluation;
    # cloth-frames that miss it are left out
    gen_tol = gen_tol or args.ws_gen_tol
    gen_mode = gen_mode or args.ws_gen_mode
    gen = ch.Solver(solver_config(ch, max_iterations=args.ws_gen_cap, tolerance=gen_tol,
                                  check_every=50,
                                  mode=ch.MODE_PD, gravity=-9.81, dt=dt, rho=rho, cheb_delay=1),
                    b.N, b.T, b.E, cloth_offsets=b.vert_off)
    wsm = ch.WarmStart.for_scene(b, dt, -9.81, seed=7)
    wsm.dataset_reserve(args.ws_frames * (b.N + 16 * b.C))
    loop = FrameLoop(b, gen, wsm, acc)
    _, gen_s = loop.run(args.ws_frames, mode=gen_mode, dataset=True)
    log(f"warm start {n}x{n}: {b.C} cloths x {args.ws_frames} frames in {gen_s:.1f} s, "
        f"{loop.kept}/{loop.offered} cloth-frames converged and kept")
    kept, offered = loop.kept, loop.offered
    del gen, loop
    t1 = time.time()
    wsm.fit_output_scale()
    losses = wsm.train(args.ws_steps, 16384, 2e-3)
    ch.check(ch.hip().clh_synchronize(), "sync")
    train_s = time.time() - t1
    params, scale = wsm.params()
    n_samples = wsm.dataset_size()
    log(f"warm start {n}x{n}: {args.ws_steps} Adam steps in {train_s:.1f} s, loss "
        f"{float(np.mean(losses[:20])):.4f} -> {float(np.mean(losses[-100:])):.4f}")
    del wsm, b
    # the gain on the network's correction, chosen on 8 held-out cloths (other
    # jitter seeds, wind directions and strengths, up to the workload's 0.5
    # N/m^2) by the iterations per frame to the tolerance; the plain start
    # runs the same frames, and when it needs the fewest iterations the
    # learned mode runs plain frames (gain None)
    t2 = time.time()
    vspecs, vacc = cfg5_specs(8, seed=9, n=n, cloth_seed=9000, wind_max=0.5)
    vb = build_scene(dict(specs=vspecs))
    vsol = ch.Solver(solver_config(ch, max_iterations=3 * args.ws_gen_cap, tolerance=args.tol,
                                   check_every=args.check_every, mode=ch.MODE_PD, gravity=-9.81,
                                   dt=dt, rho=rho, cheb_delay=1), vb.N, vb.T, vb.E,
                     cloth_offsets=vb.vert_off)
    vws = ch.WarmStart.for_scene(vb, dt, -9.81)
    vws.set_params(params, scale)
    gain, gain_table = select_gain(vb, vsol, vws, cloth_accel=vacc)
    if gain is not None:
        scale *= gain
    chosen = "the plain start" if gain is None else f"gain {gain}"
    log(f"warm start {n}x{n}: {chosen} chosen on held-out cloths in {time.time() - t2:.1f} s "
        f"(iterations per frame: " + ", ".join(f"{g}: {it:.0f}" for g, it in gain_table.items())
        + ")")
    del vb, vsol, vws
    if scene is None:
        scene = build_scene(dict(CFG2, nrows=n, ncols=n, size=1 / (n - 1)))
        pos0, vel0 = scene.get_state()
    common = dict(mode=ch.MODE_PD, gravity=-9.81, dt=dt, rho=rho, cheb_delay=1,
                  ext_accel=CFG2_WIND, tolerance=args.tol, check_every=args.check_every)
    res = {}
    cap = None
    for mode in ("plain", "extrap", "learned"):
        scene.set_state(pos0, vel0)
        full = ch.Solver(solver_config(ch, max_iterations=args.max_iters, **common), scene.N,
                         scene.T, scene.E)
        warm = None
        if mode != "plain":
            warm = ch.Solver(solver_config(ch, max_iterations=cap, **common), scene.N, scene.T,
                             scene.E)
        for sol_m in (full, warm):
            if sol_m is not None:
                sol_m.autotune(scene.step_args(), 50)
        w1 = ch.WarmStart.for_scene(scene, dt, -9.81, ext_accel=CFG2_WIND)
        w1.set_params(params, scale)
        use_ws = w1 if (mode == "extrap" or gain is not None) else None
        fl = FrameLoop(scene, warm or full, use_ws, fallback=full if warm else None)
        fl.run(args.warmup, mode=mode)
        its, sec = fl.run(args.steps, mode=mode)
        if mode == "plain":
            cap = int(1.5 * max(its))
        res[mode] = {"frames_per_s": args.steps / sec, "iterations_per_frame": float(np.mean(its)),
                     "fallback_frames": fl.fallbacks, "paused_frames": fl.paused_frames}
        log(f"  {n}x{n} {mode}: {res[mode]['frames_per_s']:.1f} frames/s, "
            f"{res[mode]['iterations_per_frame']:.0f} iterations/frame, "
            f"{fl.fallbacks} fallbacks, {fl.paused_frames} paused")
        del full, warm, w1, fl
    out = {"mesh": f"{n}x{n}", "rho": rho, "cheb_delay": 1, "training_cloths": args.ws_cloths,
           "training_frame_tolerance": gen_tol, "training_frame_start": gen_mode,
           "training_frames_per_cloth": args.ws_frames, "cloth_frames_kept": kept,
           "cloth_frames_offered": offered, "training_samples": int(n_samples),
           "training_steps": args.ws_steps, "loss_first": float(np.mean(losses[:20])),
           "loss_last": float(np.mean(losses[-100:])), "generate_seconds": gen_s,
           "train_seconds": train_s, "total_seconds": time.time() - t0,
           "warm_cap_iterations": cap, "modes": res,
           "gain": gain, "gain_held_out_iterations": {str(g): it for g, it in gain_table.items()},
           "learned_mode_start": "plain (chosen on held-out frames)" if gain is None
                                 else "extrapolation + network correction",
           "speedup_learned_vs_plain": res["learned"]["frames_per_s"] / res["plain"]["frames_per_s"],
           "speedup_learned_vs_extrapolation":
               res["learned"]["frames_per_s"] / res["extrap"]["frames_per_s"]}
    log(f"warm start {n}x{n}: learned / plain = {out['speedup_learned_vs_plain']:.2f}x frames/s")
    return out


def batch_frames(ch, scene, solver, accel, frames, warmup, ws=None, mode="plain"):
    """Frames of a batch run to the tolerance with per-cloth control: batch
    frames/s, mean batch iterations per frame, and over all cloth-frames the
    converged count and the mean iterations of the converged ones."""
    import numpy as np
    from warmstart import FrameLoop
    loop = FrameLoop(scene, solver, ws, accel)
    loop.run(warmup, mode="plain")
    its, conv_its, conv, stalled, diverged = [], [], 0, 0, 0
    ch.check(ch.hip().clh_synchronize(), "sync")
    t0 = time.perf_counter()
    for _ in range(frames):
        its.append(loop.step(mode))
        st, ci = solver.cloth_status()
        conv += int((st == ch.CLOTH_CONVERGED).sum())
        stalled += int((st == ch.CLOTH_STALLED).sum())
        diverged += int((st == ch.CLOTH_NONFINITE).sum())
        conv_its += ci[st == ch.CLOTH_CONVERGED].tolist()
    sec = time.perf_counter() - t0
    return {"frames_per_s": frames / sec, "converged_cloth_frames_per_s": conv / sec,
            "batch_iterations_per_frame": float(np.mean(its)),
            "cloth_frames": frames * scene.C, "converged": conv, "stalled": stalled,
            "diverged": diverged,
            "converged_cloth_iterations_mean": float(np.mean(conv_its)) if conv_its else None,
            "converged_cloth_iterations_median": float(np.median(conv_its)) if conv_its else None}


def warm_start_cfg5(ch, args):
    """The learned-warm-start workload: 64 cloths of 256x256 (the 256x256
    workload's material and pins) with wind U(0, 2) N/m^2 in random
    directions, frames to the tolerance with per-cloth control (a cloth
    stops when it converges, stalls for 40 checks or diverges).
      1. The plain start at Chebyshev delays 1 and --cheb-delay on 64 held-out
         cloths, with the benchmark's step scale and restart rule;
         the faster is the baseline.
      2. Training frames at that delay on 64 other cloths until at least
         args.ws_train_frames converged cloth-frames are kept (stalled and
         diverged cloth-frames get weight 0).
      3. The network trained on the GPU, then the start (gain on the
         network's correction, smoothing passes and gain on the whole
         step) chosen on 8 validation cloths by converged cloth-frames per
         second (frames that end by the stall rule make batch frames/s
         alone a misleading score).
      4. The best warm start of the validation cloths against the plain
         start on the held-out cloths; the start deployed is the plain one
         when the validation cloths chose it."""
    import numpy as np
    from helpers import CFG2, cfg5_specs
    from warmstart import FrameLoop
    dt = CFG2["dt"]
    t_all = time.time()
    common = dict(mode=ch.MODE_PD, gravity=-9.81, dt=dt, rho=args.rho, tolerance=args.tol,
                  check_every=args.check_every, stall_checks=40, max_iterations=8000,
                  **schedule(args))

    def solver_for(b, delay):
        return ch.Solver(solver_config(ch, cheb_delay=delay, **common), b.N, b.T, b.E,
                         cloth_offsets=b.vert_off)

    especs, eacc = cfg5_specs(64, seed=9, cloth_seed=9000)
    E = build_scene(dict(specs=especs))
    epos, evel = E.get_state()
    plain = {}
    for delay in (1, args.cheb_delay):
        E.set_state(epos, evel)
        sol = solver_for(E, delay)
        plain[delay] = batch_frames(ch, E, sol, eacc, args.ws_eval_frames, args.warmup)
        log(f"cfg5 plain, delay {delay}: " + json.dumps(plain[delay]))
        del sol
    best_delay = max(plain, key=lambda d: plain[d]["frames_per_s"])

    # training set
    tspecs, tacc = cfg5_specs(64)
    T = build_scene(dict(specs=tspecs))
    gen = solver_for(T, best_delay)
    wsm = ch.WarmStart.for_scene(T, dt, -9.81, seed=7)
    wsm.dataset_reserve(args.ws_max_frames * (T.N + 16 * T.C))
    loop = FrameLoop(T, gen, wsm, tacc)
    t0 = time.time()
    frames = 0
    while loop.kept < args.ws_train_frames and frames < args.ws_max_frames:
        loop.step("plain", dataset=True)
        frames += 1
    gen_s = time.time() - t0
    kept, offered = loop.kept, loop.offered
    log(f"cfg5 training set: {frames} frames of 64 cloths in {gen_s:.1f} s, {kept}/{offered} "
        f"converged cloth-frames kept")
    del gen, loop
    t1 = time.time()
    wsm.fit_output_scale()
    losses = wsm.train(args.ws_steps, 16384, 2e-3)
    ch.check(ch.hip().clh_synchronize(), "sync")
    train_s = time.time() - t1
    params, scale = wsm.params()
    del wsm, T

    # the start, chosen on 8 validation cloths
    vspecs, vacc = cfg5_specs(8, seed=11, cloth_seed=11000)
    V = build_scene(dict(specs=vspecs))
    vpos, vvel = V.get_state()
    vsol = solver_for(V, best_delay)
    vws = ch.WarmStart.for_scene(V, dt, -9.81)
    cands = [("plain", None), ("extrapolation", (0.0, 0, 1.0)), ("learned", (1.0, 0, 1.0)),
             ("learned, step smoothed 100", (1.0, 100, 1.0)),
             ("learned, step smoothed 100, gain 0.5", (1.0, 100, 0.5)),
             ("learned, step smoothed 10, gain 0.25", (1.0, 10, 0.25)),
             ("extrapolation, step smoothed 100", (0.0, 100, 1.0))]
    table = {}
    for name, c in cands:
        V.set_state(vpos, vvel)
        if c is None:
            r = batch_frames(ch, V, vsol, vacc, 6, 3)
        else:
            vws.set_params(params, scale * c[0])
            vws.set_step_filter(c[1], c[2])
            r = batch_frames(ch, V, vsol, vacc, 6, 3, ws=vws, mode="learned")
        table[name] = r
        log(f"cfg5 validation, {name}: {r['frames_per_s']:.2f} batch frames/s, "
            f"{r['converged_cloth_frames_per_s']:.1f} converged cloth-frames/s, "
            f"{r['batch_iterations_per_frame']:.0f} iterations, {r['converged']}/{r['cloth_frames']} "
            f"converged at {r['converged_cloth_iterations_mean']}")
    score = lambda n: table[n]["converged_cloth_frames_per_s"]
    deployed = max(table, key=score)
    best = max((n for n, c in cands if c is not None), key=score)
    chosen = dict(cands)[best]
    del vsol, vws, V

    # the chosen start (learned start with the chosen filter) on the held-out cloths
    E.set_state(epos, evel)
    esol = solver_for(E, best_delay)
    ews = ch.WarmStart.for_scene(E, dt, -9.81)
    ews.set_params(params, scale * chosen[0])
    ews.set_step_filter(chosen[1], chosen[2])
    learned = batch_frames(ch, E, esol, eacc, args.ws_eval_frames, args.warmup, ws=ews,
                           mode="learned")
    log(f"cfg5 best warm start of the validation cloths ({best}): " + json.dumps(learned))
    base = plain[best_delay]
    out = {"workload": "64 held-out 256x256 cloths, wind U(0, 2) N/m^2 in random directions, "
                       "frames to relative residual 1e-5 with per-cloth control (stall window 40 "
                       "checks, divergence guard)",
           "plain": {str(d): r for d, r in plain.items()}, "best_plain_delay": best_delay,
           "training_cloth_frames_kept": kept, "training_cloth_frames_offered": offered,
           "training_frames": frames, "training_generate_seconds": gen_s,
           "training_steps": args.ws_steps, "train_seconds": train_s,
           "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-100:])),
           "validation": table, "best_warm_start": best, "deployed_start": deployed,
           "best_warm_filter": {"correction_gain": chosen[0], "step_smoothing_passes": chosen[1],
                                "step_gain": chosen[2]},
           "learned": learned,
           "speedup_frames_per_s": learned["frames_per_s"] / base["frames_per_s"],
           "speedup_converged_cloth_frames_per_s":
               learned["converged_cloth_frames_per_s"] / base["converged_cloth_frames_per_s"],
           "iterations_saved_per_converged_cloth_frame":
               (base["converged_cloth_iterations_mean"] or 0.0) -
               (learned["converged_cloth_iterations_mean"] or 0.0),
           "total_seconds": time.time() - t_all}
    log(f"cfg5: learned / best plain = {out['speedup_frames_per_s']:.2f}x batch frames/s")
    return out


def ws_kernels(ch, args):
    """The learned warm start's kernels at its workload's size, for
    kernel-trace profiling: the 64-cloth 256x256 random-wind batch, 4 plain
    frames to the tolerance written into the training set, args.ws_steps
    Adam steps of 16,384 samples, then 2 frames from the learned start.
    Chunks are enqueued directly (no graph capture); prints one JSON line of
    the host-side times."""
    from helpers import CFG2, cfg5_specs
    from warmstart import FrameLoop
    specs, acc = cfg5_specs(64)
    b = build_scene(dict(specs=specs))
    sol = ch.Solver(solver_config(ch, mode=ch.MODE_PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho,
                                  cheb_delay=args.cheb_delay, tolerance=args.tol,
                                  check_every=args.check_every, max_iterations=3000,
                                  stall_checks=40, use_graph=0),
                    b.N, b.T, b.E, cloth_offsets=b.vert_off)
    ws = ch.WarmStart.for_scene(b, CFG2["dt"], -9.81, seed=7)
    ws.dataset_reserve(6 * (b.N + 16 * b.C))
    loop = FrameLoop(b, sol, ws, acc)
    out = {"vertices": b.N, "plain_frame_s": [], "learned_frame_s": []}
    for _ in range(4):
        t = time.time()
        loop.step("plain", dataset=True)
        out["plain_frame_s"].append(time.time() - t)
    ws.fit_output_scale()
    t = time.time()
    ws.train(args.ws_steps, 16384, 2e-3)
    ch.check(ch.hip().clh_synchronize(), "sync")
    out["train_s"] = time.time() - t
    out["train_steps"] = args.ws_steps
    for _ in range(2):
        t = time.time()
        loop.step("learned")
        out["learned_frame_s"].append(time.time() - t)
    print(json.dumps({"ws_kernels": out}), flush=True)


def launch_timing(ch, args):
    """The launch timing of the roofline alone: on the 256x256, 1024x1024
    and 512-cloth meshes, the autotuner's candidates and then launch_ms of
    the chosen shape (steady-phase consecutive iterations), with no frame
    stepped, and the 256 MiB copy probe; prints one JSON line of the
    per-launch times. Used to profile the iteration kernel (kernel trace,
    counters)."""
    from helpers import CFG2, CFG2_WIND, CFG3, cfg4_specs
    out = {}
    cases = [("256x256", dict(CFG2), dict(ext_accel=CFG2_WIND), 200, 500),
             ("1024x1024", dict(CFG3), {}, 20, 50)]
    if not args.no_batch:
        cases.append(("512 cloths", dict(specs=cfg4_specs()), {}, 10, 20))
    for name, spec, extra, n_tune, n_time in cases:
        sc = build_scene(spec)
        a = sc.step_args()
        dt = spec.get("dt", 1 / 60)
        kw = dict(mode=ch.MODE_PD, gravity=-9.81, dt=dt, rho=0.99, cheb_delay=1,
                  max_iterations=20, relax=args.relax, **extra)
        offs = dict(cloth_offsets=sc.vert_off) if "specs" in spec else {}
        sol = ch.Solver(solver_config(ch, **kw), sc.N, sc.T, sc.E, **offs)
        sol.autotune(a, n_tune)
        t = launch_ms(sol, a, n_time)
        pi = sol.plan_info()
        b = algorithmic_bytes(sc.N, sc.T, sc.E, pi.iters_per_launch)
        out[name] = {"us_per_launch": t * 1e3, "tile_verts": pi.tile_verts,
                     "block_threads": pi.block_threads, "element_regions": pi.element_regions,
                     "achieved_GBs": b / (t * 1e-3) / 1e9}
        log(f"{name}: {t * 1e3:.2f} us per launch at {pi.tile_verts} / {pi.block_threads}, "
            f"{pi.element_regions} region(s)")
        del sol, sc
    # the copy probe the counter passes are calibrated on (256 MiB read and
    # written per dispatch)
    copy_ms = ch.measure_copy(256 << 20, 20)
    out["copy_probe_GBs"] = 2 * (256 << 20) / (copy_ms * 1e-3) / 1e9
    print(json.dumps({"launch_timing": out}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--max-iters", type=int, default=80000)
    ap.add_argument("--check-every", type=int, default=50)
    ap.add_argument("--rho", type=float, default=0.999999)
    ap.add_argument("--cheb-delay", type=int, default=10,
                    help="plain Jacobi iterations before the Chebyshev blend starts (and after "
                         "each restart) in the frames to the tolerance (10, Wang 2015; with the "
                         "damped step any delay from 4 to 20 gives the same rate)")
    ap.add_argument("--relax", type=float, default=-1.0,
                    help="scale gamma of the Jacobi step; -1: from the Gershgorin bound of "
                         "D^-1 K (CLH_RELAX_AUTO), 1: plain Jacobi")
    ap.add_argument("--restart-rule", type=int, default=1,
                    help="Chebyshev restart rule: 0 on any growth between checks, 1 on growth "
                         "by --restart-factor over the sequence's minimum, 2 none")
    ap.add_argument("--restart-factor", type=float, default=10.0)
    ap.add_argument("--variant", type=int, default=0, help="0 auto, 1 unfused, 2 fused")
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue every solver's launches directly instead of replaying "
                         "captured hipGraphs (the kernels and their bits are the same)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-1024", action="store_true")
    ap.add_argument("--no-persist", action="store_true",
                    help="skip the persistent-kernel frames")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-ws", action="store_true", help="skip the learned warm start")
    ap.add_argument("--launch-timing", action="store_true",
                    help="only autotune and time the iteration kernel's launches on the three "
                         "meshes (no frames): for kernel-trace profiling")
    ap.add_argument("--ws-kernels", action="store_true",
                    help="only run the learned warm start's kernels at the 64-cloth 256x256 "
                         "workload's size (4 plain frames, --ws-steps training steps, 2 learned "
                         "frames; no graphs): for kernel-trace profiling")
    ap.add_argument("--ws-train-frames", type=int, default=2000,
                    help="converged cloth-frames of the 64-cloth 256x256 training set")
    ap.add_argument("--ws-max-frames", type=int, default=60,
                    help="cap on the 64-cloth training frames generated to reach them")
    ap.add_argument("--ws-eval-frames", type=int, default=6,
                    help="timed frames of each start on the 64 held-out 256x256 cloths")
    ap.add_argument("--ws-cloths", type=int, default=64)
    ap.add_argument("--ws-frames", type=int, default=32, help="training frames per cloth")
    ap.add_argument("--ws-steps", type=int, default=4000)
    ap.add_argument("--ws-gen-tol", type=float, default=1e-6,
                    help="residual tolerance of the training frames (tighter than the "
                         "evaluation's, so the targets carry little of the start's bias)")
    ap.add_argument("--ws-gen-mode", default="extrap", choices=["plain", "extrap"])
    ap.add_argument("--ws-gen-cap", type=int, default=3000,
                    help="iteration cap of a training frame (cloths missing the tolerance "
                         "are left out of the training set)")
    args = ap.parse_args()
    global USE_GRAPH
    USE_GRAPH = 0 if args.no_graph else 1

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this benchmark measures one GPU; run it with --gpus 1"}))
        sys.exit(2)

    import numpy as np

    import clothhip as ch
    from helpers import CFG2, CFG2_WIND, CFG3

    info = ch.device_info(0)
    log(f"device: {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")
    if args.launch_timing:
        launch_timing(ch, args)
        return
    if args.ws_kernels:
        ws_kernels(ch, args)
        return

    # ---- 256x256: frames/s to the residual tolerance ---------------------------
    s = build_scene(CFG2)
    N, T, E = s.N, s.T, s.E
    args_c = s.step_args()
    pos0, vel0 = s.get_state()
    common = dict(mode=ch.MODE_PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho,
                  cheb_delay=args.cheb_delay, ext_accel=CFG2_WIND, kernel_variant=args.variant,
                  use_graph=0 if args.no_graph else 1, **schedule(args))
    sol_tol = ch.Solver(solver_config(ch, max_iterations=args.max_iters, tolerance=args.tol,
                                      check_every=args.check_every, **common), N, T, E)
    rep256 = sol_tol.autotune(args_c, 200)
    tune256 = rep256.candidates()
    log("256x256 autotune: " + tune_line(tune256, 2))
    s.set_state(pos0, vel0)
    log(f"256x256: N={N} T={T} E={E}; warmup {args.warmup} frames")
    for _ in range(args.warmup):
        sol_tol.step(args_c)
    pos_w, vel_w = s.get_state()  # the state the timed frames start from
    iters, resid, conv = [], [], []
    ms_total = 0.0
    for f in range(args.steps):
        ms = sol_tol.run_frames(args_c, 1)
        st = sol_tol.stats()
        iters.append(st.iterations)
        conv.append(st.converged)
        resid.append(st.residual / st.rhs_norm if st.rhs_norm > 0 else 0.0)
        ms_total += ms
        if f % 5 == 0:
            log(f"  frame {f}: {st.iterations} iterations, rel residual {resid[-1]:.3e}, "
                f"{ms:.2f} ms")
    st_tol = sol_tol.stats()
    log(f"schedule: rho {args.rho}, delay {args.cheb_delay}, step scale {st_tol.relax:.6f} "
        f"(Gershgorin bound of D^-1 K {st_tol.jacobi_bound:.6f}), restart rule {args.restart_rule}")
    fps_tol = args.steps / (ms_total / 1e3)
    ms_per_frame = ms_total / args.steps
    mean_iters = float(np.mean(iters))
    log(f"to tol {args.tol}: {fps_tol:.2f} frames/s, {mean_iters:.0f} iterations/frame, "
        f"{sum(conv)}/{len(conv)} converged")
    persist = None
    if not args.no_persist:
        persist = persistent_frames(ch, s, args_c, pos_w, vel_w, common, args)


    # ---- 256x256: fixed 20 iterations per frame --------------------------------
    s.set_state(pos0, vel0)
    sol20 = ch.Solver(solver_config(ch, max_iterations=20, **common), N, T, E)
    sol20.autotune(args_c, 200)  # the cached choice
    s.set_state(pos0, vel0)
    for _ in range(3):
        sol20.step(args_c)
    ms20 = sol20.run_frames(args_c, 50)
    fps20 = 50 / (ms20 / 1e3)
    log(f"fixed 20 iterations: {fps20:.1f} frames/s")

    # ---- roofline of the iteration kernel -----------------------------------------
    t_iter = launch_ms(sol20, args_c, 500)
    plan = sol20.plan_info()
    ipl = plan.iters_per_launch
    b_iter = algorithmic_bytes(N, T, E, ipl)
    bw = b_iter / (t_iter * 1e-3) / 1e9
    log(f"iteration kernel: {t_iter * 1e3:.2f} us per launch of {ipl} iteration(s), "
        f"{bw:.0f} GB/s algorithmic (variant {plan.variant}, tile {plan.tile_verts}, "
        f"{plan.block_threads} threads)")
    copy_ms = ch.measure_copy(256 << 20, 20)
    copy_gbs = 2 * (256 << 20) / (copy_ms * 1e-3) / 1e9
    log(f"copy probe: {copy_gbs:.0f} GB/s")
    # the same copy kernel at small sizes: the cost of any launch (16 KiB, 4
    # workgroups), of one with the iteration kernel's 512 workgroups (2 MiB),
    # and of moving the iteration kernel's profiled traffic (read + write)
    tr256 = profiled_traffic(TRAFFIC_FILE, plan)
    small = [16 << 10, 2 << 20] + ([int(tr256 / 2) & ~15] if tr256 else [])
    copy_small = []
    for nb in small:
        us = 1e3 * min(ch.measure_copy(nb, 500) for _ in range(3))
        copy_small.append({"bytes_read": nb, "bytes_moved": 2 * nb, "us_per_launch": us})
    log("copy probe, small: " + ", ".join(f"{c['bytes_moved'] / 1e6:.2f} MB moved in "
                                           f"{c['us_per_launch']:.2f} us" for c in copy_small))
    err64 = fp64_error(ch, s, pos0, vel0, common, 20)
    log(f"FP64 oracle error after one 20-iteration frame: {err64:.2e} of extent")
    tol_err = None
    if not args.no_cpu:
        tol_err = tolerance_frame_error(ch, s, sol_tol, args_c, pos_w, vel_w, args)
        log("frame to the tolerance vs FP64 oracle: " + json.dumps(tol_err))

    roof_1024 = None
    plan_1024 = None
    if not args.no_1024:
        s3 = build_scene(CFG3)
        a3 = s3.step_args()
        # the damped step: the irregular mesh's obtuse triangles put the
        # Gershgorin bound of D^-1 K near 4.6, and plain Jacobi steps under the
        # Chebyshev weights diverge within one frame
        sol3 = ch.Solver(solver_config(ch, max_iterations=30, mode=ch.MODE_PD, gravity=-9.81,
                                       dt=CFG3["dt"], rho=0.99, cheb_delay=1, relax=args.relax,
                                       kernel_variant=args.variant), s3.N, s3.T, s3.E)
        rep3 = sol3.autotune(a3, 20)
        tune1024 = rep3.candidates()
        log("1024x1024 autotune: " + tune_line(tune1024, 1))
        sol3.step(a3)
        t3 = launch_ms(sol3, a3, 50)
        pi3 = sol3.plan_info()
        b3 = algorithmic_bytes(s3.N, s3.T, s3.E, pi3.iters_per_launch)
        ms3 = sol3.run_frames(a3, 5) / 5
        tr3 = profiled_traffic(TRAFFIC_FILE_1024, pi3)
        bw3 = b3 / (t3 * 1e-3) / 1e9
        roof_1024 = {"workload": "1024x1024 irregular mesh, 30 iterations/frame",
                     "kernel": kernel_name(pi3, "1024x1024"),
                     "N": s3.N, "T": s3.T, "E": s3.E, "us_per_launch": t3 * 1e3,
                     "iters_per_launch": pi3.iters_per_launch,
                     "us_per_iteration": t3 * 1e3 / pi3.iters_per_launch,
                     "bytes_per_launch": b3, "achieved": bw3, "frac": bw3 / HBM_PEAK_GBS,
                     "traffic": tr3, "frames_per_s_30_iterations": 1e3 / ms3,
                     "relax": float(sol3.stats().relax),
                     "jacobi_bound": float(sol3.stats().jacobi_bound),
                     "tile_verts": pi3.tile_verts, "block_threads": pi3.block_threads,
                     "autotune": tune1024}
        log(f"1024x1024: {t3 * 1e3:.1f} us per launch of {pi3.iters_per_launch} iteration(s), "
            f"{bw3:.0f} GB/s ({100 * bw3 / HBM_PEAK_GBS:.1f}% of peak); "
            f"{1e3 / ms3:.1f} frames/s at 30 iterations")
        if not args.no_cpu:
            plan_1024 = cpu_plan_1024(ch, s3, sol3, a3, host_cpu()["omp_threads"])
        del sol3, s3

    # ---- 512 batched cloths of uneven size, one launch per iteration -----------
    batch = None
    if not args.no_batch:
        from helpers import cfg4_specs
        t0 = time.time()
        specs = cfg4_specs()
        sb = build_scene(dict(specs=specs))
        ab = sb.step_args()
        log(f"batch: {sb.C} cloths, N={sb.N} T={sb.T} E={sb.E} built in {time.time() - t0:.1f} s")
        solb = ch.Solver(solver_config(ch, max_iterations=15, mode=ch.MODE_PD, gravity=-9.81,
                                       dt=specs[0]["dt"], rho=0.99, cheb_delay=1, relax=args.relax),
                         sb.N, sb.T, sb.E, cloth_offsets=sb.vert_off)
        t0 = time.time()
        rep = solb.autotune(ab, 10)
        cands = rep.candidates()
        log(f"autotune ({time.time() - t0:.1f} s): " + tune_line(cands, 0))
        for _ in range(2):
            solb.step(ab)
        msb = solb.run_frames(ab, 10) / 10
        tb = launch_ms(solb, ab, 20)
        pib = solb.plan_info()
        bb = algorithmic_bytes(sb.N, sb.T, sb.E, pib.iters_per_launch)
        bwb = bb / (tb * 1e-3) / 1e9
        batch = {"workload": "512 cloths, sides round(16*16^U) in [16, 256], stretch log-U "
                             "[1e2, 1e4], bend log-U [1e-3, 1e-1], 1-4 random boundary pins, "
                             "15 iterations/frame, graph-captured frame",
                 "cloths": sb.C, "N": sb.N, "T": sb.T, "E": sb.E,
                 "frames_per_s": 1e3 / msb, "cloth_frames_per_s": sb.C * 1e3 / msb,
                 "ms_per_frame": msb, "kernel": kernel_name(pib, "512 cloths"),
                 "us_per_launch": tb * 1e3, "iters_per_launch": pib.iters_per_launch,
                 "us_per_iteration": tb * 1e3 / pib.iters_per_launch,
                 "bytes_per_launch": bb, "achieved": bwb, "frac": bwb / HBM_PEAK_GBS,
                 "traffic": profiled_traffic(TRAFFIC_FILE_BATCH, pib),
                 "tile_verts": pib.tile_verts, "block_threads": pib.block_threads,
                 "autotune": {"best": cands[rep.best], "candidates": cands},
                 "padded_verts": int(pib.padded_verts),
                 "relax_min": float(solb.stats().relax),
                 "jacobi_bound_max": float(solb.stats().jacobi_bound)}
        log(f"batch: {1e3 / msb:.1f} frames/s ({sb.C * 1e3 / msb:.0f} cloth-frames/s), launch of "
            f"{pib.iters_per_launch} iteration(s) {tb * 1e3:.0f} us = {bwb:.0f} GB/s "
            f"({100 * bwb / HBM_PEAK_GBS:.1f}% of peak)")
        del solb, sb

    # ---- learned warm start: train on batched cloths, evaluate on the workload --
    ws = None
    if not args.no_ws:
        ws = {"cfg5_64x256": warm_start_cfg5(ch, args),
              "at_64": warm_start_study(ch, 64, 0.9999, args),
              "at_32": warm_start_study(ch, 32, 0.999, args)}
        ws["plain_start_fallback"] = (
            "a warm-started frame runs under a cap of 1.5x the plain frames' mean iterations; a "
            "frame missing the tolerance is redone from (x_n, v_n) with the plain start, and "
            "3 misses in a row pause the warm start for 20 frames")

    # ---- CPU baseline: the oracle's FP32 PD frames on the host cores -------------
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ch, s, sol_tol, pos_w, vel_w, args, plan_1024)

    value = fps_tol
    line = {
        "metric": "frames/s at 256² to fixed residual, plain vs learned warm-start; % HBM peak",
        "value": value,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / value,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 iterate, f32 storage; bf16 MLP with f32 accumulation",
        "data": "synthetic (seeded cloths, random-init network trained on solver-generated frames)",
        "config": {"workload": "256x256 PD cloth, top edge pinned, wind 0.5 N/m^2, "
                               "frames to relative residual 1e-5",
                   "value_mode": "plain start x_n at cheb_delay, relax and restart_rule "
                                 "below; warm_start.cfg5_64x256 "
                                 "compares the learned start with the best plain schedule on the "
                                 "64-cloth random-wind workload",
                   "warm_start": ws,
                   "tolerance": args.tol, "residual": "FP64 ||b - K x|| / ||b|| on the GPU",
                   "graphs": not args.no_graph,
                   "max_iterations": args.max_iters, "check_every": args.check_every,
                   "rho": args.rho, "cheb_delay": args.cheb_delay,
                   "relax": float(st_tol.relax), "jacobi_bound": float(st_tol.jacobi_bound),
                   "restart_rule": args.restart_rule,
                   "iterations_per_frame": mean_iters,
                   "frames_converged": int(sum(conv)),
                   "final_rel_residual_mean": float(np.mean(resid)),
                   "frames_per_s_fixed_20_iterations": fps20,
                   "persistent_kernel": persist,
                   "fp64_oracle_error_20_iterations": err64,
                   "fp64_oracle_error_to_tolerance": tol_err,
                   "kernel_variant": int(plan.variant), "iters_per_launch": int(ipl),
                   "element_regions": int(plan.element_regions),
                   "tile_verts": int(plan.tile_verts),
                   "block_threads": int(plan.block_threads), "autotune_256": tune256},
        "roofline": {"bound": "hbm", "achieved": bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bw / HBM_PEAK_GBS, "traffic": profiled_traffic(TRAFFIC_FILE, plan),
                     "kernel": kernel_name(plan, "256x256"),
                     "us_per_launch": t_iter * 1e3, "iters_per_launch": ipl,
                     "us_per_iteration": t_iter * 1e3 / ipl, "bytes_per_launch": b_iter,
                     "copy_probe_GBs": copy_gbs, "copy_probe_small": copy_small,
                     "mesh_1024": roof_1024,
                     "batched_512": batch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
