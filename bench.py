"""Benchmark of the projective-dynamics frame on one MI355X.

Workload: a 256x256 cloth (65,536 vertices, 130,050 triangles, 194,565 bend
hinges), 1 m x 1 m, whole top edge pinned, stretch 1e3, bend 1e-2, density
0.1 kg/m^2, dt 1/60, gravity 9.81 m/s^2, seeded +-1e-3 m jitter, uniform wind
0.5 N/m^2 (synthetic, seeded input). A "step" is one simulation frame of plain
PD (Jacobi + Chebyshev on FP64 displacement iterates with FP32 storage of the
frame data) iterated until the FP64 relative residual ||b - K x|| / ||b||
falls to the tolerance (default 1e-5). `value` is frames per second of that
loop with the state resident in HBM.

The roofline object describes the kernel that dominates the run, the fused
per-iteration kernel at 256x256: algorithmic bytes 56 N + 32 T + 40 E per
launch over its average launch time, measured here with HIP events on the
solver's stream, against the MI355X HBM peak of 8 TB/s. `traffic` is the HBM
bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a copy
kernel), read from the committed profile summary. The same figures for the
1024x1024 irregular mesh are reported under roofline.mesh_1024.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(N, T, E):
    return 56 * N + 32 * T + 40 * E


def profiled_traffic(path):
    """HBM bytes per launch recorded in a committed rocprofv3 summary."""
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


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
    sol = ch.Solver(ch.config(max_iterations=iterations, **common), s.N, s.T, s.E)
    sol.step(s.step_args())
    x_gpu, _ = s.get_state()
    m = oracle_mesh(scene_arrays(s))
    p = po.params(mode=po.PD, iterations=iterations, gravity=-9.81, dt=CFG2["dt"],
                  rho=common["rho"], cheb_delay=1, ext_accel=CFG2_WIND,
                  threads=min(os.cpu_count() or 1, 16))
    x64, *_ = po.step(m, p, pos0, vel0, "f64")
    return float(np.abs(x_gpu.astype(np.float64) - x64).max() / extent(pos0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--max-iters", type=int, default=80000)
    ap.add_argument("--check-every", type=int, default=50)
    ap.add_argument("--rho", type=float, default=0.999999)
    ap.add_argument("--variant", type=int, default=0, help="0 auto, 1 unfused, 2 fused")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-1024", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this benchmark measures one GPU; run it with --gpus 1"}))
        sys.exit(2)

    import numpy as np

    import clothhip as ch
    from helpers import CFG2, CFG2_WIND, CFG3

    info = ch.device_info(0)
    log(f"device: {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")

    # ---- 256x256: frames/s to the residual tolerance ---------------------------
    s = build_scene(CFG2)
    N, T, E = s.N, s.T, s.E
    args_c = s.step_args()
    pos0, vel0 = s.get_state()
    common = dict(mode=ch.MODE_PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho, cheb_delay=1,
                  ext_accel=CFG2_WIND, kernel_variant=args.variant)
    sol_tol = ch.Solver(ch.config(max_iterations=args.max_iters, tolerance=args.tol,
                                  check_every=args.check_every, **common), N, T, E)
    log(f"256x256: N={N} T={T} E={E}; warmup {args.warmup} frames")
    for _ in range(args.warmup):
        sol_tol.step(args_c)
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
    fps_tol = args.steps / (ms_total / 1e3)
    ms_per_frame = ms_total / args.steps
    mean_iters = float(np.mean(iters))
    log(f"to tol {args.tol}: {fps_tol:.2f} frames/s, {mean_iters:.0f} iterations/frame, "
        f"{sum(conv)}/{len(conv)} converged")

    # ---- 256x256: fixed 20 iterations per frame --------------------------------
    s.set_state(pos0, vel0)
    sol20 = ch.Solver(ch.config(max_iterations=20, **common), N, T, E)
    for _ in range(3):
        sol20.step(args_c)
    ms20 = sol20.run_frames(args_c, 50)
    fps20 = 50 / (ms20 / 1e3)
    log(f"fixed 20 iterations: {fps20:.1f} frames/s")

    # ---- roofline of the per-iteration kernel ------------------------------------
    t_iter = sol20.time_iteration_kernel(args_c, 500)
    b_iter = algorithmic_bytes(N, T, E)
    bw = b_iter / (t_iter * 1e-3) / 1e9
    plan = sol20.plan_info()
    log(f"iteration kernel: {t_iter * 1e3:.2f} us, {bw:.0f} GB/s algorithmic "
        f"(variant {plan.variant}, tile {plan.tile_verts}, {plan.block_threads} threads)")
    copy_ms = ch.measure_copy(256 << 20, 20)
    copy_gbs = 2 * (256 << 20) / (copy_ms * 1e-3) / 1e9
    log(f"copy probe: {copy_gbs:.0f} GB/s")
    err64 = fp64_error(ch, s, pos0, vel0, common, 20)
    log(f"FP64 oracle error after one 20-iteration frame: {err64:.2e} of extent")

    roof_1024 = None
    if not args.no_1024:
        s3 = build_scene(CFG3)
        a3 = s3.step_args()
        sol3 = ch.Solver(ch.config(max_iterations=30, mode=ch.MODE_PD, gravity=-9.81,
                                   dt=CFG3["dt"], rho=0.99, cheb_delay=1,
                                   kernel_variant=args.variant), s3.N, s3.T, s3.E)
        sol3.step(a3)
        t3 = sol3.time_iteration_kernel(a3, 50)
        b3 = algorithmic_bytes(s3.N, s3.T, s3.E)
        ms3 = sol3.run_frames(a3, 5) / 5
        tr3 = profiled_traffic(TRAFFIC_FILE_1024)
        bw3 = b3 / (t3 * 1e-3) / 1e9
        roof_1024 = {"workload": "1024x1024 irregular mesh, 30 iterations/frame",
                     "N": s3.N, "T": s3.T, "E": s3.E, "us_per_launch": t3 * 1e3,
                     "bytes_per_launch": b3, "achieved": bw3, "frac": bw3 / HBM_PEAK_GBS,
                     "traffic": tr3, "frames_per_s_30_iterations": 1e3 / ms3}
        log(f"1024x1024 iteration: {t3 * 1e3:.1f} us, {bw3:.0f} GB/s "
            f"({100 * bw3 / HBM_PEAK_GBS:.1f}% of peak); {1e3 / ms3:.1f} frames/s at 30 iterations")
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
        solb = ch.Solver(ch.config(max_iterations=15, mode=ch.MODE_PD, gravity=-9.81,
                                   dt=specs[0]["dt"], rho=0.99, cheb_delay=1), sb.N, sb.T, sb.E,
                         cloth_offsets=sb.vert_off)
        t0 = time.time()
        rep = solb.autotune(ab, 10)
        cands = [{"tile_verts": rep.tile_verts[i], "block_threads": rep.block_threads[i],
                  "us_per_launch": rep.us_per_launch[i]} for i in range(rep.num_candidates)]
        log(f"autotune ({time.time() - t0:.1f} s): " + ", ".join(
            f"{c['tile_verts']}/{c['block_threads']}: {c['us_per_launch']:.0f} us" for c in cands))
        for _ in range(2):
            solb.step(ab)
        msb = solb.run_frames(ab, 10) / 10
        tb = solb.time_iteration_kernel(ab, 20)
        bb = algorithmic_bytes(sb.N, sb.T, sb.E)
        bwb = bb / (tb * 1e-3) / 1e9
        pib = solb.plan_info()
        batch = {"workload": "512 cloths, sides round(16*16^U) in [16, 256], stretch log-U "
                             "[1e2, 1e4], bend log-U [1e-3, 1e-1], 1-4 random boundary pins, "
                             "15 iterations/frame, graph-captured frame",
                 "cloths": sb.C, "N": sb.N, "T": sb.T, "E": sb.E,
                 "frames_per_s": 1e3 / msb, "cloth_frames_per_s": sb.C * 1e3 / msb,
                 "ms_per_frame": msb, "us_per_iteration_launch": tb * 1e3,
                 "bytes_per_launch": bb, "achieved": bwb, "frac": bwb / HBM_PEAK_GBS,
                 "autotune": {"best": cands[rep.best], "candidates": cands},
                 "padded_verts": int(pib.padded_verts)}
        log(f"batch: {1e3 / msb:.1f} frames/s ({sb.C * 1e3 / msb:.0f} cloth-frames/s), iteration "
            f"{tb * 1e3:.0f} us = {bwb:.0f} GB/s ({100 * bwb / HBM_PEAK_GBS:.1f}% of peak)")
        del solb, sb

    # ---- CPU baseline: the oracle's FP32 PD on the host cores --------------------
    cpu = None
    if not args.no_cpu:
        import pyoracle as po
        from helpers import oracle_mesh, scene_arrays
        threads = min(os.cpu_count() or 1, 16)
        m = oracle_mesh(scene_arrays(s))
        p = po.params(mode=po.PD, iterations=20, gravity=-9.81, dt=CFG2["dt"], rho=args.rho,
                      cheb_delay=1, ext_accel=CFG2_WIND, threads=threads)
        pos, vel = pos0.copy(), vel0.copy()
        t0 = time.time()
        frames = 0
        while time.time() - t0 < args.cpu_seconds:
            pos, vel, *_ = po.step(m, p, pos, vel, "f32")
            frames += 1
        dt_cpu = time.time() - t0
        it_per_s = frames * 20 / dt_cpu
        cpu = {"value": it_per_s / mean_iters, "unit": "frames/s", "cores": threads,
               "kind": "port",
               "sample": f"{frames} frames x 20 iterations of the 256x256 PD frame (C++ oracle, "
                         f"FP32, OpenMP {threads} threads) in {dt_cpu:.1f} s = "
                         f"{it_per_s:.1f} iterations/s, divided by the GPU's "
                         f"{mean_iters:.0f} iterations/frame to the same tolerance"}
        log(f"cpu baseline: {it_per_s:.1f} iterations/s on {threads} threads")

    line = {
        "metric": "frames/s at 256² to fixed residual, plain vs learned warm-start; % HBM peak",
        "value": fps_tol,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_frame,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 iterate, f32 storage",
        "data": "synthetic (seeded 256x256 cloth)",
        "config": {"workload": "256x256 PD cloth, top edge pinned, wind 0.5 N/m^2, "
                               "plain PD to relative residual 1e-5",
                   "tolerance": args.tol, "residual": "FP64 ||b - K x|| / ||b|| on the GPU",
                   "max_iterations": args.max_iters, "check_every": args.check_every,
                   "rho": args.rho, "iterations_per_frame": mean_iters,
                   "frames_converged": int(sum(conv)),
                   "final_rel_residual_mean": float(np.mean(resid)),
                   "frames_per_s_fixed_20_iterations": fps20,
                   "fp64_oracle_error_20_iterations": err64,
                   "kernel_variant": int(plan.variant), "tile_verts": int(plan.tile_verts),
                   "block_threads": int(plan.block_threads)},
        "roofline": {"bound": "hbm", "achieved": bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bw / HBM_PEAK_GBS, "traffic": profiled_traffic(TRAFFIC_FILE),
                     "kernel": "k_fused_iter (one PD iteration, 256x256)",
                     "us_per_launch": t_iter * 1e3, "bytes_per_launch": b_iter,
                     "copy_probe_GBs": copy_gbs, "mesh_1024": roof_1024,
                     "batched_512": batch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
