"""Benchmark of the projective-dynamics frame on one MI355X.

Workload (default): config 2 — a 256x256 cloth (65,536 vertices, 130,050
triangles, 194,565 bend hinges), top row pinned, stretch 1e3, bend 1e-2,
density 0.1, dt 1/60, gravity 9.81 and 0.5 N/m^2 wind, synthetic seeded
input. A "step" is one simulation frame of plain PD (Jacobi + Chebyshev,
FP32 storage and arithmetic, displacement iterate) run until the FP64
relative residual ||b - K x|| / ||b|| falls to the tolerance (default 1e-5)
or the iteration cap. `value` is frames per second of that loop with the
state resident in HBM.

Also reported: frames/s at the config's fixed 20 iterations, and the
roofline of the per-iteration kernels at 256x256 and at config 3 (1024x1024
irregular mesh): algorithmic bytes per iteration 56 N + 32 T + 40 E over the
measured time per iteration, against the MI355X HBM peak of 8 TB/s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "cloth-cuda_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)
os.environ.setdefault("CLOTH_QUIET", "1")

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(N, T, E):
    return 56 * N + 32 * T + 40 * E


def build_scene(spec):
    import clothhip as ch
    s = ch.Scene(**spec)
    s.upload()
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--max-iters", type=int, default=20000)
    ap.add_argument("--check-every", type=int, default=50)
    ap.add_argument("--rho", type=float, default=0.9999)
    ap.add_argument("--variant", type=int, default=0, help="0 auto, 1 unfused, 2 fused")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cfg3", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np

    import clothhip as ch
    from helpers import CFG2, CFG2_WIND, CFG3

    info = ch.device_info(0)
    log(f"device: {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")

    # ---- config 2: frames/s to the residual tolerance ------------------------
    s = build_scene(CFG2)
    N, T, E = s.N, s.T, s.E
    args_c = s.step_args()
    pos0, vel0 = s.get_state()
    common = dict(mode=ch.MODE_PD, gravity=-9.81, dt=CFG2["dt"], rho=args.rho, cheb_delay=1,
                  ext_accel=CFG2_WIND, kernel_variant=args.variant)
    sol_tol = ch.Solver(ch.config(max_iterations=args.max_iters, tolerance=args.tol,
                                  check_every=args.check_every, **common), N, T, E)
    log(f"cfg2: N={N} T={T} E={E}; warmup {args.warmup} frames")
    for _ in range(args.warmup):
        sol_tol.step(args_c)
    iters = []
    ms_total = 0.0
    resid = []
    for f in range(args.steps):
        ms = sol_tol.run_frames(args_c, 1)
        st = sol_tol.stats()
        iters.append(st.iterations)
        resid.append(st.residual / st.rhs_norm if st.rhs_norm > 0 else 0.0)
        ms_total += ms
        if f % 5 == 0:
            log(f"  frame {f}: {st.iterations} iterations, rel residual {resid[-1]:.3e}, "
                f"{ms:.2f} ms")
    fps_tol = args.steps / (ms_total / 1e3)
    ms_per_frame = ms_total / args.steps
    mean_iters = float(np.mean(iters))
    log(f"cfg2 to tol {args.tol}: {fps_tol:.1f} frames/s, {mean_iters:.0f} iterations/frame")

    # ---- config 2: fixed 20 iterations (the config's own setting) -----------
    s.set_state(pos0, vel0)
    sol20 = ch.Solver(ch.config(max_iterations=20, **common), N, T, E)
    for _ in range(3):
        sol20.step(args_c)
    ms20 = sol20.run_frames(args_c, 50)
    fps20 = 50 / (ms20 / 1e3)
    log(f"cfg2 fixed 20 iterations: {fps20:.1f} frames/s")

    # ---- roofline of the per-iteration kernels -------------------------------
    t_iter2 = sol20.time_iteration_kernel(args_c, 200)
    b2 = algorithmic_bytes(N, T, E)
    bw2 = b2 / (t_iter2 * 1e-3) / 1e9
    plan2 = sol20.plan_info()
    log(f"cfg2 iteration: {t_iter2 * 1e3:.2f} us, {bw2:.0f} GB/s algorithmic "
        f"(variant {plan2.variant})")
    copy_ms = ch.measure_copy(256 << 20, 20)
    copy_gbs = 2 * (256 << 20) / (copy_ms * 1e-3) / 1e9
    log(f"copy probe: {copy_gbs:.0f} GB/s")

    roof3 = None
    if not args.no_cfg3:
        s3 = build_scene(CFG3)
        a3 = s3.step_args()
        sol3 = ch.Solver(ch.config(max_iterations=30, mode=ch.MODE_PD, gravity=-9.81,
                                   dt=CFG3["dt"], rho=0.99, cheb_delay=1,
                                   kernel_variant=args.variant), s3.N, s3.T, s3.E)
        sol3.step(a3)
        t3 = sol3.time_iteration_kernel(a3, 50)
        b3 = algorithmic_bytes(s3.N, s3.T, s3.E)
        ms3 = sol3.run_frames(a3, 5) / 5
        roof3 = dict(N=s3.N, T=s3.T, E=s3.E, us_per_iter=t3 * 1e3,
                     achieved_GBs=b3 / (t3 * 1e-3) / 1e9, bytes_per_iter=b3,
                     frames_per_s_30it=1e3 / ms3)
        log(f"cfg3 iteration: {t3 * 1e3:.1f} us, {roof3['achieved_GBs']:.0f} GB/s; "
            f"{roof3['frames_per_s_30it']:.1f} frames/s at 30 iterations")
        del sol3, s3

    # ---- CPU baseline: the oracle's FP32 PD on the host cores ----------------
    cpu = None
    if not args.no_cpu:
        import pyoracle as po
        from helpers import oracle_mesh, scene_arrays
        threads = min(os.cpu_count() or 1, 16)
        arr = scene_arrays(s)
        m = oracle_mesh(arr)
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
               "sample": f"{frames} frames x 20 iterations of 256x256 PD FP32 (oracle, OpenMP "
                         f"{threads} threads) in {dt_cpu:.1f} s = {it_per_s:.1f} iterations/s, "
                         f"divided by the GPU's {mean_iters:.0f} iterations/frame to the same "
                         f"tolerance"}
        log(f"cpu baseline: {it_per_s:.1f} iterations/s on {threads} threads")

    line = {
        "metric": "frames/s at 256^2 to fixed residual (plain PD); % HBM peak",
        "value": fps_tol,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_frame,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 256x256 cloth, config 2)",
        "config": {"workload": "config 2: 256x256 PD cloth, top row pinned, wind 0.5 N/m^2",
                   "tolerance": args.tol, "residual": "FP64 ||b - K x|| / ||b||",
                   "max_iterations": args.max_iters, "check_every": args.check_every,
                   "rho": args.rho, "iterations_per_frame": mean_iters,
                   "final_rel_residual_mean": float(np.mean(resid)),
                   "frames_per_s_fixed_20_iterations": fps20,
                   "kernel_variant": int(plan2.variant)},
        "roofline": {"bound": "hbm", "achieved": bw2, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bw2 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "one PD iteration at 256x256",
                     "us_per_iteration": t_iter2 * 1e3, "bytes_per_iteration": b2,
                     "copy_probe_GBs": copy_gbs, "config3": roof3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
