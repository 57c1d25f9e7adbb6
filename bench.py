"""Benchmark of the MI355X Z4c RK4 path.

Workload (the robust-stability configuration on one GPU): one box of 192^3
interior points on [-8,8]^3 (h = 1/12) with 3 periodic ghost layers
(198^3 stored points, 25 FP64 grid functions per register), flat space plus
seeded uniform noise of amplitude 1e-10 on every evolved variable,
1+log slicing with shift advection, Gamma driver eta = 2, 4th-order
centred / lop-sided advection finite differences, KO dissipation sigma=0.1,
dt = h/4.  A step is one full RK4 step (4 RHS evaluations, boundary copies,
fused stage updates, algebraic constraints) on device-resident data.

Prints one JSON line (metric = grid-point-updates/s on the GPU) with the
roofline of the dominant kernel (the Z4c RHS bulk kernel, timed with HIP
events on the library stream during the timed region) and a CPU baseline
(the repository's OpenMP CPU oracle on a bounded sample of the same
workload, on the host cores of the GPU machine).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Z4c grid-point-updates/s (FP64, batched FMR) + con2prim pts/s learned vs plain"
FLOP_PER_POINT_RHS = 5.5e3      # algorithmic FP64 FLOP per point of one Z4c RHS
BYTES_PER_POINT_RHS = 400.0     # 25 fields read + 25 written, FP64
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=192, help="interior points per direction")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def flat_noise(nstore, g, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1e-10, 1e-10, (25, nstore, nstore, nstore))
    for c in (0, 3, 5, 6, 18):      # gxx gyy gzz chi alpha
        u[c] += 1.0
    # periodic consistency of the ghost layers (period = interior)
    for ax in (1, 2, 3):
        idx = [slice(None)] * 4
        src = [slice(None)] * 4
        idx[ax] = slice(0, g)
        src[ax] = slice(nstore - 2 * g, nstore - g)
        u[tuple(idx)] = u[tuple(src)]
        idx[ax] = slice(nstore - g, nstore)
        src[ax] = slice(g, 2 * g)
        u[tuple(idx)] = u[tuple(src)]
    return u


def gpu_run(args):
    import baum_amd as B
    g = 3
    n = args.n
    nst = n + 2 * g
    h = 16.0 / n
    dt = 0.25 * h
    box = B.make_box(nst, nst, nst, -8.0 - (g - 0.5) * h, -8.0 - (g - 0.5) * h, -8.0 - (g - 0.5) * h, h,
                     skip=(g,) * 6, bflag=(B.PERBOUND,) * 6)
    lev = B.Level([box], 100)
    lev.enable(range(100))
    U, P, R, V = list(range(25)), list(range(25, 50)), list(range(50, 75)), list(range(75, 100))
    for v in range(100):
        lev.fill(v, 0.0)
    u0 = flat_noise(nst, g, seed=20261202)
    lev.put_state(U, u0.reshape(25, -1))
    del u0
    par = B.z4c_params()
    for _ in range(args.warmup):
        lev.z4c_rk4_step(par, dt, U, P, R, V, pgh=g)
    B.synchronize()
    B.timer_reset()
    B.timer_enable(True)
    B.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lev.z4c_rk4_step(par, dt, U, P, R, V, pgh=g)
    B.synchronize()
    t1 = time.perf_counter()
    B.timer_enable(False)
    ms_rhs, n_rhs = B.timer_get("z4_rhs")
    ms_upd, n_upd = B.timer_get("rk4_update")
    ms_bc, n_bc = B.timer_get("set_boundary")
    ms_alg, n_alg = B.timer_get("algcon")
    ms_cpy, n_cpy = B.timer_get("vladd")
    l2 = lev.z4c_ham_l2(U)
    lev.free()
    return dict(n=n, h=h, dt=dt, elapsed=t1 - t0, ms_rhs=ms_rhs, n_rhs=n_rhs, ms_upd=ms_upd, n_upd=n_upd,
                ms_bc=ms_bc, ms_alg=ms_alg, ms_cpy=ms_cpy, ham_l2=l2)


def cpu_baseline(target_s):
    """the OpenMP CPU oracle (test infrastructure, used here only as the
    reported baseline) on a 48^3-interior periodic noise box"""
    import oracle as O
    g = 3
    n = 48
    nst = n + 2 * g
    h = 16.0 / 192
    ob = O.make_box(nst, nst, nst, 0.0, 0.0, 0.0, h, skip=(g,) * 6)
    u = flat_noise(nst, g, seed=7)
    p = O.z4c_params()
    work = np.zeros((75, nst, nst, nst))
    steps = 0
    t0 = time.perf_counter()
    while True:
        O.z4c_rk4_step(ob, p, 0.25 * h, u, bc=2, pgh=g, work=work)
        steps += 1
        el = time.perf_counter() - t0
        if el >= target_s or steps >= 200:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=steps * n ** 3 / el, unit="grid-point-updates/s", cores=cores, kind="port",
                sample="%d RK4 steps of the same Z4c configuration on a 48^3-interior periodic box "
                       "(OpenMP C oracle), %.1f s" % (steps, el))


def traffic_from_profile():
    """HBM bytes per RHS launch from the committed rocprofv3 PMC summary, if any"""
    path = os.path.join(ROOT, "profiles", "z4c_rhs_bulk_192_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1].split()[0])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU box decomposition is a design "
              "described in DESIGN.md, not an implemented path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    r = gpu_run(args)
    npts = r["n"] ** 3
    value = npts * args.steps / r["elapsed"]
    launch_ms = r["ms_rhs"] / max(r["n_rhs"], 1)
    flops = FLOP_PER_POINT_RHS * npts
    achieved = flops / (launch_ms * 1e-3) / 1e12
    roof = {"bound": "fp64", "kernel": "z4c_rhs_bulk", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profile(),
            "launch_ms": launch_ms, "algorithmic_flop_per_launch": flops,
            "hbm_gbs_algorithmic": BYTES_PER_POINT_RHS * npts / (launch_ms * 1e-3) / 1e9}
    out = {"metric": METRIC, "value": value, "unit": "grid-point-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * r["elapsed"] / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "Z4c RK4, one %d^3 periodic box (+3 ghosts), flat + 1e-10 noise, "
                                  "1+log / Gamma driver, 4th-order FD, KO 0.1, dt=h/4" % r["n"],
                      "points_per_step": npts, "parallelism": "single GPU"},
           "roofline": roof,
           "breakdown_ms_per_step": {"z4c_rhs": r["ms_rhs"] / args.steps, "rk4_update": r["ms_upd"] / args.steps,
                                     "set_boundary": r["ms_bc"] / args.steps, "algcon": r["ms_alg"] / args.steps,
                                     "copy_up": r["ms_cpy"] / args.steps},
           "ham_l2_after": r["ham_l2"]}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
