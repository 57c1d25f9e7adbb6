"""Benchmark of the MI355X Z4c / FMR / con2prim path (one JSON line).

Workloads (synthetic data of the configurations' shapes, FP64 throughout):

* fmr (the headline `value`): Berger-Oliger Z4c evolution of a 7-level FMR
  hierarchy.  Level 0 is one 128^3 box on [-15.875, 15.875]^3 (h0 = 0.25,
  radiative boundary); levels 1..6 each hold two boxes around a
  Brill-Lindquist binary (m = 0.5 at x = -3, +3) with seeded uneven sizes in
  48..96 per direction (clipped so that every box nests in its parent with
  a 4-cell margin and the two boxes of a level do not overlap).  Precollapsed
  lapse, zero shift; 1+log / Gamma driver, 4th-order FD with lop-sided
  advection, KO 0.1 on level 0 and 0.02 on levels 1..6 (the reference's
  get_dissipation_factor with amr_move_lcube 0), chi floor 1e-5 in the
  undifferentiated 1/chi terms, dt_l = h_l / 4 with subcycling, time interpolation of
  order 3, amr_nbuffer 6, order_RP 4.  A step is one coarse step (level 0
  once, level l 2^l times, each followed by restriction / prolongation with
  its parent), captured once as a hipGraph and replayed.  Unit: interior
  grid-point updates per second, sum over levels of 2^l x interior points.
* box192: one 192^3 periodic box (+3 ghosts), flat space + 1e-10 noise,
  full RK4 steps; it times the Z4c RHS bulk kernel for the `roofline`.
* c1: the single-box configuration (BL m = 1, 64^3 interior on [-8, 8]^3,
  8 RK4 steps), on the device and with the CPU oracle.
* box192 also runs the robust-stability test: 100 RK4 steps from the noise
  data with the Hamiltonian constraint's L2 norm every 10 steps.
* c2p: con2prim on 2^24 points of the con2prim distribution (ideal gas,
  Gamma 2, tolerance 1e-12): the reference's plain p-root, the bracketed
  solver from a cold start, and the learned guess (3-64-64-1 network trained
  on the GPU on 2^22 solved samples) polished by the bracketed solver; the
  analytic recovery grhd_c2p_vanal against the original primitives (on the
  distribution's metric and on a flat one); the CPU oracle's hybrid and
  bracketed solvers on 2^20 points.

`roofline` is for the dominant kernel, z4c_rhs_bulk: achieved = 5.5e3
algorithmic FP64 FLOP per point x points per launch / the launch's average
duration from HIP events on the library stream.  `fmr.rhs_fp64` gives the
same rate for the headline workload: 5.5e3 FLOP x 4 RHS evaluations per point
update x the updates of one coarse step / the time of all its RHS launches
(bulk and shell tiles, RK4 update fused into the store) in the eager split.
`cpu_baseline` times the
OpenMP CPU oracle (the restatement used as the parity checker) on a bounded
sample of the fmr workload's per-box work on the host cores.
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
BINARY = ((0.5, (-3.0, 0.0, 0.0)), (0.5, (3.0, 0.0, 0.0)))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed coarse steps of the fmr workload")
    ap.add_argument("--warmup", type=int, default=2, help="untimed coarse steps (the second is captured)")
    ap.add_argument("--box-steps", type=int, default=10)
    ap.add_argument("--rhs-launches", type=int, default=5, help="plain RHS launches timed for the roofline")
    ap.add_argument("--c2p-n", type=int, default=1 << 24)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rhs-overlap", type=int, default=1,
                    help="1: box-edge shell kernels concurrent with the bulk kernel (second stream)")
    ap.add_argument("--rk-fusion", type=int, default=2,
                    help="RK4 stage update inside the RHS kernels: 0 never, 1 always, 2 on levels >= 2^19 points")
    ap.add_argument("--xcd-map", type=int, default=1, help="1: XCD-aware tile numbering in the RHS kernels")
    ap.add_argument("--only", default="fmr,box192,c1,c2p,c5")
    return ap.parse_args()


# ------------------------------------------------------------------ data
def puncture_fields(x, y, z):
    """BL binary: chi = psi^-4, precollapsed lapse psi^-2, flat gt, zero rest"""
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    psi = np.ones_like(X)
    for m, c in BINARY:
        psi += m / (2 * np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2))
    u = np.zeros((25,) + X.shape)
    u[0] = u[3] = u[5] = 1.0
    u[6] = psi ** -4
    u[18] = psi ** -2
    return u


def c3_geometry(make_box, phy, ref, seed=20261203):
    """boxes per level and the parent of every box (see the module doc)"""
    rng = np.random.default_rng(seed)
    n0, H0 = 128, 0.25
    x0 = -(n0 - 1) / 2 * H0
    levels = [[dict(m=n0, n=n0, o=n0, x0=x0, y0=x0, z0=x0, h=H0)]]
    parents = [[0]]
    for l in range(1, 7):
        boxes, par = [], []
        for b, cx in enumerate((-3.0, 3.0)):
            p = levels[l - 1][b if l > 1 else 0]
            H = p["h"]
            size = []
            for d, key in enumerate(("m", "n", "o")):
                s = int(rng.integers(24, 49)) * 2
                s = min(s, 2 * (p[key] - 8))
                if l == 1 and d == 0:
                    s = min(s, int(round(6.0 / (H / 2))) // 2 * 2)
                size.append(s)
            c0 = []
            for d, (key, ctr) in enumerate((("m", cx), ("n", 0.0), ("o", 0.0))):
                org = p["x0" if d == 0 else ("y0" if d == 1 else "z0")]
                c = int(round((ctr - org) / H - (size[d] - 2) / 4.0))
                c = max(4, min(c, p[key] - 4 - size[d] // 2))
                c0.append(c)
            box = dict(m=size[0], n=size[1], o=size[2], h=H / 2,
                       x0=p["x0"] + (c0[0] - 0.25) * H, y0=p["y0"] + (c0[1] - 0.25) * H,
                       z0=p["z0"] + (c0[2] - 0.25) * H)
            boxes.append(box)
            par.append(b if l > 1 else 0)
        # the two boxes of a level must not overlap in x
        a, bb = boxes
        gap = bb["x0"] - (a["x0"] + (a["m"] - 1) * a["h"])
        if gap <= 0:
            shift = int(np.ceil((-gap + a["h"]) / (2 * a["h"])))
            bb["x0"] += shift * 2 * a["h"]
        levels.append(boxes)
        parents.append(par)
    out = []
    for l, bl in enumerate(levels):
        fl = phy if l == 0 else ref
        out.append([make_box(b["m"], b["n"], b["o"], b["x0"], b["y0"], b["z0"], b["h"], bflag=(fl,) * 6)
                    for b in bl])
    return out, parents


def inner_points(boxes):
    return sum((b.m - 2) * (b.n - 2) * (b.o - 2) for b in boxes)


# ------------------------------------------------------------------ fmr
def run_fmr(args):
    import baum_amd as B
    U, UP, UPP, R, V = (list(range(25 * k, 25 * k + 25)) for k in range(5))
    FP, FR, NV = 125, 126, 127
    boxes, parents = c3_geometry(B.make_box, B.PHYBOUND, B.REFBOUND)
    levels = []
    for l, bl in enumerate(boxes):
        lv = B.Level(bl, NV, l=l)
        lv.enable(range(NV))
        for v in range(NV):
            lv.fill(v, 0.0)
        u = np.concatenate([puncture_fields(b.x0 + b.dx * np.arange(b.m), b.y0 + b.dy * np.arange(b.n),
                                            b.z0 + b.dz * np.arange(b.o)).reshape(25, -1) for b in bl], axis=1)
        lv.put_state(U, u)
        del u
        levels.append(lv)
    for l in range(1, len(levels)):
        B.set_rp_flags(levels[l - 1], levels[l], parents[l], 6, FP, FR)
    # chi floor of the reference's binary-puncture parameter files (z4_chi_div_floor = 1e-5):
    # without it the 1/chi terms overflow at the punctures on levels >= 4
    # KO 0.1 with the reference's level rule (get_dissipation_factor with
    # amr_move_lcube = 0 and dissipation_factor_move unset: 0.2 x on levels >= 1)
    pars = []
    for l in range(len(levels)):
        p = B.z4c_params(z4_chi_div_floor=1e-5)
        p.dissfactor = B.dissipation_factor(l, 0.1)
        pars.append(p)
    dt0 = 0.25 * boxes[0][0].dx
    # the punctures are tracked on the finest level after each of its steps (analyze_level)
    punc0 = np.array([c for _, c in BINARY], dtype=np.float64)
    punc = B.DeviceArray.from_host(punc0)
    H = B.Hierarchy(levels, pars, dt0, U, UP, UPP, R, V, order_RP=4, timeinterpolation=3, punctures=punc)
    H.set_check(0)
    updates = sum(2 ** l * inner_points(bl) for l, bl in enumerate(boxes))
    # warm-up: the first coarse step eagerly (reduced-order time interpolation), then capture the next
    H.advance(0)
    g = H.capture(0)
    H.launch(g)
    for _ in range(max(args.warmup - 2, 0)):
        H.launch(g)
    B.synchronize()
    assert H.nonfinite(reset=False) == 0, "non-finite values in the warm-up steps"
    B.timer_reset()
    B.timer_enable(True)
    B.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        H.launch(g)
    B.synchronize()
    t1 = time.perf_counter()
    B.timer_enable(False)
    bad = H.nonfinite(reset=False)
    B.Hierarchy.free_graph(g)
    # per-kernel split from an eager coarse step (timers pause inside graphs)
    B.timer_reset()
    B.timer_enable(True)
    H.advance(0)
    B.synchronize()
    B.timer_enable(False)
    split = {k: B.timer_get(k)[0] for k in ("z4_rhs", "z4_rhs_edge", "rk4_update", "vladd", "algcon",
                                            "set_boundary", "prolong", "restrict", "puncture")}
    desc = [[(b.m, b.n, b.o) for b in bl] for bl in boxes]
    punc_end = punc.get()
    punc.free()
    for lv in levels:
        lv.free()
    el = t1 - t0
    return dict(value=updates * args.steps / el, ms_per_step=1e3 * el / args.steps, updates_per_step=updates,
                nonfinite=bool(bad), boxes=desc, eager_split_ms=split,
                punctures_end=[[float("%.9g" % c) for c in p] for p in punc_end])


# ------------------------------------------------------------------ box192
def flat_noise(nstore, g, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1e-10, 1e-10, (25, nstore, nstore, nstore))
    for c in (0, 3, 5, 6, 18):      # gxx gyy gzz chi alpha
        u[c] += 1.0
    for ax in (1, 2, 3):            # periodic ghosts (period = interior)
        idx = [slice(None)] * 4
        src = [slice(None)] * 4
        idx[ax] = slice(0, g)
        src[ax] = slice(nstore - 2 * g, nstore - g)
        u[tuple(idx)] = u[tuple(src)]
        idx[ax] = slice(nstore - g, nstore)
        src[ax] = slice(g, 2 * g)
        u[tuple(idx)] = u[tuple(src)]
    return u


def run_box(args):
    import baum_amd as B
    g, n = 3, 192
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
    lev.z4c_rk4_step(par, dt, U, P, R, V, pgh=g)
    B.synchronize()
    B.timer_reset()
    B.timer_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.box_steps):
        lev.z4c_rk4_step(par, dt, U, P, R, V, pgh=g)
    B.synchronize()
    t1 = time.perf_counter()
    B.timer_enable(False)
    t = {k: B.timer_get(k) for k in ("z4_rhs", "rk4_update", "set_boundary", "algcon", "vladd")}
    l2 = lev.z4c_constraints_l2(U)
    # the roofline kernel: the plain RHS (the bulk tile kernel over the whole
    # 192^3 interior, as on the FMR levels), timed alone on the same data
    lev.z4c_rhs(par, R, U)
    B.synchronize()
    B.timer_reset()
    B.timer_enable(True)
    for _ in range(args.rhs_launches):
        lev.z4c_rhs(par, R, U)
    B.synchronize()
    B.timer_enable(False)
    ms, cnt = B.timer_get("z4_rhs")
    # the robust-stability run: 100 RK4 steps from the noise data, the
    # Hamiltonian constraint's L2 norm every 10 steps
    lev.put_state(U, flat_noise(nst, g, seed=20261202).reshape(25, -1))
    ham_hist = [lev.z4c_constraints_l2(U)[0]]
    for step in range(1, 101):
        lev.z4c_rk4_step(par, dt, U, P, R, V, pgh=g)
        if step % 10 == 0:
            ham_hist.append(lev.z4c_constraints_l2(U)[0])
    lev.free()
    npts = n ** 3
    return dict(ham_l2_every_10_steps=[float("%.6g" % x) for x in ham_hist],
                value=npts * args.box_steps / (t1 - t0), ms_per_step=1e3 * (t1 - t0) / args.box_steps,
                rhs_launch_ms=ms / max(cnt, 1), npts=npts, ham_mom_l2_after=[float(x) for x in l2],
                breakdown_ms_per_step={k: v[0] / args.box_steps for k, v in t.items()})


# ------------------------------------------------------------------ c2p
def c2p_prims(n, seed, metric_noise=1e-3):
    """rho log-uniform [1e-10, 1e-2], eps log-uniform [1e-6, 1], W uniform
    [1, 10] in a random direction, metric flat + 1e-3 seeded perturbation"""
    rng = np.random.default_rng(seed)
    rho = 10.0 ** rng.uniform(-10, -2, n)
    eps = 10.0 ** rng.uniform(-6, 0, n)
    W = rng.uniform(1.0, 10.0, n)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    gm = np.zeros((6, n))
    gm[0] = gm[3] = gm[5] = 1.0
    gm += metric_noise * rng.uniform(-1, 1, (6, n))
    gdd = (gm[0] * d[0] ** 2 + gm[3] * d[1] ** 2 + gm[5] * d[2] ** 2
           + 2 * (gm[1] * d[0] * d[1] + gm[2] * d[0] * d[2] + gm[4] * d[1] * d[2]))
    v = np.sqrt((1.0 - 1.0 / W ** 2) / gdd) * d
    return np.array([rho, eps, v[0], v[1], v[2]]), gm


def run_c2p(args):
    import baum_amd as B
    P = B.c2p_params(errormax=1e-12, countmax=100, rhoatm=1e-20)

    def cons_of(n, seed):
        prim5, gm = c2p_prims(n, seed)
        dp, dg = B.DeviceArray.from_host(prim5), B.DeviceArray.from_host(gm)
        dc = B.DeviceArray((5, n))
        B.p2c_batch(P, dp, dg, dc)
        dp.free()
        return dc, dg, prim5[0]

    def timed(fn, reps=3):
        best = 1e30
        for _ in range(reps):
            B.synchronize()
            B.timer_reset()
            B.timer_enable(True)
            fn()
            B.synchronize()
            B.timer_enable(False)
            best = min(best, B.timer_get("c2p")[0])
        return best

    n = args.c2p_n
    dc0, dg, rho = cons_of(n, 4242)
    cons0 = dc0.get()
    dc = B.DeviceArray((5, n))
    dpr = B.DeviceArray((8, n))
    ds, di = B.DeviceArray((n,), np.int32), B.DeviceArray((n,), np.int32)
    out = {"n": n, "tolerance": 1e-12}

    def reset():
        dc.put(cons0)
        pr = np.zeros((8, n))
        pr[0] = rho * 0.9    # density guess of the reference cold branch
        dpr.put(pr)

    for name, method in (("plain", B.C2P_PLAIN), ("bracketed", B.C2P_BRACKETED)):
        def go():
            reset()
            B.c2p_batch(P, method, dc, dg, dpr, ds, di)
        ms = timed(go)
        st, it = ds.get(), di.get()
        out[name] = dict(ms=ms, pts_per_s=n / (ms * 1e-3), converged=float(np.mean(st == B.C2P_OK)),
                         mean_iters=float(it.mean()), max_iters=int(it.max()),
                         iter_hist=np.bincount(np.minimum(it, 12), minlength=13).tolist())
    p_brk = dpr.get()[5]
    # learned guess: training data solved by the bracketed solver, training on the device
    ntr = 1 << 22
    tc, tg, _ = cons_of(ntr, 777)
    tp = B.DeviceArray.from_host(np.zeros((8, ntr)))
    ts, ti = B.DeviceArray((ntr,), np.int32), B.DeviceArray((ntr,), np.int32)
    B.c2p_batch(P, B.C2P_BRACKETED, tc, tg, tp, ts, ti)
    feat, tgt = B.DeviceArray((ntr, 3), np.float32), B.DeviceArray((ntr,), np.float32)
    B.mlp_prepare(P, tc, tg, tp, ts, feat, tgt)
    f, t = feat.get(), tgt.get()
    ok = np.isfinite(t)
    mu, sd = f[ok].mean(axis=0), f[ok].std(axis=0)
    nrm = np.array([mu[0], mu[1], mu[2], 1 / sd[0], 1 / sd[1], 1 / sd[2], t[ok].mean(), t[ok].std()], np.float32)
    net = B.Mlp(seed=11)
    net.set(None, nrm)
    B.synchronize()
    t0 = time.perf_counter()
    losses = net.train(feat, tgt, epochs=30, batch=4096, lr=3e-3, seed=12)
    B.synchronize()
    train_s = time.perf_counter() - t0
    for a in (tc, tg, tp, ts, ti, feat, tgt):
        a.free()
    fb = B.DeviceArray.from_host(np.zeros(1, np.int32))

    def go_l():
        reset()
        fb.put(np.zeros(1, np.int32))
        B.c2p_learned_batch(P, net, dc, dg, dpr, ds, di, fb, polish_max=8)
    ms = timed(go_l)
    st, it, pl = ds.get(), di.get(), dpr.get()[5]
    agree = float(np.max(np.abs(pl / p_brk - 1)))
    out["learned"] = dict(ms=ms, pts_per_s=n / (ms * 1e-3), converged=float(np.mean(st == B.C2P_OK)),
                          mean_iters=float(it.mean()), max_iters=int(it.max()),
                          iter_hist=np.bincount(np.minimum(it, 12), minlength=13).tolist(),
                          fallback_fraction=float(fb.get()[0]) / n, max_rel_diff_p_vs_bracketed=agree,
                          training=dict(samples=ntr, epochs=30, batch=4096, seconds=train_s,
                                        loss_first=float(losses[0]), loss_last=float(losses[-1])))
    # the analytic recovery grhd_c2p_vanal as a cross-check of the primitives
    out["vanal"] = {name: vanal_report(B, P, 1 << 22, 4343, noise)
                    for name, noise in (("c4_metric", 1e-3), ("flat_metric", 0.0))}
    out["learned_vs_plain_speedup"] = out["learned"]["pts_per_s"] / out["plain"]["pts_per_s"]
    out["learned_vs_bracketed_speedup"] = out["learned"]["pts_per_s"] / out["bracketed"]["pts_per_s"]
    for a in (dc0, dg, dc, dpr, ds, di, fb):
        a.free()
    return out, net


# ------------------------------------------------------------------ C5: coupled Z4c + GRHD
def run_move():
    """move_box on a 5-level hierarchy of 96^3 boxes nested around one
    puncture (coarse 192^3, h0 = 1): the puncture jumps by one coarse cell
    and every fine level is rebuilt (allocation, flags, prolongation
    everywhere of u, up, upp, overlap copy of 75 variables)"""
    import time
    import baum_amd as B
    nvar, nxyz, nlev = 127, 96, 5
    U, UP, UPP, R, V = (list(range(25 * k, 25 * k + 25)) for k in range(5))
    levels = [B.Level([B.make_box(192, 192, 192, -95.5, -95.5, -95.5, 1.0, bflag=(B.PHYBOUND,) * 6)], nvar)]
    # amr_move_nxyz 96 with amr_buffer_addextra = no: 96^3 fine boxes
    sizes = B.NestedSizes(nx=192, move_nxyz=nxyz, nbuffer=6, addextra=False)
    for l in range(1, nlev):
        hc = levels[-1].boxes[0].dx
        bb, ib = B.find_nested_bboxes(sizes, l, [[0.0, 0.0, 0.0]], hc)
        d = hc / 2
        b = B.make_box(2 * (ib[0][1] + 1), 2 * (ib[0][3] + 1), 2 * (ib[0][5] + 1), bb[0][0] - d / 2, bb[0][2] - d / 2,
                       bb[0][4] - d / 2, d, bflag=(B.REFBOUND,) * 6)
        levels.append(B.Level([b], nvar, l=l))
    for lv in levels:
        lv.enable(range(nvar))
        for v in range(nvar):
            lv.fill(v, 1.0 if v in (6, 18) else 0.0)
    for l in range(1, nlev):
        B.set_rp_flags(levels[l - 1], levels[l], [0], 6, 125, 126)
    punc = B.DeviceArray.from_host(np.zeros((1, 3)))
    H = B.Hierarchy(levels, [B.z4c_params()] * nlev, 0.25, U, UP, UPP, R, V, punctures=punc, sizes=sizes,
                    rp_flags=(6, 125, 126))
    res = []
    for shift in (1.0, 2.0):
        punc.put(np.array([[shift, 0.0, 0.0]]))
        B.synchronize()
        t0 = time.perf_counter()
        moved = H.move_box(1)
        B.synchronize()
        res.append({"shift": shift, "moved": moved, "ms": 1e3 * (time.perf_counter() - t0)})
    pts = sum(lv.npoints for lv in H.levels[1:])
    for lv in H.levels:
        lv.free()
    punc.free() if hasattr(punc, "free") else None
    return {"workload": "5 levels of 96^3 around one puncture, the puncture moved by one coarse cell of level 0; "
                        "all 4 fine levels rebuilt", "fine_points": pts, "moves": res}


def run_c5(args, net):
    """the coupled Z4c + GRHD system on a 128^3 box holding a TOV star
    (Gamma 2, K 100, rho_c 1.28e-3, 1e-4 density noise): one right-hand side
    evaluation and one RK4 step, each with the reference hybrid c2p, the
    bracketed solver and the learned guess"""
    import baum_amd as B
    n, L = 128, 16.0
    h = 2 * L / n
    ns = n + 2
    x = -L - 0.5 * h + h * np.arange(ns)
    K, G, rhoc = 100.0, 2.0, 1.28e-3
    hc = float(np.log1p(G * K / (G - 1) * rhoc ** (G - 1)))     # log enthalpy of rho_c
    star = B.tov_h(hc, K, G, 100000)
    UZ, UQ = list(range(25)), list(range(25, 30))
    RZ, RQ = list(range(30, 55)), list(range(55, 60))
    ADMG, ADMK, PRIM, MADM = list(range(60, 66)), list(range(66, 72)), list(range(72, 80)), list(range(80, 91))
    UP, V = list(range(91, 121)), list(range(121, 151))
    lev = B.Level([B.make_box(ns, ns, ns, x[0], x[0], x[0], h)], 151)
    lev.enable(range(151))
    # the star on the grid (set_TOV): psi^4, alpha and rho read back for the noisy matter below
    lev.tov_set(star, ADMG, ADMK, 18, PRIM[:6])
    psi4, alpha, rho_star = lev.get(ADMG[0]), lev.get(18), lev.get(PRIM[0])
    M = star["M"]
    rhoatm = 1e-10 * rhoc
    rho0 = np.maximum(rho_star, rhoatm)
    rng = np.random.default_rng(20261202)
    rho = rho0 * (1 + 1e-4 * rng.standard_normal(rho0.shape))
    eps = K * rho ** (G - 1) / (G - 1)
    sg = psi4 ** 1.5
    npts = ns ** 3
    for v in range(151):
        lev.fill(v, 0.0)

    def reset_geometry():
        for v in range(25):
            lev.fill(v, 1.0 if v in (0, 3, 5) else 0.0)
        lev.put(6, 1.0 / psi4)
        lev.put(18, alpha)

    reset_geometry()
    cons = np.array([sg * rho, sg * rho * eps])       # static fluid: D, tau; S_i = 0
    z4 = B.z4c_params()
    P = B.c2p_params(eos="ideal", gamma=G, K=K, rhoatm=rhoatm, errormax=1e-12, countmax=100)
    hp = B.hrsc_params(eos="ideal", gamma=G, K=K)
    status, iters = B.DeviceArray((npts,), np.int32), B.DeviceArray((npts,), np.int32)
    fb = B.DeviceArray.from_host(np.zeros(1, np.int32))
    out = {"grid": f"{n}^3 interior + 1 boundary layer, h = {h}", "tov": {"M": star["M"], "R_schwarzschild": star["R"],
                                                                         "R_isotropic": star["Riso"]}}

    def reset():
        lev.put(UQ[0], cons[0])
        lev.put(UQ[1], cons[1])
        for v in UQ[2:]:
            lev.fill(v, 0.0)
        lev.put(PRIM[0], rho0)
        lev.put(PRIM[5], K * rho0 ** G)
        fb.put(np.zeros(1, np.int32))

    res_p = {}
    for name, method in (("hybrid", B.C2P_HYBRID), ("bracketed", B.C2P_BRACKETED), ("learned", B.C2P_LEARNED)):
        kw = dict(method=method, mlp=net if method == B.C2P_LEARNED else None, fallbacks=fb)
        reset()
        lev.z4c_grhd_rhs(z4, P, hp, RZ + RQ, UZ + UQ, ADMG, ADMK, PRIM, MADM, status, iters, **kw)   # tunes
        reps, tot, c2p_ms = 3, [], []
        for _ in range(reps):
            reset()
            B.synchronize()
            B.timer_reset()
            B.timer_enable(True)
            t0 = time.perf_counter()
            lev.z4c_grhd_rhs(z4, P, hp, RZ + RQ, UZ + UQ, ADMG, ADMK, PRIM, MADM, status, iters, **kw)
            B.synchronize()
            tot.append(time.perf_counter() - t0)
            B.timer_enable(False)
            c2p_ms.append(B.timer_get("c2p")[0])
        st, it = status.get(), iters.get()
        res_p[name] = (lev.get(PRIM[5]), st)
        out[name] = dict(rhs_ms=1e3 * min(tot), c2p_ms=min(c2p_ms),
                         status_hist=np.bincount(st, minlength=5).tolist(), mean_iters=float(it.mean()),
                         max_iters=int(it.max()))
        if method == B.C2P_LEARNED:
            out[name]["fallbacks"] = int(fb.get()[0])
    # one RK4 step of the coupled system (four RHS evaluations, four c2p)
    dt = 0.25 * h
    for name, method in (("hybrid", B.C2P_HYBRID), ("bracketed", B.C2P_BRACKETED), ("learned", B.C2P_LEARNED)):
        kw = dict(method=method, mlp=net if method == B.C2P_LEARNED else None, fallbacks=fb)
        tot, c2p_ms = [], []
        for rep in range(3):
            reset_geometry()
            reset()
            B.synchronize()
            B.timer_reset()
            B.timer_enable(True)
            t0 = time.perf_counter()
            lev.z4c_grhd_rk4_step(z4, P, hp, dt, UZ + UQ, UP, RZ + RQ, V, ADMG, ADMK, PRIM, MADM, status, iters, **kw)
            B.synchronize()
            if rep:                      # the first step is a warm-up
                tot.append(time.perf_counter() - t0)
                c2p_ms.append(B.timer_get("c2p")[0])
            B.timer_enable(False)
        out[name].update(step_ms=1e3 * min(tot), step_c2p_ms=min(c2p_ms), step_updates_per_s=n ** 3 / min(tot),
                         step_status_hist=np.bincount(status.get(), minlength=5).tolist())
        if method == B.C2P_LEARNED:
            out[name]["step_fallbacks"] = int(fb.get()[0])
    out["step_c2p_learned_vs_hybrid_speedup"] = out["hybrid"]["step_c2p_ms"] / out["learned"]["step_c2p_ms"]
    # where the reference's solver converged (not its atmosphere or cold branches)
    ok = res_p["hybrid"][1] == B.C2P_OK
    out["hybrid"]["converged_points"] = int(ok.sum())
    for name in ("bracketed", "learned"):
        out[name]["max_rel_diff_p_vs_hybrid_where_converged"] = float(
            np.max(np.abs(res_p[name][0][ok] / res_p["hybrid"][0][ok] - 1)))
    out["c2p_learned_vs_hybrid_speedup"] = out["hybrid"]["c2p_ms"] / out["learned"]["c2p_ms"]
    for a in (status, iters, fb):
        a.free()
    lev.free()
    return out


def vanal_report(B, P, n, seed, metric_noise):
    """grhd_c2p_vanal on the device for n points of the con2prim
    distribution: the fraction whose (rho, eps, v^i) match the original
    primitives to 1e-10, the fraction where the recovery stops (eps < 0),
    and of the rest the fraction within 64 eps times the conditioning of
    the quartic's root (kappa_v v^2 / (1 - v^2) for rho, times the
    cancellation (E + |S||v| + rho) / |E - |S||v| - rho| for eps)"""
    prim5, gm = c2p_prims(n, seed, metric_noise)
    dp, dg = B.DeviceArray.from_host(prim5), B.DeviceArray.from_host(gm)
    dc = B.DeviceArray((5, n))
    B.p2c_batch(P, dp, dg, dc)
    dpr = B.DeviceArray.from_host(np.zeros((8, n)))
    ds, di = B.DeviceArray((n,), np.int32), B.DeviceArray((n,), np.int32)
    B.synchronize()
    t0 = time.perf_counter()
    B.c2p_batch(P, B.C2P_VANAL, dc, dg, dpr, ds, di)
    B.synchronize()
    ms = 1e3 * (time.perf_counter() - t0)
    cons, pr, st = dc.get(), dpr.get(), ds.get()
    for a in (dp, dg, dc, dpr, ds, di):
        a.free()
    ok = st == B.C2P_OK
    rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    vmag = np.sqrt((prim5[2:5] ** 2).sum(axis=0))
    err_rho, err_eps = rel(pr[0], prim5[0]), rel(pr[1], prim5[1])
    err_v = np.abs(pr[2:5] - prim5[2:5]).max(axis=0) / np.maximum(vmag, 1e-300)
    match = ok & (err_rho <= 1e-10) & (err_eps <= 1e-10) & (err_v <= 1e-10)
    # conditioning from the invariants (undensitised D, E = tau + D, |S|)
    g = gm
    det = g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[2] * g[4]) + g[2] * (g[1] * g[4] - g[2] * g[3])
    sq = np.sqrt(det)
    D, E = cons[0] / sq, (cons[0] + cons[1]) / sq
    W = 1 / np.sqrt(1 - (g[0] * prim5[2] ** 2 + g[3] * prim5[3] ** 2 + g[5] * prim5[4] ** 2 + 2 * (
        g[1] * prim5[2] * prim5[3] + g[2] * prim5[2] * prim5[4] + g[4] * prim5[3] * prim5[4])))
    v = np.sqrt(1 - 1 / W ** 2)
    h = 1 + prim5[1] + (P.gamma - 1) * prim5[1]
    S = prim5[0] * h * W ** 2 * v
    G, gmo = P.gamma, P.gamma - 1
    div = 1 / (gmo ** 2 * (S ** 2 + D ** 2))
    a3, a2 = -2 * G * gmo * S * E * div, (G * G * E * E + 2 * gmo * S ** 2 - gmo ** 2 * D ** 2) * div
    a1, a0 = -2 * G * S * E * div, S ** 2 * div
    q1 = 4 * v ** 3 + 3 * a3 * v ** 2 + 2 * a2 * v + a1
    kv = (v ** 4 + np.abs(a3) * v ** 3 + np.abs(a2) * v ** 2 + np.abs(a1) * v + np.abs(a0)) / np.maximum(
        np.abs(q1) * v, 1e-300) + 1
    eps_m = np.finfo(float).eps
    krho = 1 + kv * v * v / np.maximum(1 - v * v, 1e-300)
    keps = krho * (E + S * v + D / W) / np.maximum(np.abs(E - S * v - D / W), 1e-300)
    rest = ok & ~match
    within = rest & (err_rho <= 64 * eps_m * krho) & (err_eps <= 64 * eps_m * keps)
    return dict(n=n, metric_noise=metric_noise, ms=ms, match_1e_10=float(match.mean()),
                stopped_eps_negative=float((st == B.C2P_FAIL).mean()), other=float(rest.mean()),
                other_within_64_eps_kappa=float(within.sum() / max(rest.sum(), 1)))


def run_c1(args):
    """the single-box configuration: BL m = 1 at the origin, 64^3 interior
    on [-8, 8]^3 (h = 0.25, cell centred) with one radiative boundary layer,
    precollapsed lapse, 1+log / Gamma driver, KO 0.1, dt = h/4, 8 RK4 steps
    on the device, and the same steps with the OpenMP CPU oracle (bounded
    by the CPU sample time)"""
    import baum_amd as B
    h, n = 0.25, 66
    x = -8.0 - 0.5 * h + h * np.arange(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    psi = 1 + 1.0 / (2 * np.sqrt(X ** 2 + Y ** 2 + Z ** 2))
    u = np.zeros((25, n, n, n))
    u[0] = u[3] = u[5] = 1.0
    u[6] = psi ** -4
    u[18] = psi ** -2
    lev = B.Level([B.make_box(n, n, n, x[0], x[0], x[0], h)], 100)
    lev.enable(range(100))
    for v in range(100):
        lev.fill(v, 0.0)
    U, P, R, V = (list(range(25 * k, 25 * k + 25)) for k in range(4))
    par = B.z4c_params()
    dt = 0.25 * h
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_rk4_step(par, dt, U, P, R, V)          # tunes the tile kernels for this box
    lev.put_state(U, u.reshape(25, -1))
    B.synchronize()
    t0 = time.perf_counter()
    for _ in range(8):
        lev.z4c_rk4_step(par, dt, U, P, R, V)
    B.synchronize()
    el = time.perf_counter() - t0
    l2 = lev.z4c_constraints_l2(U)
    lev.free()
    out = dict(value=8 * 64 ** 3 / el, unit="grid-point-updates/s", ms_per_step=1e3 * el / 8,
               ham_mom_l2_after_8_steps=[float("%.6g" % x) for x in l2])
    if not args.no_cpu:
        import oracle as O
        ob = O.make_box(n, n, n, x[0], x[0], x[0], h)
        work = np.zeros((75, n, n, n))
        steps, t0 = 0, time.perf_counter()
        while steps < 8:
            O.z4c_rk4_step(ob, O.z4c_params(), dt, u, bc=1, pgh=0, work=work)
            steps += 1
            if time.perf_counter() - t0 >= args.cpu_seconds:
                break
        el = time.perf_counter() - t0
        out["cpu"] = dict(value=steps * 64 ** 3 / el, unit="grid-point-updates/s",
                          cores=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), kind="port",
                          sample="%d RK4 steps of the 64^3 box with the OpenMP C oracle, %.1f s" % (steps, el))
    return out


def c4_cpu_baseline(target_s):
    """the OpenMP CPU oracle's con2prim (the reference hybrid and the
    bracketed solver) on a bounded sample of the con2prim distribution"""
    import baum_amd as B
    import oracle as O
    n = 1 << 20
    prim5, gm = c2p_prims(n, 4242)
    Po = O.c2p_params(errormax=1e-12, countmax=100, rhoatm=1e-20)
    cons = O.p2c(Po, prim5, gm)
    out = {}
    for name in ("hybrid", "bracketed"):
        reps, t0 = 0, time.perf_counter()
        while True:
            _, st, it, _ = O.c2p(Po, cons, gm, np.zeros(n), rhoguess=prim5[0], method=name)
            reps += 1
            el = time.perf_counter() - t0
            if el >= target_s / 2 or reps >= 20:
                break
        out[name] = dict(pts_per_s=reps * n / el, converged=float(np.mean(st == B.C2P_OK)),
                         mean_iters=float(it.mean()),
                         iter_hist=np.bincount(np.minimum(it, 12), minlength=13).tolist())
    out.update(cores=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), kind="port",
               sample="2^20 points of the con2prim distribution (tolerance 1e-12), cold start, OpenMP C oracle")
    return out


# ------------------------------------------------------------------ CPU baseline
def cpu_baseline(target_s):
    """the OpenMP CPU oracle (test infrastructure, used here only as the
    reported baseline): RK4 steps of one 48^3-interior box of the fmr
    workload's level-3 data (binary puncture, radiative boundary)"""
    import oracle as O
    n = 50
    h = 0.25 / 8
    x0 = -3.0 - (n - 1) / 2 * h
    ob = O.make_box(n, n, n, x0, -(n - 1) / 2 * h, -(n - 1) / 2 * h, h)
    u = puncture_fields(x0 + h * np.arange(n), -(n - 1) / 2 * h + h * np.arange(n), -(n - 1) / 2 * h + h * np.arange(n))
    p = O.z4c_params()
    work = np.zeros((75, n, n, n))
    steps = 0
    t0 = time.perf_counter()
    while True:
        O.z4c_rk4_step(ob, p, 0.25 * h, u, bc=1, pgh=0, work=work)
        steps += 1
        el = time.perf_counter() - t0
        if el >= target_s or steps >= 500:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=steps * (n - 2) ** 3 / el, unit="grid-point-updates/s", cores=cores, kind="port",
                sample="%d RK4 steps of one 48^3-interior box of the FMR workload (binary puncture data, "
                       "radiative boundary) with the OpenMP C oracle, %.1f s; restriction/prolongation "
                       "(7%% of the reference's FMR time) not included" % (steps, el))


def traffic_from_profile():
    """HBM bytes per RHS launch from the committed rocprofv3 PMC summary, if any"""
    path = os.path.join(ROOT, "profiles", "z4c_rhs_bulk_192_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1].split()[0])
    return None


def progress(msg):
    """one line on stderr as each workload starts and ends (stdout carries only the JSON line)"""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU decomposition is a design, not an "
              "implemented path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    only = set(args.only.split(","))
    import baum_amd as B
    B.init()
    B.set_overlap(args.rhs_overlap != 0)
    B.set_rk_fusion(args.rk_fusion)
    B.set_xcd_map(args.xcd_map != 0)
    out = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    progress("fmr hierarchy")
    fmr = run_fmr(args) if "fmr" in only else None
    progress("192^3 box")
    box = run_box(args) if "box192" in only else None
    progress("64^3 puncture box")
    c1 = run_c1(args) if "c1" in only else None
    if fmr:
        out.update(value=fmr["value"], unit="grid-point-updates/s", ms_per_step=fmr["ms_per_step"])
        out["config"] = {"workload": "Z4c BO evolution of a 7-level FMR hierarchy: 128^3 level 0 + 2 boxes of "
                                     "48..96^3 per level 1..6 around a BL binary (m=0.5 at x=+-3), nbuffer 6, "
                                     "order_RP 4, time interpolation 3, z4_chi_div_floor 1e-5, dissipation_factor 0.1 with "
                                     "get_dissipation_factor's rule (amr_move_lcube 0, dissipation_factor_move "
                                     "unset: 0.1 on level 0, 0.02 on levels 1..6), puncture tracking on level 6, "
                                     "one hipGraph per coarse step",
                         "updates_per_coarse_step": fmr["updates_per_step"], "boxes": fmr["boxes"],
                         "parallelism": "single GPU"}
        out["fmr"] = {k: fmr[k] for k in ("nonfinite", "eager_split_ms", "punctures_end")}
        rhs_ms = fmr["eager_split_ms"]["z4_rhs"]
        rhs_flop = FLOP_PER_POINT_RHS * 4 * fmr["updates_per_step"]
        if rhs_ms > 0:
            rate = rhs_flop / (rhs_ms * 1e-3) / 1e12
            out["fmr"]["rhs_fp64"] = {"achieved": rate, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                      "frac": rate / PEAK_FP64_TFLOPS, "rhs_ms_per_coarse_step": rhs_ms,
                                      "flop_per_coarse_step": rhs_flop,
                                      "kernels": "every Z4c RHS launch of one eagerly run coarse step: bulk and "
                                                 "shell tiles on two streams, RK4 update fused into the store"}
    elif box:
        out.update(value=box["value"], unit="grid-point-updates/s", ms_per_step=box["ms_per_step"])
        out["config"] = {"workload": "Z4c RK4, one 192^3 periodic box", "parallelism": "single GPU"}
    if box:
        flops = FLOP_PER_POINT_RHS * box["npts"]
        achieved = flops / (box["rhs_launch_ms"] * 1e-3) / 1e12
        out["roofline"] = {"bound": "fp64", "kernel": "z4c_rhs_tile (bulk, plain RHS)", "achieved": achieved,
                           "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                           "traffic": traffic_from_profile(), "launch_ms": box["rhs_launch_ms"],
                           "algorithmic_flop_per_launch": flops,
                           "hbm_achieved_gbs": BYTES_PER_POINT_RHS * box["npts"] / (box["rhs_launch_ms"] * 1e-3) / 1e9,
                           "bound_note": "FP64 vector ALU (78.6 TFLOP/s, the figure FP64 MFMA shares); the stencils "
                                         "are not contractions, so the kernel runs on the vector units",
                           "workload": "192^3 periodic box (+3 ghosts), one RHS launch"}
        out["box192"] = {"value": box["value"], "unit": "grid-point-updates/s", "ms_per_step": box["ms_per_step"],
                         "breakdown_ms_per_step": box["breakdown_ms_per_step"],
                         "ham_mom_l2_after": box["ham_mom_l2_after"],
                         "ham_l2_every_10_steps_of_100": box["ham_l2_every_10_steps"]}
    if c1:
        out["c1"] = c1
    # tile kernel chosen by the autotuner for each tuned box set, with the time of every variant
    out["rhs_tuning"] = B.tune_report().splitlines()
    if "c2p" in only or "c5" in only:
        progress("con2prim on 2^24 points, network training")
        c2p_out, net = run_c2p(args)
        if "c2p" in only:
            out["c2p"] = c2p_out
            if not args.no_cpu:
                progress("con2prim CPU baseline")
                out["c2p"]["cpu"] = c4_cpu_baseline(args.cpu_seconds)
        if "c5" in only:
            progress("coupled Z4c + GRHD on the TOV star")
            out["c5"] = run_c5(args, net)
        net.free()
    if "move" in only:
        progress("moving boxes")
        try:
            out["move"] = run_move()
        except Exception as e:          # reported, never fatal for the bench line
            out["move"] = {"error": str(e)}
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    progress("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
