#!/usr/bin/env python3
"""Benchmark of the Kohn-Sham eigensolver hot path on one MI355X.

A step is one pass of the Kohn-Sham eigensolver's loop: one ChFSI pass
(Chebyshev filter over all blocks + Cholesky-Gram-Schmidt + Rayleigh-Ritz, the
reference's "Chebyshev Filter" / "OrthoNormalization" / "RR Step" split,
src/linearAlgebra/ChebyshevFilteredEigenSolver.t.cpp:222-397) followed by the
residual norms of every state (KohnShamEigenSolver::getLinearEigenSolveResidual,
src/ksdft/KohnShamEigenSolver.t.cpp:574-684), on the 32-atom enriched workload
(c4: 32^3 cells at p=4, ~2.1M dofs, 200 states, block 64, degree 30), mixed
precision (FP32 filter bulk, FP64 orthonormalization / Rayleigh-Ritz /
residuals).  All inputs live in HBM before the timed region; the problem
setup and the Lanczos bounds are not timed.

Prints one JSON line (see the keys below).  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dft-efe_amd"))

PEAK_FP32_MFMA = 157.3   # TFLOP/s, MI355X dense FP32 matrix
PEAK_FP64_MFMA = 78.6    # TFLOP/s
PEAK_HBM = 8000.0        # GB/s
PROF_EVERY = 4           # cell applies timed by events: every 4th of the kind


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def traffic_from_profiles(kind):
    """HBM bytes per launch of the cell-apply kernel from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "cell_apply_c4_pmc.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals[f"traffic_bytes_per_colour_step_{kind}"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(prob, n_states, block, degree, threads=16, sample_every=8, reps=5):
    """The reference's CPU apply (the oracle's threaded cell-wise apply: gather,
    per-cell product, scatter into a slab-local vector, parallel merge, plus
    the constraint work) on the C4 operator in FP64 at `block` columns:
      leg 1: `threads` threads, the full mesh, median of `reps` after a warm-up;
      leg 2: 1 thread (the reference's per-rank behaviour), every
             `sample_every`-th cell, median of `reps`, the cell part scaled by
             the sample's share of sum n_e^2.
    A pass is (degree + 5) column-applies per state (residual seed H and M,
    degree - 1 filter applies, CholGS M, RR H, residual norms H and M); the
    O(N B) vector updates, Gram matrices and dense factorizations are not
    counted.  Test infrastructure (oracle/) is used here as the CPU baseline
    only."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nd = prob.get("cellNdofs").astype(np.int64)
    cmap = prob.get("cellMap")
    cons = dict(rows=prob.get("rows"), rowSizes=prob.get("rowSizes"),
                rowOffsets=prob.get("rowOffsets"), cols=prob.get("cols"),
                weights=prob.get("weights"), inhom=prob.get("inhom"))
    X = O.rng_uniform(9, -1, 1, prob.nDofs * block).reshape(prob.nDofs, block)
    col_applies = n_states * (degree + 5)
    H = prob.get("H")
    P = O.Problem.from_arrays(nd, cmap, prob.nDofs, prob.nClassical, H=H, p=prob.p, **cons)
    del H
    t_full, t_cells, t_vec, _ = P.time_apply(O.OP_H, X, reps=reps, threads=threads)
    del P
    # 1-thread leg on a sample of cells
    off = np.concatenate([[0], np.cumsum(nd)])
    moff = np.concatenate([[0], np.cumsum(nd * nd)])
    sample = np.arange(0, len(nd), sample_every)
    smap = np.concatenate([cmap[off[e]:off[e + 1]] for e in sample])
    H = prob.get("H")
    sH = np.concatenate([H[moff[e]:moff[e + 1]] for e in sample])
    del H
    Ps = O.Problem.from_arrays(nd[sample], smap, prob.nDofs, prob.nClassical, H=sH, p=prob.p,
                               **cons)
    _, t1_cells, t1_vec, _ = Ps.time_apply(O.OP_H, X, reps=reps, threads=1)
    frac = float((nd[sample] ** 2).sum()) / float((nd ** 2).sum())
    t1_apply = t1_cells / frac + t1_vec
    t_pass = t_full * col_applies / block
    t1_pass = t1_apply * col_applies / block
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                 if l.startswith("model name")][0]
    except (OSError, IndexError):
        model = "unknown"
    return {"value": 1.0 / t_pass, "unit": "passes/s", "cores": threads, "kind": "port",
            "value_1_thread": 1.0 / t1_pass,
            "apply_s": {str(threads) + "_threads_full_mesh": t_full,
                        "1_thread_extrapolated": t1_apply},
            "cpu": model,
            "sample": (f"oracle cell-wise apply of the C4 Hamiltonian, FP64, B={block}: "
                       f"{threads} threads on the full mesh ({len(nd)} cells), median of {reps} "
                       f"= {t_full:.3f} s (cells {t_cells:.3f} s, constraints and zeroing "
                       f"{t_vec:.3f} s); 1 thread on every {sample_every}th cell, median of "
                       f"{reps}, cells {t1_cells:.3f} s / {frac:.4f} + vector work "
                       f"{t1_vec:.3f} s = {t1_apply:.2f} s; a pass = {col_applies / block:.1f} "
                       f"such block applies (filter, CholGS, RR, residual norms); vector "
                       f"updates, Gram matrices and dense factorizations not counted")}


def serving_run(n_requests, rate, slots, seed=5, use_graph=True, saturate=False):
    """The C5 stream (or its first n_requests): Poisson arrivals, continuous
    batching on `slots` streams, FP32 filter / FP64 RR.  Latency = result time
    minus arrival time."""
    import numpy as np
    import dfx
    from dfx import configs
    reqs = configs.serving_stream(seed=seed, n_requests=n_requests, rate=rate)
    if saturate:   # every request waiting at t = 0: the sustained capacity
        reqs = [dict(r, arrival=0.0) for r in reqs]
    # first-use costs (code objects, library handles) outside the measurement
    dfx.serve([dict(reqs[0], arrival=0.0)], max_slots=1)
    t0 = time.perf_counter()
    energies, res = dfx.serve(reqs, max_slots=slots, precision=dfx.PREC_MIXED,
                              use_graph=use_graph)
    wall = time.perf_counter() - t0
    lat = np.array([x["doneSeconds"] - r["arrival"] for x, r in zip(res, reqs)])
    ok = sum(1 for x in res if x["status"] == 0)
    last = max(x["doneSeconds"] for x in res)
    return {"requests": n_requests, "arrival_rate_per_s": None if saturate else rate,
            "arrivals": "all at t = 0 (saturation)" if saturate else "Poisson process",
            "slots": slots,
            "solves_per_s": n_requests / last, "wall_s": round(wall, 3),
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "converged": ok, "mean_passes": float(np.mean([x["passes"] for x in res])),
            "degree_loop": "captured once per request, replayed every pass" if use_graph
                           else "eager launches",
            "stream": f"SplitMix64 seed {seed}: 1-8 atoms (Z in 1,6), 16-28 bohr, "
                      "6^3-14^3 cells at p=3, 5-40 states, 1e-6 residual"}


def c3_run(ctx, max_passes=40):
    """C3: enriched FE ChFSI (4 atoms, 16^3 cells at p = 3, 40 eigenpairs,
    degree 40) solved to a 1e-6 residual once in FP64 and once with the FP32
    filter, from the same start: passes, time per pass, eigenvalue
    agreement.  Both replay their captured degree loops (the operator is
    small enough for launch gaps to matter); each precision is solved once
    untimed first (first-use allocations and the captures)."""
    import numpy as np
    import dfx
    from dfx import configs
    cfg = configs.config("c3")
    prob = dfx.Problem(ctx, p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"],
                       atoms=cfg["atoms"], enrich=cfg["enrich"], orthogonalize=True)
    H, M, K, Minv, Mlump, cons = prob.operators(keep_fp32=True)
    N, n = prob.nDofs, cfg["states"]
    out = {}
    for prec in (dfx.PREC_FP64, dfx.PREC_MIXED):
        for timed in (False, True):
            G = dfx.MultiVector(ctx, N, n)
            G.fill_random(1, -1.0, 1.0)
            cons.distribute_parent_to_child(G)
            L = dfx.MultiVector(ctx, N, 1)
            L.fill_random(2, -1.0, 1.0)
            cons.distribute_parent_to_child(L)
            W = dfx.MultiVector(ctx, N, n)
            ks = dfx.KohnShamEigenSolver(ctx, n, 1e-6, max_passes, G, L,
                                         polynomialDegree=cfg["degree"],
                                         waveFunctionBatchSize=cfg["block"], precision=prec,
                                         numElectrons=cfg["electrons"], useGraph=True)
            e = np.zeros(n)
            ctx.synchronize()
            t0 = time.perf_counter()
            rc = ks.solve(H, e, W, M, Minv, Mlump)
            ctx.synchronize()
            dt = time.perf_counter() - t0
        out[prec] = (rc, e, ks.result, dt)
    # the FP32 cell kernel on C3's operator at the filter's width with M^-1
    # folded in (the degree loop's apply), 20 eager applies profiled per
    # colour step (the solves above replay graphs, which are not profiled)
    info = H.info()
    X = dfx.MultiVector(ctx, N, n, dfx.F32)
    X.fill_random(3, -1.0, 1.0)
    Y = dfx.MultiVector(ctx, N, n, dfx.F32)
    H.apply_after_inverse(Minv, X, Y)
    H.set_profiling(True, dfx.F32)
    H.profile(reset=True)
    for _ in range(20):
        H.apply_after_inverse(Minv, X, Y)
    ctx.synchronize()
    pr = H.profile(reset=True)
    H.set_profiling(False)
    nrows = len(prob.get("rows"))
    fl = 2.0 * n * info["sumNe2"] / info["nColors"]
    by = (4.0 * info["sumNe2"] + 4.0 * (info["sumNe"] + info["nCells"]) + 2.0 * 4 * N * n +
          4.0 * 2 * nrows + 4.0 * nrows * n) / info["nColors"]
    ms = pr["kernelMs"] / max(pr["launches"], 1)
    t_m, t_h = fl / (PEAK_FP32_MFMA * 1e12), by / (PEAK_HBM * 1e9)
    ach = fl / (ms * 1e-3) / 1e12 if ms else 0.0
    c3_roof = {"bound": "mfma" if t_m > t_h else "hbm", "kernel": "cellApplyBufK<float> "
               "(M^-1 folded), one colour step per launch unit",
               "achieved_tflops": ach, "frac_of_fp32_mfma_peak": ach / PEAK_FP32_MFMA,
               "achieved_gbps": by / (ms * 1e-3) / 1e9 if ms else 0.0,
               "frac_of_hbm_peak": (by / (ms * 1e-3) / 1e9 / PEAK_HBM) if ms else 0.0,
               "avg_launch_ms": ms, "ideal_us_mfma_hbm": [t_m * 1e6, t_h * 1e6]}
    occ = int(out[dfx.PREC_FP64][2]["numOccupied"])
    f64, mix = out[dfx.PREC_FP64], out[dfx.PREC_MIXED]
    return {"workload": f"c3: {len(cfg['atoms'])} atoms, 16^3 cells, p=3, {N} dofs, "
                        f"{n} states, degree {cfg['degree']}",
            "fp64_rc": int(f64[0]), "fp64_passes": int(f64[2]["passes"]),
            "fp64_ms_per_pass": f64[3] * 1e3 / max(int(f64[2]["passes"]), 1),
            "mixed_rc": int(mix[0]), "mixed_passes": int(mix[2]["passes"]),
            "mixed_ms_per_pass": mix[3] * 1e3 / max(int(mix[2]["passes"]), 1),
            "occupied_states": occ, "degree_loop": "captured graph, replayed every pass",
            "max_eig_err_mixed_vs_fp64_Ha": float(np.abs(mix[1][:occ] - f64[1][:occ]).max()),
            "roofline": c3_roof}


def poisson_run(ctx):
    """C2: batched Poisson CG on the electrostatics path -- 16^3 cells at p=3
    in a 20-bohr cube, zero Dirichlet, Jacobi preconditioner, relative
    tolerance 1e-10, 8 right-hand sides 4 pi rho_k with rho_k four normalized
    Gaussians (width 0.3 bohr, charges 1-6, centres from seed 100+k) at the
    quadrature points, integrated against the basis by the library's Poisson
    function (PoissonLinearSolverFunctionFE); CGLinearSolver::solve(function)
    is timed.  The same cell kernel applies K."""
    import numpy as np
    import dfx
    from dfx import configs
    cfg = configs.config("c2")
    t0 = time.perf_counter()
    prob = dfx.Problem(ctx, p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"], atoms=[],
                       enrich=False, orthogonalize=False)
    H, M, K, Minv, Mlump, cons = prob.operators(keep_fp32=False)
    N, nr, w = prob.nDofs, cfg["rhs"], cfg["width"]
    xq = prob.get("quadXYZ").reshape(-1, 3)
    f = np.zeros((xq.shape[0], nr))
    for k in range(nr):
        cen, q = configs.poisson_rhs_charges(100 + k)
        for c, qq in zip(cen, q):
            r2 = ((xq - c) ** 2).sum(axis=1)
            f[:, k] += 4 * np.pi * qq * (2 * np.pi * w * w) ** -1.5 * np.exp(-r2 / (2 * w * w))
    fn = dfx.PoissonLinearSolverFunctionFE(prob, K, cons, dfx.MultiVector.from_numpy(ctx, f))
    setup = time.perf_counter() - t0
    b = fn.getRhs()
    cg = dfx.CGLinearSolver(ctx, 10000, 1e-30, cfg["rel_tol"], 1e10)
    ctx.synchronize()
    t0 = time.perf_counter()
    rc = cg.solve(fn)
    ctx.synchronize()
    t = time.perf_counter() - t0
    bn = b.l2norms()
    info = K.info()
    # one CG iteration: about 17 N B FP64 vector accesses (PCG's dots,
    # updates and Jacobi scaling) plus the minimum bytes of one K apply
    # (cell matrices once, the map, X read and Y written once)
    ncons = len(prob.get("rows"))
    bytes_apply = (8.0 * info["sumNe2"] + 4.0 * (info["sumNe"] + info["nCells"]) +
                   2.0 * 8 * N * nr + 4.0 * 2 * ncons + 8.0 * ncons)
    bytes_iter = 17.0 * 8 * N * nr + bytes_apply
    t_iter = t / max(cg.iterations, 1)
    achieved = bytes_iter / t_iter / 1e9
    return {"workload": "c2: 16^3 cells, p=3, %d dofs, %d right-hand sides, Jacobi, rel. tol %g"
                        % (N, nr, cfg["rel_tol"]),
            "rc": int(rc), "iterations": int(cg.iterations), "solve_ms": t * 1e3,
            "ms_per_iteration": t_iter * 1e3,
            "max_relative_residual": float(np.max(cg.residualNorms / bn)),
            "rhs": "integral of 4 pi rho against the basis at the quadrature points",
            "setup_s": round(setup, 3),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM, "unit": "GB/s",
                         "frac": achieved / PEAK_HBM, "algorithmic_bytes_per_iteration": bytes_iter,
                         "note": "17 N B FP64 vector accesses + one K apply's minimum bytes "
                                 "per iteration, over the measured ms per iteration"}}


def batch_widths(n, bs):
    """Column counts of the filter's launches (the library's partition,
    filterLaunchWidths in solve.hip): batches of bs, a remainder of at most
    bs / 4 joining the last full batch, then consecutive batches grouped
    while a group holds at most DFX_FILTER_GROUP_COLS columns (default 128;
    0 keeps the batches).  Each column's result does not depend on it."""
    batches, s0 = [], 0
    while s0 < n:
        nb = min(bs, n - s0)
        if n - s0 > nb and (n - s0 - nb) * 4 <= bs:
            nb = n - s0
        batches.append(nb)
        s0 += nb
    group = int(os.environ.get("DFX_FILTER_GROUP_COLS", "128"))
    out, acc = [], 0
    for w in batches:
        if acc and acc + w > group:
            out.append(acc)
            acc = 0
        acc += w
    if acc:
        out.append(acc)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c4")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="built-in cell-kernel chunk widths (profiling runs: no tuning "
                         "launches in the trace)")
    ap.add_argument("--no-convergence", action="store_true")
    ap.add_argument("--convergence-passes", type=int, default=25)
    ap.add_argument("--serve-requests", type=int, default=256,
                    help="requests of the C5 stream served after the C4 passes (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--no-serve-saturation", dest="serve_saturation", action="store_false",
                    help="skip the run with every request waiting at t = 0")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy 5 times (the reference kernel for "
                         "scaling HBM byte counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build runs the eigensolver hot path on a "
              "single MI355X (multi-GPU sharding is a design, not a measured path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import dfx
    from dfx import configs

    if args.config == "c5":
        r = serving_run(256, args.serve_rate, args.serve_slots)
        log(f"served {r['requests']}: {r['solves_per_s']:.2f} solves/s, p50 "
            f"{r['latency_p50_s']:.3f} s, p99 {r['latency_p99_s']:.3f} s")
        print(json.dumps({"metric": "serving solves/s (C5 stream; p50 / p99 latency)",
                          "value": r["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
                          "higher_is_better": True, "vs_baseline": None,
                          "dtype": "fp32 filter / fp64 ortho+RR (mixed)",
                          "data": "synthetic (seeded request stream)",
                          "config": {"workload": "c5: 256 requests", **r}}), flush=True)
        return

    if args.calibrate_copy:
        # 2^22 rows x 32 FP64 = 1 GiB read + 1 GiB written per copy
        ctx = dfx.Context(0)
        A = dfx.MultiVector(ctx, 1 << 22, 32)
        Bv = dfx.MultiVector(ctx, 1 << 22, 32)
        A.set(1.0)
        for _ in range(5):
            Bv.copy_from(A)
        ctx.synchronize()
        log("copy calibration: 5 copies of 1073741824 bytes (convertK<double, double>)")
        return

    cfg = configs.config(args.config)
    ctx = dfx.Context(0)
    log(f"config {args.config}: p={cfg['p']} cells={cfg['n']} atoms={len(cfg['atoms'])}")
    t0 = time.time()
    prob = dfx.Problem(ctx, p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"],
                       atoms=cfg["atoms"], enrich=cfg["enrich"], orthogonalize=True)
    H, M, K, Minv, Mlump, cons = prob.operators(keep_fp32=True)
    info = H.info()
    setup_s = time.time() - t0
    N, n, B, m = prob.nDofs, cfg["states"], cfg["block"], cfg["degree"]
    log(f"setup {setup_s:.1f}s: N={N} (enriched {prob.nEnriched}) sum n_e^2={info['sumNe2']} "
        f"max n_e={info['maxCellDofs']} colours={info['nColors']}")

    # column-chunk widths of the cell kernel per size class, tuned on the
    # device for every column count the pass uses: the filter batches (FP32
    # H; FP64 H and M for the residual seed) and all states (FP64 M for
    # CholGS, FP64 H for Rayleigh-Ritz)
    t0 = time.time()
    tuned = {}
    widths = sorted(set(batch_widths(n, B)))
    for name, op, dt, cols in (("H", H, dfx.F32, widths), ("H", H, dfx.F64, widths + [n]),
                               ("M", M, dfx.F64, widths + [n])):
        for bw in cols if not args.no_autotune else ():
            tuned[f"{name} {'f32' if dt == dfx.F32 else 'f64'} B={bw}"] = \
                op.autotune(dt, bw, reps=3)
    if tuned:
        log(f"autotune {time.time() - t0:.1f}s: " +
            "; ".join(f"{k}: {v}" for k, v in tuned.items()))

    def start_state():
        G = dfx.MultiVector(ctx, N, n)
        G.fill_random(1, -1.0, 1.0)
        cons.distribute_parent_to_child(G)
        Lg = dfx.MultiVector(ctx, N, 1)
        Lg.fill_random(2, -1.0, 1.0)
        cons.distribute_parent_to_child(Lg)
        return G, Lg

    G, Lg = start_state()
    lo, hi, beta = dfx.lanczos_extreme(ctx, H, Mlump, Minv, Lg, 20)
    r = beta / 10.0
    a = lo
    b = (hi + r - lo) * (n * 200.0 / N) + lo
    if b >= hi + r:
        b = (hi + r + lo) * 0.5
    c = hi + r
    log(f"Lanczos bounds [{lo:.4f}, {hi:.2f}] beta {beta:.3g}; filter [{a:.4f}, {b:.4f}, {c:.2f}]")
    eig = np.zeros(n)
    V = dfx.MultiVector(ctx, N, n)
    warm = dfx.ChebyshevFilteredEigenSolver(ctx, a, b, c, 1, 1e-14, G, False, B)
    ctx.check(warm.solve(H, eig, V, True, M, Minv))
    G.copy_from(V)
    bounds = [eig[0], eig[-1]]

    def make_solver(prec, guess):
        return dfx.ChebyshevFilteredEigenSolver(ctx, bounds[0], bounds[1], c, m, 1e-14, guess,
                                                True, B, prec)

    res_ms = []

    def one_pass(solver, guess, ev, out, bnd):
        rc = solver.solve(H, ev, out, True, M, Minv)
        if rc != 0:
            raise RuntimeError(f"ChFSI pass failed: {dfx.EIG_CODES.get(rc, rc)}")
        # the pass's convergence test: residual norms of every state (FP64),
        # from the H X and M X the solver keeps of its output
        tr = time.perf_counter()
        solver.residual_norms(H, M, out, ev)
        res_ms.append((time.perf_counter() - tr) * 1e3)
        guess.copy_from(out)
        bnd[0], bnd[1] = ev[0], ev[-1]
        solver.reinit(bnd[0], bnd[1], c, m, 1e-14, guess)

    # ---- FP64 reference state (same start) for the per-pass comparison
    G64 = dfx.MultiVector(ctx, N, n)
    G64.copy_from(G)
    eig64, bnd64 = eig.copy(), list(bounds)

    # ---- mixed precision: warm-up, then the timed passes
    mixed = make_solver(dfx.PREC_MIXED, G)
    for _ in range(args.warmup):
        one_pass(mixed, G, eig, V, bounds)
    log(f"warm-up done ({args.warmup} passes); timing {args.steps} passes")
    # every PROF_EVERY-th FP32 apply of the timed passes has its colour steps
    # bracketed by HIP events (a sample: the events cost GPU time of their own)
    H.set_profiling(True, dfx.F32, every=PROF_EVERY)
    H.profile(reset=True)
    res_ms.clear()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_pass(mixed, G, eig, V, bounds)
    ctx.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    prof32 = H.profile(reset=True)
    H.set_profiling(False)
    tim = dict(mixed.timings())
    tim["residual_norms_ms"] = float(np.median(res_ms))
    log(f"mixed pass {t_mixed * 1e3:.1f} ms (filter {tim['filter_ms']:.1f} / ortho "
        f"{tim['ortho_ms']:.1f} / RR {tim['rr_ms']:.1f} / residual norms "
        f"{tim['residual_norms_ms']:.1f} ms)")

    # ---- FP64 passes from the same start (reported beside; not the timed region)
    fp64 = dfx.ChebyshevFilteredEigenSolver(ctx, bnd64[0], bnd64[1], c, m, 1e-14, G64, True,
                                            B, dfx.PREC_FP64)
    V64 = dfx.MultiVector(ctx, N, n)
    H.set_profiling(True, dfx.F64, every=PROF_EVERY)
    H.profile(reset=True)
    n64 = max(2, min(args.steps, 4))
    for _ in range(args.warmup):
        one_pass(fp64, G64, eig64, V64, bnd64)
    H.profile(reset=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(n64):
        one_pass(fp64, G64, eig64, V64, bnd64)
    ctx.synchronize()
    t_fp64 = (time.perf_counter() - t0) / n64
    prof64 = H.profile(reset=True)
    H.set_profiling(False)
    log(f"FP64 pass {t_fp64 * 1e3:.1f} ms; speed-up mixed/FP64 {t_fp64 / t_mixed:.2f}x")

    def tflops(p):
        return p["flops"] / (p["kernelMs"] * 1e-3) / 1e12 if p["kernelMs"] else 0.0

    # ---- mixed vs FP64 eigenvalues at convergence of the occupied states
    conv = None
    if not args.no_convergence:
        out = {}
        for prec in (dfx.PREC_FP64, dfx.PREC_MIXED):
            Gc, Lc = start_state()
            W = dfx.MultiVector(ctx, N, n)
            ks = dfx.KohnShamEigenSolver(ctx, n, 1e-6, args.convergence_passes, Gc, Lc,
                                         polynomialDegree=m, waveFunctionBatchSize=B,
                                         precision=prec, numElectrons=cfg["electrons"])
            e = np.zeros(n)
            t0 = time.time()
            rc = ks.solve(H, e, W, M, Minv, Mlump)
            out[prec] = (rc, e, ks.result, ks.residualNorms, time.time() - t0, W)
            log(f"KS solve {'mixed' if prec else 'FP64'}: rc {rc} passes "
                f"{ks.result['passes']} occupied {ks.result['numOccupied']} "
                f"max occ. residual {ks.residualNorms[:ks.result['numOccupied']].max():.2e} "
                f"({time.time() - t0:.1f}s)")
        occ = out[dfx.PREC_FP64][2]["numOccupied"]
        err = float(np.abs(out[dfx.PREC_MIXED][1][:occ] - out[dfx.PREC_FP64][1][:occ]).max())
        conv = {"occupied_states": int(occ), "max_eig_err_mixed_vs_fp64_Ha": err,
                "fp64_rc": int(out[0][0]), "fp64_passes": int(out[0][2]["passes"]),
                "mixed_rc": int(out[1][0]), "mixed_passes": int(out[1][2]["passes"]),
                "residual_tol": 1e-6,
                "fp64_solve_s": round(out[0][4], 2), "mixed_solve_s": round(out[1][4], 2)}
        # the electron density of the occupied states (occupation 1 each)
        # from both converged subspaces: rho = sum_i 2 f_i psi_i^2
        try:
            fo = np.zeros(n)
            fo[:occ] = 1.0
            r64 = prob.density(fo, out[0][5]).to_numpy().ravel()
            rmx = prob.density(fo, out[1][5]).to_numpy().ravel()
            wq = prob.get("quadW")
            conv["density_err_max_rel"] = float(np.abs(rmx - r64).max() / np.abs(r64).max())
            conv["density_err_integrated_electrons"] = float(np.sum(wq * np.abs(rmx - r64)))
            conv["electrons_from_fp64_density"] = float(np.sum(wq * r64))
            log(f"density mixed vs FP64: max rel {conv['density_err_max_rel']:.2e}, "
                f"integrated |drho| {conv['density_err_integrated_electrons']:.2e} e "
                f"(of {conv['electrons_from_fp64_density']:.6f})")
            del r64, rmx, wq
        except dfx.DfxError as ex:
            log(f"density comparison skipped: {ex}")
        del out

    # electron density of the 200-state subspace at every quadrature point
    # (next to the eigensolve: DensityCalculator::computeRho), first 70
    # states occupied
    occ = np.zeros(n)
    occ[:70] = 1.0
    dens = None
    try:
        ctx.synchronize()
        t0 = time.perf_counter()
        rho = prob.density(occ, V)
        ctx.synchronize()
        t_rho = time.perf_counter() - t0
        nquad = int(prob.get("quadOffsets")[-1])
        dens = {"states": n, "quadrature_points": nquad, "ms": t_rho * 1e3,
                "interpolation_gflop": 2.0 * n * float(np.sum(
                    np.diff(prob.get("quadOffsets")).astype(np.float64) *
                    prob.get("cellNdofs"))) / 1e9}
        log(f"density: {nquad} quadrature points, {n} states, {t_rho * 1e3:.1f} ms")
        del rho
    except dfx.DfxError as e:
        log(f"density skipped: {e}")

    c3 = None
    if not args.no_convergence:
        log("C3 enriched ChFSI, FP64 and mixed ...")
        c3 = c3_run(ctx)
        log(f"C3: FP64 {c3['fp64_passes']} passes {c3['fp64_ms_per_pass']:.1f} ms/pass, mixed "
            f"{c3['mixed_passes']} passes {c3['mixed_ms_per_pass']:.1f} ms/pass, max occupied "
            f"eigenvalue difference {c3['max_eig_err_mixed_vs_fp64_Ha']:.2e} Ha")

    log("Poisson CG (c2) ...")
    poisson = poisson_run(ctx)
    log(f"Poisson: rc {poisson['rc']}, {poisson['iterations']} iterations, "
        f"{poisson['solve_ms']:.1f} ms ({poisson['ms_per_iteration']:.3f} ms / iteration)")

    serving = None
    if args.serve_requests:
        log(f"serving {args.serve_requests} requests of the C5 stream ...")
        serving = serving_run(args.serve_requests, args.serve_rate, args.serve_slots)
        log(f"served: {serving['solves_per_s']:.2f} solves/s, p50 "
            f"{serving['latency_p50_s']:.3f} s, p99 {serving['latency_p99_s']:.3f} s")
        if args.serve_saturation:
            log(f"saturation: {args.serve_requests} requests waiting at t = 0 ...")
            sat = serving_run(args.serve_requests, args.serve_rate, args.serve_slots,
                              saturate=True)
            serving = {"stream": serving, "saturation": sat}
            log(f"saturation: {sat['solves_per_s']:.2f} solves/s sustained, p50 "
                f"{sat['latency_p50_s']:.3f} s, p99 {sat['latency_p99_s']:.3f} s")

    base = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, sampled cells) ...")
        base = cpu_baseline(prob, n, B, m)
        log(f"CPU baseline {base['value']:.4g} passes/s on {base['cores']} threads")

    # ---- roofline of the dominant kernel: the FP32 cell apply of the filter's
    # degree loop with the residual step fused in.  A launch unit is one
    # colour step (its size-class launches side by side), timed with HIP
    # events on the library's stream over a sample of the timed passes'
    # applies.  Algorithmic flops 2 B sum n_e^2 per apply; algorithmic bytes
    # the minimum per apply: cell matrices once (s sum n_e^2), the map
    # (4 (sum n_e + E)), X read once, and per column of a row finished in the
    # kernel res and r0 read and res written, of the other rows (constrained
    # and enriched: the separate pass finishes them) y written, plus the
    # constraint rows; both split evenly over the colours.
    nrows = len(prob.get("rows"))
    cons_bytes = 4.0 * (2 * nrows + prob.nEntries) + 8.0 * prob.nEntries
    n_fin = N - nrows - prob.nEnriched

    def apply_bytes(bw):
        return (4.0 * info["sumNe2"] + 4.0 * (info["sumNe"] + info["nCells"]) +
                4.0 * N * bw + 3 * 4.0 * n_fin * bw + 4.0 * (N - n_fin) * bw + cons_bytes +
                4.0 * nrows * bw)

    apply_widths = [bw for bw in batch_widths(n, B) for _ in range(m - 1)] * args.steps
    nc = info["nColors"]
    sampled = apply_widths[::PROF_EVERY]
    launches = prof32["launches"]
    flops_sampled = sum(2.0 * bw * info["sumNe2"] for bw in sampled)
    if launches != len(sampled) * nc or abs(prof32["flops"] - flops_sampled) > 1e-9 * flops_sampled:
        log(f"warning: profiled launches {launches} / flops {prof32['flops']:.4g} differ from "
            f"the expected {len(sampled) * nc} / {flops_sampled:.4g}")
    per_launch_flops = flops_sampled / max(len(sampled) * nc, 1)
    per_launch_bytes = sum(apply_bytes(bw) for bw in sampled) / max(len(sampled) * nc, 1)
    per_launch_ms = prof32["kernelMs"] / max(launches, 1)
    # the roofline that bounds the kernel is the one with the longer ideal time
    t_mfma = per_launch_flops / (PEAK_FP32_MFMA * 1e12)
    t_hbm = per_launch_bytes / (PEAK_HBM * 1e9)
    if t_hbm >= t_mfma:
        bound, peak, unit = "hbm", PEAK_HBM, "GB/s"
        achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms else 0.0
    else:
        bound, peak, unit = "mfma", PEAK_FP32_MFMA, "TFLOP/s"
        achieved = per_launch_flops / (per_launch_ms * 1e-3) / 1e12 if per_launch_ms else 0.0
    traffic = traffic_from_profiles("fp32")
    result = {
        "metric": "ChFSI pass time + cell-apply TFLOP/s, mixed vs FP64 @1e-6 Ha; p99 serve",
        "value": 1.0 / t_mixed,
        "unit": "passes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 filter / fp64 ortho+RR (mixed)",
        "data": "synthetic (seeded atoms and initial subspace)",
        "config": {"workload": (f"{args.config}: {len(cfg['atoms'])} atoms, "
                                f"{cfg['n'][0]}^3 cells, p={cfg['p']}, {N} dofs, {n} states, "
                                f"block {B}, degree {m}"),
                   "states": n, "block": B, "filter_launch_columns": batch_widths(n, B),
                   "degree": m, "dofs": N,
                   "enriched_dofs": prob.nEnriched, "sum_ne2": info["sumNe2"],
                   "max_cell_dofs": info["maxCellDofs"], "parallelism": "1 GPU",
                   "cell_chunk_tiles_autotuned": {k: {str(c): w for c, w in v.items()}
                                                  for k, v in tuned.items()}},
        "fp64_pass_ms": t_fp64 * 1e3,
        "speedup_mixed_vs_fp64": t_fp64 / t_mixed,
        "pass_split_ms": tim,
        "density": dens,
        "c3": c3,
        "cell_apply_tflops_fp32": tflops(prof32),
        "cell_apply_tflops_fp64": tflops(prof64),
        "cell_apply_fp64_frac_of_peak": tflops(prof64) / PEAK_FP64_MFMA,
        "convergence": conv,
        "setup_s": setup_s,
        "serving": serving,
        "poisson": poisson,
        "roofline": {"bound": bound,
                     "kernel": ("cellApplyBufK<float> (fused gather / FP32 MFMA / scatter "
                                "with the residual Chebyshev step); one launch unit = "
                                "one colour step, its size-class launches running side "
                                "by side"),
                     "achieved": achieved, "peak": peak, "unit": unit,
                     "frac": achieved / peak, "traffic": traffic,
                     "traffic_over_algorithmic": (traffic / per_launch_bytes
                                                  if traffic else None),
                     "ideal_us_mfma_hbm": [t_mfma * 1e6, t_hbm * 1e6],
                     "launches": prof32["launches"],
                     "launches_sampled": f"the colour steps of every {PROF_EVERY}th FP32 "
                                         f"apply of the {args.steps} timed passes",
                     "avg_launch_ms": per_launch_ms,
                     "algorithmic_flops_per_launch": per_launch_flops,
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "achieved_gbps": per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
                                      if per_launch_ms else 0.0,
                     "achieved_tflops": (per_launch_flops / (per_launch_ms * 1e-3) / 1e12
                                         if per_launch_ms else 0.0),
                     "frac_of_fp32_mfma_peak": (per_launch_flops / (per_launch_ms * 1e-3)
                                                / 1e12 / PEAK_FP32_MFMA
                                                if per_launch_ms else 0.0)},
        "cpu_baseline": base,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
