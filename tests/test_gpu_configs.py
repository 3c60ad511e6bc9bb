"""Parity at the benchmark configurations' own sizes and kernel shapes.

C1 (8^3 cells, p = 3, 10 states, block 10, degree 50) at full size: GPU
builder maps bit-exact with the oracle's, one ChFSI pass with each filter and
a full Kohn-Sham eigensolve against the oracle (eigenvalues 1e-10 Ha,
subspace angles 1e-8).  C4's cell-kernel shapes (p = 4 cells carrying 5-45
enrichment functions: 130-170 dofs, the three-M-tile class) in FP32 at 64
and 72 columns, M^-1 folded into the gather included, and FP64 at 200
columns, against the oracle's applies.  C3 at full size in FP64 and mixed
precision (eigenvalues within 1e-6 Ha, residuals within tolerance).  A few
requests of the C5 stream served alone and in the mix."""
import numpy as np
import pytest
import scipy.linalg as sl

import dfx
import oracle as O
from dfx import configs
from gpu_helpers import Ops
from problems import guess

pytestmark = pytest.mark.gpu

TOL_EIG64 = 1e-10
TOL_ANGLE64 = 1e-8
TOL_APPLY64 = 1e-13
TOL_APPLY32 = 2e-6


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def mv(ctx, a, dtype=dfx.F64):
    return dfx.MultiVector.from_numpy(ctx, a, dtype)


@pytest.fixture(scope="module")
def c1():
    cfg = configs.config("c1")
    P = O.Problem.build(p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"], atoms=cfg["atoms"])
    return cfg, P


def c1_bounds(P, n):
    """KohnShamEigenSolver's filter bounds (KohnShamEigenSolver.t.cpp:
    264-290) from the oracle's Lanczos run."""
    lo, hi, beta = P.lanczos(O.rng_uniform(2, -1, 1, P.nDofs))
    r = beta / 10.0
    a, c = lo, hi + r
    b = (hi + r - lo) * (n * 200.0 / P.nDofs) + lo
    if b >= hi + r:
        b = (hi + r + lo) * 0.5
    return a, b, c


def test_c1_builder_matches_oracle(ctx, c1):
    cfg, P = c1
    G = dfx.Problem(ctx, p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"],
                    atoms=cfg["atoms"], enrich=False, orthogonalize=False)
    d = P.arrays()
    assert (G.nCells, G.nDofs) == (512, 15625) and len(d["rows"]) == 25 ** 3 - 23 ** 3
    for key in ("cellNdofs", "cellMap", "rows", "rowSizes"):
        assert np.array_equal(G.get(key), d[key]), key
    for key in ("H", "M", "K"):
        assert rel(G.get(key), d[key]) < 1e-12, key


@pytest.mark.parametrize("residual", [True, False])
def test_c1_chfsi_pass_matches_oracle(ctx, c1, residual):
    cfg, P = c1
    ops = Ops(ctx, P)
    n, m = cfg["states"], cfg["degree"]
    X = guess(P, n, seed=1)
    a, b, c = c1_bounds(P, n)
    lam = np.linspace(a, b, n)
    G, V = mv(ctx, X), dfx.MultiVector(ctx, P.nDofs, n)
    s = dfx.ChebyshevFilteredEigenSolver(ctx, a, b, c, m, 1e-14, G, residual, cfg["block"])
    eig = lam.copy()
    assert s.solve(ops.H, eig, V, True, ops.M, ops.Minv) == 0
    rco, eigo, Xo, _ = P.chfsi(a, b, c, m, X, lam, residual=residual, batch=cfg["block"])
    assert rco == 0
    assert np.abs(eig - eigo).max() < TOL_EIG64
    assert sl.subspace_angles(V.to_numpy(), Xo).max() < TOL_ANGLE64


def test_c1_ks_eigensolve_matches_oracle(ctx, c1):
    """The full C1 solve: Lanczos bounds, one degree-1 plain pass, residual
    filter passes at degree 50 until every occupied state's residual is
    within 1e-9."""
    cfg, P = c1
    ops = Ops(ctx, P)
    n, tol = cfg["states"], 1e-9
    X0 = guess(P, n, seed=1)
    lg = P.p2c(O.rng_uniform(77, -1, 1, P.nDofs).reshape(-1, 1))
    ref = P.ks_eigensolve(X0, lg[:, 0], degree=cfg["degree"], tol=tol, max_pass=30,
                          batch=cfg["block"], num_electrons=cfg["electrons"])
    W = dfx.MultiVector(ctx, P.nDofs, n)
    ks = dfx.KohnShamEigenSolver(ctx, n, tol, 30, mv(ctx, X0), mv(ctx, lg),
                                 polynomialDegree=cfg["degree"],
                                 waveFunctionBatchSize=cfg["block"],
                                 numElectrons=cfg["electrons"])
    e = np.zeros(n)
    rc = ks.solve(ops.H, e, W, ops.M, ops.Minv, ops.Mlump)
    assert rc == 0 and ref["rc"] == 0
    assert ks.result["passes"] == ref["passes"]
    assert ks.result["numOccupied"] == ref["nocc"]
    # the states converged to 1e-6 up to a spectral gap: 1s, 2p (threefold)
    # and 2s (the next pair of states is split by only 1e-10 Ha)
    k = 5
    assert np.all(ref["res"][:k] <= 1e-6)
    assert ref["eig"][k] - ref["eig"][k - 1] > 1e-2
    assert np.abs(e[:k] - ref["eig"][:k]).max() < TOL_EIG64
    assert sl.subspace_angles(W.to_numpy()[:, :k], ref["X"][:, :k]).max() < TOL_ANGLE64


# ------------------------------------------------------ C4 kernel shapes
def abs_apply(P, key, Z):
    """|A| |Z| assembled with the problem's maps and constraints: the scale
    of the rounding error of A Z."""
    d = P.arrays()
    Pa = O.Problem.from_arrays(d["cellNdofs"], d["cellMap"], P.nDofs, P.nClassical,
                               H=np.abs(d[key]), rows=d["rows"], rowSizes=d["rowSizes"],
                               rowOffsets=d["rowOffsets"], cols=d["cols"],
                               weights=np.abs(d["weights"]), inhom=d["inhom"], p=P.p)
    return Pa.apply(O.OP_H, np.abs(Z))


@pytest.fixture(scope="module")
def c4cells():
    """p = 4 (125 classical dofs per cell) with six overlapping enriched atoms:
    cells of 130-170 dofs, as in C4 (the three-M-tile class, 129-192)."""
    atoms = [(-1.0, 0.5, 0.2, 1), (1.2, -0.8, 0.6, 6), (0.3, 1.4, -1.1, 8),
             (-0.6, -1.3, 0.9, 1), (1.6, 1.1, 1.3, 6), (-1.7, 0.2, -1.4, 8)]
    P = O.Problem.build(p=4, n=(4, 4, 4), lo=(-12.,) * 3, hi=(12.,) * 3, atoms=atoms,
                        enrich=True, nq_enr=8)
    nd = P.arrays()["cellNdofs"]
    assert nd.min() == 125 and 129 <= nd.max() <= 192
    assert (nd > 128).sum() >= 16
    return P


@pytest.mark.parametrize("dtype", [dfx.F32, dfx.F64])
@pytest.mark.parametrize("B", [8, 72, 128, 200])
def test_split_column_chunks_give_the_same_bits(ctx, c4cells, dtype, B, monkeypatch):
    """A cell's column chunks on workgroups of their own (one XCD) or looped
    in one workgroup (DFX_CELL_SPLIT=0): the same bits for H, for M^-1
    folded into H's gather, and for the enriched rows, at one chunk, a chunk
    and a narrow tail (72 = 64 + 8 in FP64) and several chunks (128, 200)."""
    P = c4cells
    ops = Ops(ctx, P)
    npd = np.float32 if dtype == dfx.F32 else np.float64
    X = guess(P, B, seed=B + 1).astype(npd)
    outs = []
    for split in ("1", "0"):
        monkeypatch.setenv("DFX_CELL_SPLIT", split)
        Y = dfx.MultiVector(ctx, P.nDofs, B, dtype)
        ops.H.apply(mv(ctx, X, dtype), Y)
        Yf = dfx.MultiVector(ctx, P.nDofs, B, dtype)
        ops.H.apply_after_inverse(ops.Minv, mv(ctx, X, dtype), Yf)
        outs.append((Y.to_numpy(), Yf.to_numpy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("B", [64, 72])
def test_c4_fp32_cell_shapes_match_oracle(ctx, c4cells, B):
    P = c4cells
    ops = Ops(ctx, P)
    X = guess(P, B, seed=B)
    X32 = X.astype(np.float32)
    Y = dfx.MultiVector(ctx, P.nDofs, B, dfx.F32)
    ops.H.apply(mv(ctx, X32, dfx.F32), Y)
    ref = P.apply(O.OP_H, X32.astype(np.float64))
    assert rel(Y.to_numpy().astype(np.float64), ref) < TOL_APPLY32
    # M^-1 folded into the gather (shared homogeneous constraints) against the
    # oracle's M^-1 apply followed by its H apply
    Yf = dfx.MultiVector(ctx, P.nDofs, B, dfx.F32)
    ops.H.apply_after_inverse(ops.Minv, mv(ctx, X32, dfx.F32), Yf)
    Z = P.apply(O.OP_MINV, X32.astype(np.float64))
    ref = P.apply(O.OP_H, Z)
    # FP32 error model of the composite, element by element: the products
    # d x rounded to FP32, H_e rounded to FP32 and the n_e-term FP32
    # accumulation, then at most 8 cells added per row (an FP32 rounding
    # each): |err| <= (n_max + 10) u32 (|H| |Z|)
    err = np.abs(Yf.to_numpy().astype(np.float64) - ref)
    bound = (P.arrays()["cellNdofs"].max() + 10) * 2.0 ** -24 * abs_apply(P, "H", Z)
    assert np.all(err <= bound)
    assert rel(Yf.to_numpy().astype(np.float64), ref) < 1e-5
    # the folded path gives the bits of M^-1 then H launched apart (M^-1 on a
    # second, identical constraints object is not folded)
    d = ops.d
    cons2 = dfx.Constraints(ctx, P.nDofs, d["rows"], d["rowSizes"], d["rowOffsets"],
                            d["cols"], d["weights"], d["inhom"])
    Minv2 = dfx.DiagonalOperator(ctx, d["dinv"], P.nEnriched, d["enrMinv"], cons2)
    Yu = dfx.MultiVector(ctx, P.nDofs, B, dfx.F32)
    ops.H.apply_after_inverse(Minv2, mv(ctx, X32, dfx.F32), Yu)
    assert np.array_equal(Yf.to_numpy(), Yu.to_numpy())


@pytest.mark.parametrize("prec", [dfx.PREC_FP64, dfx.PREC_MIXED])
def test_c4_residual_step_fused_is_bit_identical(ctx, c4cells, prec, monkeypatch):
    """One ChFSI pass of 200 states in C4's batches of 64, launched as C4's
    are (128 + 72 columns, degree 30): the residual step fused into the cell
    kernels of C4's classes gives the bits of the separate step
    (DFX_FUSE_RES=0)."""
    P = c4cells
    ops = Ops(ctx, P)
    nv = 200
    X = guess(P, nv, seed=136)
    lo, hi, beta = P.lanczos(O.rng_uniform(2, -1, 1, P.nDofs))
    a, c = lo, hi + beta / 10
    b = a + 0.1 * (c - a)
    outs = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("DFX_FUSE_RES", fuse)
        G = mv(ctx, X)
        V = dfx.MultiVector(ctx, P.nDofs, nv)
        s = dfx.ChebyshevFilteredEigenSolver(ctx, a, b, c, 30, 1e-14, G, True, 64, prec)
        eig = np.linspace(a, b, nv)
        assert s.solve(ops.H, eig, V, True, ops.M, ops.Minv) == 0
        outs.append((eig.copy(), V.to_numpy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_c4_fp64_200_columns_match_oracle(ctx, c4cells):
    """The FP64 applies of C4's CholGS and Rayleigh-Ritz (all 200 states)."""
    P = c4cells
    ops = Ops(ctx, P)
    X = guess(P, 200, seed=200)
    for kind, op in ((O.OP_H, ops.H), (O.OP_M, ops.M)):
        Xg, Y = mv(ctx, X), dfx.MultiVector(ctx, P.nDofs, 200)
        op.apply(Xg, Y)
        assert rel(Y.to_numpy(), P.apply(kind, X.copy())) < TOL_APPLY64
    Y = dfx.MultiVector(ctx, P.nDofs, 200)
    ops.H.apply_after_inverse(ops.Minv, mv(ctx, X), Y)
    ref = P.apply(O.OP_H, P.apply(O.OP_MINV, X.copy()))
    assert rel(Y.to_numpy(), ref) < TOL_APPLY64


# ------------------------------------------------------------------- C3
def test_c3_mixed_matches_fp64(ctx):
    cfg = configs.config("c3")
    G = dfx.Problem(ctx, p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"],
                    atoms=cfg["atoms"], enrich=cfg["enrich"], orthogonalize=True)
    H, M, K, Minv, Mlump, cons = G.operators(keep_fp32=True)
    N, n, tol = G.nDofs, cfg["states"], 1e-6
    out = {}
    for prec in (dfx.PREC_FP64, dfx.PREC_MIXED):
        Gs = dfx.MultiVector(ctx, N, n)
        Gs.fill_random(1, -1.0, 1.0)
        cons.distribute_parent_to_child(Gs)
        L = dfx.MultiVector(ctx, N, 1)
        L.fill_random(2, -1.0, 1.0)
        cons.distribute_parent_to_child(L)
        W = dfx.MultiVector(ctx, N, n)
        ks = dfx.KohnShamEigenSolver(ctx, n, tol, 40, Gs, L, polynomialDegree=cfg["degree"],
                                     waveFunctionBatchSize=cfg["block"], precision=prec,
                                     numElectrons=cfg["electrons"])
        e = np.zeros(n)
        rc = ks.solve(H, e, W, M, Minv, Mlump)
        assert rc == 0
        occ = ks.result["numOccupied"]
        assert np.all(ks.residualNorms[:occ] <= tol)
        out[prec] = (e, occ)
    occ = out[dfx.PREC_FP64][1]
    assert out[dfx.PREC_MIXED][1] == occ
    err = np.abs(out[dfx.PREC_MIXED][0][:occ] - out[dfx.PREC_FP64][0][:occ]).max()
    print(f"C3: {N} dofs, {occ} occupied, mixed vs FP64 {err:.2e} Ha", flush=True)
    assert err <= 1e-6


# ------------------------------------------------------------------- C5
def test_c5_requests_alone_and_in_the_mix():
    """The first four requests of the C5 stream (seed 5), served one at a
    time and together on four slots: the same passes and eigenvalues, bit for
    bit."""
    reqs = configs.serving_stream(seed=5, n_requests=4, rate=100.0)
    alone = [dfx.serve([dict(r, arrival=0.0)], max_slots=1, precision=dfx.PREC_FP64)
             for r in reqs]
    mix_e, mix_r = dfx.serve(reqs, max_slots=4, precision=dfx.PREC_MIXED)
    alone_mixed = [dfx.serve([dict(r, arrival=0.0)], max_slots=1, precision=dfx.PREC_MIXED)
                   for r in reqs]
    for (e1, r1), e2, x2 in zip(alone_mixed, mix_e, mix_r):
        assert r1[0]["status"] == 0 and x2["status"] == 0
        assert r1[0]["passes"] == x2["passes"]
        assert np.array_equal(e1[0], e2)
    # the occupied states (at least Ne / 2) agree between FP64 and mixed
    for r, (e64, r64), (emx, rmx) in zip(reqs, alone, alone_mixed):
        assert r64[0]["status"] == 0
        k = int(np.ceil(r["electrons"] / 2))
        assert np.abs(e64[0][:k] - emx[0][:k]).max() < 1e-6
