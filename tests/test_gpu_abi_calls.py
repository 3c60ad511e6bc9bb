"""Entry points of the C ABI beyond the apply and the solvers, each against
the oracle or an independent restatement: the SCF re-initialization of a
cell-wise operator (dfx_op_update_cells), set_constrained_nodes, axpby /
axpby_blocked, column norms and dots; and the captured degree-loop graphs'
life cycle (captured again when an operator's buffers move, released with
the operator)."""
import numpy as np
import pytest

import dfx
import oracle as O
from gpu_helpers import Ops
from problems import enriched, guess, hanging

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def mv(ctx, a, dtype=dfx.F64):
    return dfx.MultiVector.from_numpy(ctx, a, dtype)


@pytest.fixture(scope="module")
def enr():
    return enriched(p=2, n=4)


@pytest.mark.parametrize("dtype", [dfx.F64, dfx.F32])
def test_update_cells_equals_a_fresh_operator(ctx, enr, dtype):
    """KohnShamOperatorContextFE::reinit (KohnShamOperatorContextFE.t.cpp:
    1247-1310): new cell matrices on an existing operator give the bits of an
    operator created with them."""
    P = enr
    d = P.arrays()
    cons = dfx.Constraints(ctx, P.nDofs, d["rows"], d["rowSizes"], d["rowOffsets"], d["cols"],
                           d["weights"], d["inhom"])
    mk = lambda key: dfx.CellOperator(ctx, d["cellNdofs"], d["cellMap"], d[key], P.nDofs,
                                      P.nClassical, cons, keep_fp32=True)
    A, Mfresh = mk("H"), mk("M")
    npd = np.float64 if dtype == dfx.F64 else np.float32
    X = guess(P, 24, seed=3).astype(npd)
    Y0 = dfx.MultiVector(ctx, P.nDofs, 24, dtype)
    A.apply(mv(ctx, X, dtype), Y0)          # used (and scratch sized) before the update
    A.update_cells(d["M"])
    Y1, Y2 = (dfx.MultiVector(ctx, P.nDofs, 24, dtype) for _ in range(2))
    A.apply(mv(ctx, X, dtype), Y1)
    Mfresh.apply(mv(ctx, X, dtype), Y2)
    assert np.array_equal(Y1.to_numpy(), Y2.to_numpy())
    assert not np.array_equal(Y0.to_numpy(), Y1.to_numpy())


def test_set_constrained_nodes(ctx):
    """ConstraintsInternal::setConstrainedNodes (ConstraintsInternal.cpp:
    196-217): every constrained row set to the value, other rows untouched."""
    P = hanging(p=3)
    ops = Ops(ctx, P)
    d = ops.d
    X = O.rng_uniform(4, -1, 1, P.nDofs * 7).reshape(-1, 7)
    Xg = mv(ctx, X)
    ops.cons.set_constrained_nodes(Xg, 2.5)
    ref = X.copy()
    ref[d["rows"]] = 2.5
    assert np.array_equal(Xg.to_numpy(), ref)
    assert len(d["rows"]) > 0 and d["rowSizes"].max() > 0     # hanging rows too


def test_axpby_and_blocked(ctx):
    """blasLapack::axpby / axpbyBlocked (BlasLapack.t.cpp:228-257): z = a x + b y
    and z_ij = a1 a_j x_ij + b1 b_j y_ij, to one rounding of the reference
    (the GPU may fuse the multiply-add)."""
    N, B = 1000, 9
    x = O.rng_uniform(1, -2, 2, N * B).reshape(N, B)
    y = O.rng_uniform(2, -2, 2, N * B).reshape(N, B)
    X, Y, Z = mv(ctx, x), mv(ctx, y), dfx.MultiVector(ctx, N, B)
    Z.axpby(0.7, X, -1.3, Y)
    ref = 0.7 * x + (-1.3) * y
    bound = 2 * EPS * (np.abs(0.7 * x) + np.abs(1.3 * y))
    assert np.all(np.abs(Z.to_numpy() - ref) <= bound)
    a = np.linspace(0.5, 1.5, B)
    b = np.linspace(-1.0, 2.0, B)
    Z.axpby_blocked(1.5, a, X, -0.25, b, Y)
    ref = 1.5 * a * x + (-0.25) * b * y
    bound = 4 * EPS * (np.abs(1.5 * a * x) + np.abs(0.25 * b * y))
    assert np.all(np.abs(Z.to_numpy() - ref) <= bound)
    # in place (z aliases x), as the filters use it
    X.axpby(2.0, X, 1.0, Y)
    assert np.all(np.abs(X.to_numpy() - (2 * x + y)) <= 2 * EPS * (np.abs(2 * x) + np.abs(y)))


@pytest.mark.parametrize("dtype", [dfx.F64, dfx.F32])
def test_l2norms_and_dots(ctx, dtype):
    """MultiVector::l2Norms and linearAlgebra::dot (MultiVector.h:620), FP64
    accumulation in a fixed order: against numpy to rounding."""
    N, B = 50001, 13
    npd = np.float64 if dtype == dfx.F64 else np.float32
    x = O.rng_uniform(5, -1, 1, N * B).reshape(N, B).astype(npd)
    y = O.rng_uniform(6, -1, 1, N * B).reshape(N, B).astype(npd)
    X, Y = mv(ctx, x, dtype), mv(ctx, y, dtype)
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    dots = X.dots(Y)
    ref = (xd * yd).sum(axis=0)
    assert np.all(np.abs(dots - ref) <= 1e-12 * (np.abs(xd) * np.abs(yd)).sum(axis=0))
    nrm = X.l2norms()
    assert np.allclose(nrm, np.sqrt((xd * xd).sum(axis=0)), rtol=1e-13, atol=0)
    assert np.array_equal(X.l2norms(), nrm)              # deterministic


@pytest.mark.parametrize("prec", [dfx.PREC_FP64, dfx.PREC_MIXED])
def test_graph_is_captured_again_when_operator_buffers_move(enr, prec, monkeypatch):
    """A captured degree loop holds the addresses of the operators' buffers.
    On fresh enriched operators with batches narrower than the subspace
    (launched one by one), the first pass captures the loop at the batch
    width and Rayleigh-Ritz then grows the operator's enriched-partials
    buffer (all states at once): the second pass must capture again (not
    replay into freed memory) and gives the eager bits; destroying the
    operators releases their graphs."""
    monkeypatch.setenv("DFX_FILTER_GROUP_COLS", "0")
    P = enr
    ctx = dfx.Context(0)
    try:
        nv = 12
        X = guess(P, nv, seed=41)
        lo, hi, beta = P.lanczos(O.rng_uniform(2, -1, 1, P.nDofs))
        a, c = lo, hi + beta / 10
        b = a + 0.05 * (c - a)
        outs, stats = [], []
        for use_graph in (True, False):       # the graph run first, on fresh operators
            ops = Ops(ctx, P)
            G = mv(ctx, X)
            V = dfx.MultiVector(ctx, P.nDofs, nv)
            s = dfx.ChebyshevFilteredEigenSolver(ctx, a, b, c, 9, 1e-14, G, True, 4, prec,
                                                 useGraph=use_graph)
            eig = np.linspace(a, b, nv)
            res = []
            for _ in range(3):
                assert s.solve(ops.H, eig, V, True, ops.M, ops.Minv) == 0
                res.append((eig.copy(), V.to_numpy()))
                if use_graph:
                    stats.append(ctx.graph_stats())
                G.copy_from(V)
                s.reinit(eig[0], eig[-1], c, 9, 1e-14, G)
            outs.append(res)
            s.close()
            del ops
            import gc
            gc.collect()
            if use_graph:
                assert ctx.graph_stats()[2] == 0      # released with the operators
        for (e0, v0), (e1, v1) in zip(outs[0], outs[1]):
            assert np.array_equal(e0, e1) and np.array_equal(v0, v1)
        # pass 1: captured once (three batches: one capture, two replays);
        # pass 2: captured again (the enriched buffer grew in pass 1's RR);
        # pass 3: replays only
        (c1, r1, _), (c2, r2, _), (c3, r3, _) = stats
        assert c1 == 1 and r1 == 2
        assert c2 == 2 and r2 == 4
        assert c3 == 2 and r3 == 7
    finally:
        ctx.close()


def test_chfsi_intermediate_subspaces_and_orthogonalization_type(ctx, enr):
    """storeIntermediateSubspaces keeps the filtered subspace and its
    orthonormalized form (ChebyshevFilteredEigenSolver.t.cpp:336-381): the
    latter is CholGS of the former and matches the oracle's orthonormalized
    guess; without the flag the getters fail; only CHOLESKY_GRAMSCHMIDT
    orthogonalization is accepted."""
    P = enr
    ops = Ops(ctx, P)
    nv = 8
    X = guess(P, nv, seed=27)
    lo, hi, beta = P.lanczos(O.rng_uniform(2, -1, 1, P.nDofs))
    a, c = lo, hi + beta / 10
    b = a + 0.05 * (c - a)
    lam = np.linspace(a, b, nv)
    G, V = mv(ctx, X), dfx.MultiVector(ctx, P.nDofs, nv)
    s = dfx.ChebyshevFilteredEigenSolver(ctx, a, b, c, 10, 1e-14, G, True,
                                         storeIntermediateSubspaces=True)
    eig = lam.copy()
    assert s.solve(ops.H, eig, V, True, ops.M, ops.Minv) == 0
    F, Fo = s.getFilteredSubspace(), s.getOrthogonalizedFilteredSubspace()
    Xo = dfx.MultiVector(ctx, P.nDofs, nv)
    assert dfx.cholesky_gram_schmidt(ctx, F, Xo, ops.M) == 0
    assert np.array_equal(Xo.to_numpy(), Fo.to_numpy())
    rco, eigo, Xr, Gr = P.chfsi(a, b, c, 10, X, lam, residual=True)
    assert rco == 0
    assert np.abs(Fo.to_numpy() - Gr).max() < 1e-10 * np.abs(Gr).max()
    s2 = dfx.ChebyshevFilteredEigenSolver(ctx, a, b, c, 10, 1e-14, G, True)
    assert s2.solve(ops.H, lam.copy(), V, True, ops.M, ops.Minv) == 0
    with pytest.raises(dfx.DfxError, match="storeIntermediateSubspaces"):
        s2.getFilteredSubspace()
    for t in (dfx.ORTHO_MULTIPASS_LOWDIN, dfx.ORTHO_MULTIPASS_CGS):
        with pytest.raises(dfx.DfxError) as ex:
            dfx.ChebyshevFilteredEigenSolver(ctx, a, b, c, 10, 1e-14, G, True, orthoType=t)
        assert ex.value.code == -5


def test_eigen_residual_norms_match_oracle(ctx, enr):
    """KohnShamEigenSolver::getLinearEigenSolveResidual (KohnShamEigenSolver.t.cpp:
    574-684): ||H x_j - lambda_j M x_j|| per column against the oracle's."""
    P = enr
    ops = Ops(ctx, P)
    X = guess(P, 6, seed=13)
    lam = np.linspace(-0.5, 0.3, 6)
    res = dfx.eigen_residual_norms(ctx, ops.H, ops.M, mv(ctx, X), lam)
    ref = P.residual_norms(X, lam)
    assert np.abs(res - ref).max() < 1e-12 * np.abs(ref).max()
