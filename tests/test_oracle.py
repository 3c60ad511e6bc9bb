"""Pins of the CPU oracle (test infrastructure) against the reference's own
known answers and against what the mathematics fixes: closed forms, dense
brute force on tiny meshes, invariants.  CPU only."""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
from problems import c1_like, enriched, guess, hanging

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------- dense helpers
def dense_cell_matrix(P, which):
    """Global assembled matrix sum_e Q_e^T A_e Q_e (no constraints)."""
    d = P.arrays()
    N = P.nDofs
    A = np.zeros((N, N))
    mats = d[which]
    off = 0
    moff = 0
    for n in d["cellNdofs"]:
        ids = d["cellMap"][off:off + n]
        Ae = mats[moff:moff + n * n].reshape(n, n, order="F")
        A[np.ix_(ids, ids)] += Ae
        off += n
        moff += n * n
    return A


def constraint_matrix(P):
    """C with x_full = C x: identity on free rows, weights on constrained rows."""
    d = P.arrays()
    N = P.nDofs
    Cm = np.eye(N)
    for i, r in enumerate(d["rows"]):
        Cm[r, :] = 0.0
        o = d["rowOffsets"][i]
        for j in range(d["rowSizes"][i]):
            Cm[r, d["cols"][o + j]] = d["weights"][o + j]
    return Cm


# ---------------------------------------------------------------- golden
def test_gemm_known_answer_from_reference_test():
    g = json.load(open(os.path.join(GOLD, "gemm_colmajor_known_answer.json")))
    A = np.array(g["A_colmajor"]); B = np.array(g["B_colmajor"])
    Cm = np.zeros(g["m"] * g["n"])
    O.gemm("N", "N", g["m"], g["n"], g["k"], 1.0, A, g["lda"], B, g["ldb"], 0.0, Cm, g["ldc"])
    assert np.abs(Cm - np.array(g["C_colmajor"])).max() < g["tol"]


def test_gemm_transposes_against_numpy():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((7, 4)); B = rng.standard_normal((7, 5))
    Cm = np.zeros(4 * 5)
    O.gemm("T", "N", 4, 5, 7, 1.0, np.asfortranarray(A).ravel(order="F"), 7,
           np.asfortranarray(B).ravel(order="F"), 7, 0.0, Cm, 4)
    assert np.allclose(Cm.reshape(5, 4).T, A.T @ B, atol=1e-13)


# ------------------------------------------------------------ quadrature
def test_gll_closed_forms():
    x, w = O.gll(4)
    assert np.allclose(x, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], atol=1e-15)
    assert np.allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], atol=1e-15)
    x, w = O.gll(5)
    assert np.allclose(x, [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], atol=1e-15)
    assert np.allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], atol=1e-15)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8])
def test_quadrature_exactness(n):
    xg, wg = O.gauss(n)
    xl, wl = O.gll(n)
    for k in range(2 * n):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.sum(wg * xg ** k) - exact) < 1e-13
        if k <= 2 * n - 3:
            assert abs(np.sum(wl * xl ** k) - exact) < 1e-13


# ----------------------------------------------------------------- maps
def test_config1_sizes_and_dirichlet_rows():
    P = c1_like(n=8, p=3)
    assert P.nCells == 512 and P.nDofs == 25 ** 3
    d = P.arrays()
    assert P.nRows == 25 ** 3 - 23 ** 3 == 3458
    assert np.all(d["rowSizes"] == 0) and np.all(np.diff(d["rows"].astype(np.int64)) > 0)
    # every dof of the lattice is used; interior vertex nodes sit in 8 cells
    counts = np.bincount(d["cellMap"], minlength=P.nDofs)
    assert counts.min() >= 1 and counts.max() == 8
    xyz = P.node_xyz()
    on_bnd = np.any(np.isclose(np.abs(xyz), 10.0), axis=1)
    assert set(np.where(on_bnd)[0]) == set(d["rows"].tolist())


def test_cell_local_nodes_are_gll_tensor_points():
    P = c1_like(n=3, p=3)
    d = P.arrays(); xyz = P.node_xyz()
    lo, h = P.cell_boxes()
    g, _ = O.gll(4)
    e = 5
    ids = d["cellMap"][e * 64:(e + 1) * 64]
    for c in range(4):
        for b in range(4):
            for a in range(4):
                l = a + 4 * (b + 4 * c)
                want = lo[e] + h[e] * (np.array([g[a], g[b], g[c]]) + 1) / 2
                assert np.allclose(xyz[ids[l]], want, atol=1e-13)


@pytest.mark.parametrize("p", [2, 3, 4])
def test_hanging_constraints_reproduce_polynomials(p):
    P = hanging(p=p)
    xyz = P.node_xyz(); hang = P.is_hanging()
    assert hang.sum() > 0
    f = lambda X: (X[:, 0] ** p + 0.3 * X[:, 1] ** (p - 1) * X[:, 2]
                   - 2 * X[:, 0] * X[:, 1] * X[:, 2] + 1.0)
    x = f(xyz).reshape(-1, 1).copy()
    x[hang] = -7.0
    P.p2c(x)
    scale = np.abs(f(xyz)).max()
    assert np.abs(x[hang, 0] - f(xyz)[hang]).max() <= 1e-13 * scale
    d = P.arrays()
    assert set(np.unique(d["rowSizes"])) == {0, p + 1, (p + 1) ** 2}


# -------------------------------------------------------- cell matrices
def _lagrange_1d(nodes, x):
    """L_a(x) for the Lagrange polynomials on `nodes`, shape (len(nodes), len(x))."""
    L = np.ones((len(nodes), len(x)))
    for a, xa in enumerate(nodes):
        for b, xb in enumerate(nodes):
            if b != a:
                L[a] *= (x - xb) / (xa - xb)
    return L


@pytest.mark.parametrize("p", [2, 4])
def test_potential_block_constant_and_linear(p):
    """The potential block of H_e (the local potential's int N_i V N_j,
    src/ksdft/KohnShamOperatorContextFE.t.cpp, computed on the cell's Gauss
    rule of p + 1 points per direction) pinned by two potentials:
    V = 1.7 must give exactly 1.7 M_e, and V = v0 + g . x, whose integrand
    N_i V N_j is of degree 2p + 1 per direction, must equal its exact
    integral, formed here from 1-D factors on a 12-point Gauss rule in
    numpy (the nodes of N are the GLL points).  A missing factor, a wrong
    Jacobian or a potential evaluated at the wrong physical points fails."""
    lo, hi = (-1.5, -0.5, 0.25), (2.5, 1.5, 3.25)
    P = O.Problem.build_affine([1.7, 0, 0, 0], p=p, n=(2, 3, 2), lo=lo, hi=hi)
    d = P.arrays()
    V = d["H"] - 0.5 * d["K"]
    assert np.abs(V - 1.7 * d["M"]).max() <= 1e-13 * np.abs(d["M"]).max()
    v = np.array([0.3, -1.1, 0.7, 2.3])
    P = O.Problem.build_affine(v, p=p, n=(2, 3, 2), lo=lo, hi=hi)
    d = P.arrays()
    V = d["H"] - 0.5 * d["K"]
    np1 = p + 1
    nodes, _ = O.gll(np1)
    xq, wq = np.polynomial.legendre.leggauss(12)
    L = _lagrange_1d(nodes, xq)
    M1 = (L * wq) @ L.T                       # int L_a L_b over [-1, 1]
    X1 = (L * wq * xq) @ L.T                  # int xi L_a L_b
    clo, ch = P.cell_boxes()
    off = 0
    for e in range(P.nCells):
        ne = np1 ** 3
        assert d["cellNdofs"][e] == ne
        half = ch[e] / 2
        mid = clo[e] + half
        # x_d = mid_d + half_d xi_d: int x_d L L = mid_d M1 + half_d X1
        F = [[M1 * half[k], (mid[k] * M1 + half[k] * X1) * half[k]] for k in range(3)]
        Ve = np.zeros((ne, ne))
        for term in range(4):
            f = [F[k][1 if term == k + 1 else 0] for k in range(3)]
            # local index a + np (b + np c): x fastest
            Ve += v[term] * np.kron(f[2], np.kron(f[1], f[0]))
        got = V[off:off + ne * ne].reshape(ne, ne, order="F")
        assert np.abs(got - Ve).max() <= 1e-12 * np.abs(Ve).max()
        off += ne * ne


def test_coulomb_potential_block_integrates_the_smeared_potential():
    """With the partition of unity sum_i N_i = 1, the entries of a classical
    cell's potential block add up to the integral of V over the cell.  Over
    8^3 cells at p = 4 around one Z = 3 centre, V = -Z erf(r / 0.5) / r (a
    smooth function) summed so must match the integral that numpy forms on
    a 20-point Gauss rule per cell direction, to the p + 1-point rule's
    error: a missing Z, another smearing width or sign fails.  V is linear
    in Z: the block for Z = 3 is 3 times the block for Z = 1."""
    c, Z, sig = np.array([0.3, -0.2, 0.1]), 3, 0.5
    P = O.Problem.build(p=4, n=(8, 8, 8), lo=(-4.,) * 3, hi=(4.,) * 3,
                        atoms=[(c[0], c[1], c[2], Z)], v_smear=sig)
    d = P.arrays()
    got = float(np.sum(d["H"] - 0.5 * d["K"]))
    from scipy.special import erf
    xq, wq = np.polynomial.legendre.leggauss(20)
    clo, ch = P.cell_boxes()
    ref = 0.0
    for e in range(P.nCells):
        pts = [clo[e][k] + ch[e][k] * (xq + 1) / 2 for k in range(3)]
        X, Y, Zc = np.meshgrid(*pts, indexing="ij")
        r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Zc - c[2]) ** 2)
        w = np.einsum("i,j,k->ijk", wq, wq, wq) * np.prod(ch[e]) / 8
        ref += float(np.sum(w * -Z * erf(r / sig) / r))
    # the p + 1-point rule is within 1e-8 of it here; a smearing width 10 %
    # off moves the integral by 1e-3
    assert abs(got - ref) <= 1e-6 * abs(ref)
    # the integral scales with Z exactly
    P1 = O.Problem.build(p=4, n=(8, 8, 8), lo=(-4.,) * 3, hi=(4.,) * 3,
                         atoms=[(c[0], c[1], c[2], 1)], v_smear=sig)
    d1 = P1.arrays()
    assert np.abs((d["H"] - 0.5 * d["K"]) - Z * (d1["H"] - 0.5 * d1["K"])).max() <= \
        1e-13 * np.abs(d["H"]).max()


def test_cell_matrix_invariants():
    P = c1_like(n=3, p=3)
    d = P.arrays()
    for e in (0, 13):
        K = d["K"][e * 4096:(e + 1) * 4096].reshape(64, 64, order="F")
        M = d["M"][e * 4096:(e + 1) * 4096].reshape(64, 64, order="F")
        H = d["H"][e * 4096:(e + 1) * 4096].reshape(64, 64, order="F")
        assert np.abs(K.sum(axis=1)).max() < 1e-12
        assert abs(M.sum() - (20 / 3) ** 3) < 1e-10
        assert np.abs(H - H.T).max() < 1e-14
        assert np.linalg.eigvalsh(M).min() > 0


def test_apply_equals_dense_CtAC():
    for P in (hanging(p=2, n=3), c1_like(n=2, p=2)):
        A = dense_cell_matrix(P, "H")
        Cm = constraint_matrix(P)
        X = guess(P, 3, seed=7)
        X0 = X.copy()
        Y = P.apply(O.OP_H, X)
        ref = Cm.T @ A @ Cm @ X0
        assert np.abs(Y - ref).max() < 1e-12 * np.abs(ref).max()


def test_enriched_apply_equals_dense():
    P = enriched(p=2, n=3, nq_enr=5)
    assert P.nEnriched == 10
    A = dense_cell_matrix(P, "H")
    Cm = constraint_matrix(P)
    X = guess(P, 2, seed=3)
    Y = P.apply(O.OP_H, X.copy())
    ref = Cm.T @ A @ Cm @ X
    assert np.abs(Y - ref).max() < 1e-12 * np.abs(ref).max()


def test_enrichment_function_gradients_and_cutoff():
    atom = np.array([0.3, -0.2, 0.5])
    rng = np.random.default_rng(1)
    for k in range(5):
        for Z in (1, 6):
            for _ in range(5):
                x = atom + rng.uniform(-3, 3, 3)
                v, g = O.enrichment(atom, Z, k, x)
                fd = np.zeros(3)
                for d in range(3):
                    e = np.zeros(3); e[d] = 1e-6
                    fd[d] = (O.enrichment(atom, Z, k, x + e)[0] -
                             O.enrichment(atom, Z, k, x - e)[0]) / 2e-6
                assert np.allclose(g, fd, rtol=1e-5, atol=1e-9)
        # support ends at rcut (1 + 1/smooth) = 6
        assert O.enrichment(atom, 1, k, atom + [0, 0, 6.01])[0] == 0.0
    # 1s inside the cutoff: Y00 e^{-Zr}
    v, _ = O.enrichment(atom, 1, 0, atom + [1.0, 0, 0])
    assert abs(v - 0.5 / np.sqrt(np.pi) * np.exp(-1.0)) < 1e-15
    # 2p_x ~ x: sqrt(3/4pi) x e^{-Zr}
    v, _ = O.enrichment(atom, 1, 4, atom + [0.6, 0.8, 0])
    assert abs(v - np.sqrt(3 / (4 * np.pi)) * 0.6 * np.exp(-1.0)) < 1e-15


def test_gll_overlap_inverse_diagonal():
    P = hanging(p=3)
    d = P.arrays()
    # the GLL mass diagonal integrates 1 over the domain
    assert abs(d["mgll"].sum() - 16.0 ** 3) < 1e-9
    assert np.all(d["dinv"][d["rows"]] == 0.0)
    free = np.setdiff1d(np.arange(P.nDofs), d["rows"])
    assert np.all(d["dinv"][free] > 0)


# ------------------------------------------------------- dense LA pieces
def test_syev_jacobi_against_numpy():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((12, 12)); A = A + A.T
    w, Z = O.syev(A)
    assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-12)
    assert np.allclose(A @ Z, Z * w, atol=1e-11)


def test_filter_on_diagonal_operator_closed_form():
    lam = np.linspace(0.0, 10.0, 40)
    N = len(lam)
    P = O.Problem.from_arrays(np.ones(N), np.arange(N), N, N, H=lam, M=np.ones(N),
                              dinv=np.ones(N), mgll=np.ones(N), p=1)
    a, b, c, m = 0.0, 2.0, 10.0, 12
    X = np.ones((N, 2)); X[:, 1] = np.linspace(1, 2, N)
    e, c0 = (c - b) / 2, (c + b) / 2
    cheb = lambda t: np.cos(m * np.arccos(np.clip(t, -1, 1))) if np.all(np.abs(t) <= 1) else \
        np.where(np.abs(t) <= 1, np.cos(m * np.arccos(np.clip(t, -1, 1))),
                 np.cosh(m * np.arccosh(np.maximum(np.abs(t), 1))) * np.sign(t) ** m)
    scale = cheb((lam - c0) / e) / cheb(np.array([(a - c0) / e]))
    Y = P.chebyshev_filter(X, m, a, b, c)
    assert np.allclose(Y, X * scale[:, None], rtol=1e-10, atol=1e-12)
    Yr = P.residual_filter(np.array([0.7, 3.0]), X, m, a, b, c)
    assert np.allclose(Yr, X * scale[:, None], rtol=1e-9, atol=1e-11)


def test_cholgs_orthonormal_and_rr_matches_dense():
    P = c1_like(n=3, p=2)
    X = guess(P, 6, seed=2)
    rc, Xo = P.cholgs(X.copy())
    assert rc == 0
    M = constraint_matrix(P).T @ dense_cell_matrix(P, "M") @ constraint_matrix(P)
    H = constraint_matrix(P).T @ dense_cell_matrix(P, "H") @ constraint_matrix(P)
    Xo_p = P.p2c(Xo.copy())
    assert np.abs(Xo_p.T @ M @ Xo_p - np.eye(6)).max() < 1e-12
    rc, eig, Xr = P.rayleigh_ritz(Xo)
    want = sl.eigh(Xo_p.T @ H @ Xo_p, Xo_p.T @ M @ Xo_p, eigvals_only=True)
    assert np.allclose(eig, want, atol=1e-12)


def test_cholgs_reports_rank_deficiency():
    P = c1_like(n=2, p=2)
    X = guess(P, 3, seed=4)
    X[:, 2] = X[:, 1]
    rc, _ = P.cholgs(X)
    assert rc == 1  # OrthonormalizationErrorCode::LAPACK_ERROR


def test_cg_matches_direct_solve():
    P = hanging(p=2, n=3)
    d = P.arrays()
    K = constraint_matrix(P).T @ dense_cell_matrix(P, "K") @ constraint_matrix(P)
    free = np.setdiff1d(np.arange(P.nDofs), d["rows"])
    diagK = np.diag(dense_cell_matrix(P, "K")).copy().reshape(-1, 1)
    diagK = P.c2p(diagK)[:, 0]
    diagK[d["rows"]] = 1.0
    b = guess(P, 3, seed=5)
    b = P.c2p(b)
    rc, x, it = P.cg(O.OP_K, 1.0 / diagK, b, abs_tol=1e-14, rel_tol=1e-13)
    assert rc == 0 and it > 2
    xd = np.zeros_like(b)
    xd[free] = np.linalg.solve(K[np.ix_(free, free)], b[free])
    cond = np.linalg.cond(K[np.ix_(free, free)])
    assert np.abs(x[free] - xd[free]).max() <= 1e-13 * cond * np.abs(xd).max()


def test_cg_max_iter_reports_failed_to_converge():
    P = c1_like(n=2, p=2)
    d = P.arrays()
    b = P.c2p(guess(P, 2, seed=9))
    rc, x, it = P.cg(O.OP_K, np.ones(P.nDofs), b, abs_tol=1e-30, rel_tol=1e-30, max_iter=1)
    assert rc == 1  # LinearSolverErrorCode::FAILED_TO_CONVERGE


def _one_d_gep(p, n, L):
    """1-D consistent stiffness / mass on GLL-node Lagrange bases, Dirichlet ends,
    built independently of the oracle (numpy Gauss rule, finite-free derivative)."""
    xg, wg = np.polynomial.legendre.leggauss(p + 1)
    nodes = np.array(O.gll(p + 1)[0])
    V = np.vander(nodes, p + 1, increasing=True)
    coef = np.linalg.inv(V)  # columns: monomial coefficients of each Lagrange basis
    Lq = np.vander(xg, p + 1, increasing=True) @ coef
    dV = np.hstack([np.zeros((p + 1, 1)),
                    np.vander(xg, p, increasing=True) * np.arange(1, p + 1)])
    Dq = dV @ coef
    h = L / n
    Ne = n * p + 1
    K = np.zeros((Ne, Ne)); M = np.zeros((Ne, Ne))
    for e in range(n):
        s = slice(e * p, e * p + p + 1)
        M[s, s] += (Lq.T * wg) @ Lq * h / 2
        K[s, s] += (Dq.T * wg) @ Dq * 2 / h
    return K[1:-1, 1:-1], M[1:-1, 1:-1]


def test_chfsi_v0_tensor_closed_form():
    """V = 0 on a uniform mesh: eigenvalues of (1/2 K, M) are sums of three 1-D
    generalized eigenvalues (exact tensor structure)."""
    p, n, L = 3, 3, 20.0
    K1, M1 = _one_d_gep(p, n, L)
    l1 = sl.eigh(0.5 * K1, M1, eigvals_only=True)
    all3 = np.sort((l1[:, None, None] + l1[None, :, None] + l1[None, None, :]).ravel())
    P = O.Problem.build(p=p, n=(n, n, n), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[])
    nev = 10  # 4 checked + buffer vectors past the degenerate cluster
    X = guess(P, nev, seed=1)
    lo, hi, beta = P.lanczos(O.rng_uniform(2, -1, 1, P.nDofs))
    c = hi + beta / 10
    a = lo
    b = min((c - a) * nev * 200 / P.nDofs + a, (c + a) / 2)
    eig = np.zeros(nev)
    for it in range(12):
        rc, eig, Xn, _ = P.chfsi(a, b, c, 25, X, eig, residual=(it > 0))
        assert rc == 0
        X = Xn
        a, b = eig[0], eig[-1]
    assert np.abs(eig[:4] - all3[:4]).max() < 1e-10


def test_lanczos_bound_matches_dense_lumped_spectrum():
    P = c1_like(n=2, p=3)
    d = P.arrays()
    free = np.setdiff1d(np.arange(P.nDofs), d["rows"])
    H = constraint_matrix(P).T @ dense_cell_matrix(P, "H") @ constraint_matrix(P)
    Hf = H[np.ix_(free, free)]
    w = sl.eigh(Hf, np.diag(1.0 / d["dinv"][free]), eigvals_only=True)
    lo, hi, beta = P.lanczos(O.rng_uniform(11, -1, 1, P.nDofs), max_krylov=40)
    assert abs(hi - w[-1]) < 1e-6 * abs(w[-1])
    assert lo >= w[0] - 1e-8


def test_cpu_baseline_timing_loop_computes_the_apply():
    """The threaded CPU timing loop (slab-local accumulation + parallel merge)
    produces the operator apply itself, on a mesh with hanging nodes."""
    P = hanging()
    X = guess(P, 5, seed=3)
    ref = P.apply(O.OP_H, X.copy())
    for threads in (1, 3):
        t, tc, tv, Y = P.time_apply(O.OP_H, X.copy(), reps=1, threads=threads)
        assert t > 0 and tc > 0 and tv >= 0
        assert np.abs(Y - ref).max() <= 1e-12 * np.abs(ref).max()


# ------------------------------------------------------------ KS driver pieces
def test_fermi_level_counts_electrons():
    """At the Newton-Raphson Fermi level sum_i 2 f_i = Ne; with a gap much
    larger than kT it lies between HOMO and LUMO, and a symmetric spectrum
    with a half-filled degenerate pair puts it exactly at the pair."""
    eig = np.array([-3.0, -2.0, -1.0, -0.5, 0.2, 0.9])
    T = 500.0
    kT = O.K_BOLTZMANN_HARTREE * T
    for ne in (2, 4, 6, 8):
        ef = O.fermi_level(eig, ne, T, 1e-12)
        occ = sum(2 * O.fermi_dirac(e, ef, O.K_BOLTZMANN_HARTREE, T) for e in eig)
        assert abs(occ - ne) < 1e-9
        homo, lumo = eig[ne // 2 - 1], eig[ne // 2]
        assert homo + 20 * kT < ef < lumo - 20 * kT
    sym = np.array([-2.0, -1.0, -1.0, 0.0])  # 4 electrons: 2 in -2, 2 shared by the pair
    ef = O.fermi_level(sym, 4, T, 1e-12)
    assert abs(ef - (-1.0)) < 1e-9


def test_fermi_dirac_limits_and_derivative():
    kb, T, mu = O.K_BOLTZMANN_HARTREE, 300.0, 0.1
    assert O.fermi_dirac(mu, mu, kb, T) == 0.5
    assert O.fermi_dirac(mu - 1.0, mu, kb, T) == 1.0
    assert O.fermi_dirac(mu + 1.0, mu, kb, T) < 1e-300
    for e in (0.099, 0.1, 0.1005):
        h = 1e-7
        fd = (O.fermi_dirac(e, mu + h, kb, T) - O.fermi_dirac(e, mu - h, kb, T)) / (2 * h)
        assert abs(fd - O.fermi_dirac_der(e, mu, kb, T)) < 1e-5 * max(1.0, abs(fd))


def test_cheby_degree_lookup_boundaries():
    """lower_bound over the table keys: a key equal to a table entry takes
    that entry, one past it the next; beyond the last entry 1250."""
    assert O.cheby_degree(0) == 6 and O.cheby_degree(10) == 6 and O.cheby_degree(11) == 9
    assert O.cheby_degree(750) == 30 and O.cheby_degree(751) == 39
    assert O.cheby_degree(500000) == 1000 and O.cheby_degree(500001) == 1250
    assert O.cheby_degree(10.9) == 6  # truncation to the integer key


def test_residual_norms_on_diagonal_operator():
    """||H x - eps M x|| for a unit vector e_i of a diagonal pencil is |eps m_i - lam_i|."""
    lam = np.linspace(-1.0, 4.0, 12)
    mdiag = np.linspace(1.0, 2.0, 12)
    N = len(lam)
    P = O.Problem.from_arrays(np.ones(N), np.arange(N), N, N, H=lam, M=mdiag,
                              dinv=1.0 / mdiag, mgll=mdiag, p=1)
    X = np.zeros((N, 3)); X[2, 0] = 1.0; X[5, 1] = 1.0; X[7, 2] = 1.0
    eps = np.array([lam[2], lam[5] + 0.25, lam[7] - 1.0])
    res = P.residual_norms(X, eps)
    want = np.abs(eps * mdiag[[2, 5, 7]] - lam[[2, 5, 7]])
    assert np.allclose(res, want, rtol=1e-14, atol=1e-15)


def test_orthogonalized_enrichment_is_orthogonal_on_interior_nodes():
    """After the projection M_II alpha = d, int N_c phi~_g = 0 for every
    classical node c interior to atom g's support (all of c's cells enriched
    by g, c unconstrained): checked on the assembled M^ce."""
    P = O.Problem.build(p=2, n=(6, 6, 6), lo=(-12.,) * 3, hi=(12.,) * 3,
                        atoms=[(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 6)],
                        enrich=True, orthogonalize=True)
    d = P.arrays()
    nd, cmap, M = d["cellNdofs"], d["cellMap"], d["M"]
    nC, nE = P.nClassical, P.nEnriched
    Mce = np.zeros((nC, nE))
    inc = np.zeros(nC, dtype=np.int64)
    ncl = (P.p + 1) ** 3
    cnt = {}
    off = moff = 0
    for e, n in enumerate(nd):
        ids = cmap[off:off + n]
        Me = M[moff:moff + n * n].reshape(n, n, order="F")
        inc[ids[:ncl]] += 1
        for jj in range(ncl, n):
            g = ids[jj] - nC
            Mce[ids[:ncl], g] += Me[:ncl, jj]
            for c in ids[:ncl]:
                cnt[(g // 5, c)] = cnt.get((g // 5, c), 0) + 1
        off += n
        moff += n * n
    rows = set(d["rows"].tolist())
    worst, scale, checked = 0.0, np.abs(Mce).max(), 0
    for (a, c), k in cnt.items():
        if k // 5 == inc[c] and c not in rows:
            worst = max(worst, np.abs(Mce[c, 5 * a:5 * a + 5]).max())
            checked += 1
    assert checked > 10 and worst < 1e-10 * scale
    # and the projection did something: without it the same entries are O(1)
    Q = O.Problem.build(p=2, n=(6, 6, 6), lo=(-12.,) * 3, hi=(12.,) * 3,
                        atoms=[(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 6)],
                        enrich=True, orthogonalize=False)
    assert np.abs(Q.arrays()["M"] - M).max() > 1e-3 * scale


def test_ks_eigensolve_v0_converges_to_tensor_closed_form():
    """The KS driver (Lanczos bounds, warm-up pass, residual-filter passes,
    Fermi level, occupancy-based convergence) on V = 0: the occupied states
    converge to sums of 1-D generalized eigenvalues."""
    p, n, L = 3, 3, 20.0
    K1, M1 = _one_d_gep(p, n, L)
    l1 = sl.eigh(0.5 * K1, M1, eigvals_only=True)
    all3 = np.sort((l1[:, None, None] + l1[None, :, None] + l1[None, None, :]).ravel())
    P = O.Problem.build(p=p, n=(n, n, n), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[])
    X0 = guess(P, 10, seed=1)
    lg = P.p2c(O.rng_uniform(2, -1, 1, P.nDofs).reshape(-1, 1))[:, 0]
    r = P.ks_eigensolve(X0, lg, degree=25, tol=1e-8, max_pass=40, num_electrons=2)
    assert r["rc"] == 0
    occ = r["nocc"]
    assert occ >= 1 and np.all(r["res"][:occ] <= 1e-8)
    assert abs(r["eig"][0] - all3[0]) < 1e-10


def _assembled_mass_apply(P, X):
    """sum_e Q_e^T M_e Q_e X with no constraints (the density's own basis)."""
    d = P.arrays()
    MX = np.zeros_like(X)
    offm = offc = 0
    for n in d["cellNdofs"]:
        ids = d["cellMap"][offc:offc + n]
        Me = d["M"][offm:offm + n * n].reshape(n, n, order="F")
        MX[ids] += Me @ X[ids]
        offc += n
        offm += n * n
    return MX


@pytest.mark.parametrize("enrich", [False, True])
def test_density_integrates_to_the_occupied_mass(enrich):
    """sum_q JxW rho = sum_i 2 f_i x_i^T M x_i holds exactly for the
    density's interpolation (M is assembled from the same basis values and
    weights): a wrong factor, square, occupation index, weight or basis
    layout breaks it.  The weights integrate the box volume."""
    P = O.Problem.build(p=2, n=(3, 3, 3), lo=(-6.,) * 3, hi=(6.,) * 3,
                        atoms=[(0.3, -0.2, 0.1, 6)], enrich=enrich, orthogonalize=enrich,
                        keep_basis=True)
    off, w = P.quadrature()
    assert abs(w.sum() - 12.0 ** 3) < 1e-10
    X = O.rng_uniform(3, -1, 1, P.nDofs * 5).reshape(P.nDofs, 5)
    occ = np.array([1.0, 0.8, 0.5, 0.2, 0.0])
    rho = P.density(X, occ)
    assert rho.min() >= 0.0
    rhs = float(np.sum(2.0 * occ * np.sum(X * _assembled_mass_apply(P, X), axis=0)))
    assert abs((w * rho).sum() - rhs) < 1e-13 * abs(rhs)


def test_density_of_the_constant_function():
    """The Lagrange basis is a partition of unity: X = 1 on every node gives
    psi = 1 at every quadrature point, so rho = 2 f exactly."""
    P = O.Problem.build(p=3, n=(2, 3, 2), lo=(-3.,) * 3, hi=(3.,) * 3, keep_basis=True)
    rho = P.density(np.ones((P.nDofs, 1)), np.array([0.75]))
    assert np.abs(rho - 1.5).max() < 1e-13
