"""Integration against the basis (FEBasisOperations::integrateWithBasisValues)
and the Poisson linear-solver function (PoissonLinearSolverFunctionFE) on the
GPU against the oracle: right-hand side, Jacobi diagonal, CG iterations and
solution, at the C2 size (16^3 cells, p = 3, 117 649 dofs, 8 right-hand
sides) and on a Gaussian charge with inhomogeneous Dirichlet data."""
import numpy as np
import pytest

import dfx
import oracle as O
from dfx import configs
from problems import quad_xyz

pytestmark = pytest.mark.gpu

SPECS = {
    "c1": dict(p=3, n=(4, 4, 4), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[(0., 0., 0., 1)]),
    "enr": dict(p=2, n=(6, 6, 6), lo=(-12.,) * 3, hi=(12.,) * 3,
                atoms=[(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 6)], enrich=True),
}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def mv(ctx, a):
    return dfx.MultiVector.from_numpy(ctx, a)


@pytest.mark.parametrize("name", list(SPECS))
def test_integrate_matches_oracle(ctx, name):
    spec = SPECS[name]
    orth = spec.get("enrich", False)
    G = dfx.Problem(ctx, orthogonalize=orth, **spec)
    R = O.Problem.build(orthogonalize=orth, keep_basis=True, **spec)
    off, w = R.quadrature()
    assert np.array_equal(G.get("quadOffsets"), off)
    assert rel(G.get("quadW"), w) < 1e-14
    xyz = G.get("quadXYZ").reshape(-1, 3)
    assert np.abs(xyz - quad_xyz(R)).max() < 1e-13 * np.abs(xyz).max()
    f = O.rng_uniform(3, -1, 1, int(off[-1]) * 3).reshape(-1, 3)
    H, M, K, Minv, Mlump, cons = G.operators(keep_fp32=False)
    got = G.integrate(mv(ctx, f), cons).to_numpy()
    ref = R.integrate(f, c2p=True)
    # enriched: the orthogonalization coefficients agree to 1e-9
    assert rel(got, ref) < (1e-12 if not orth else 1e-8)
    raw = G.integrate(mv(ctx, f)).to_numpy()
    assert rel(raw, R.integrate(f, c2p=False)) < (1e-12 if not orth else 1e-8)


def _c2_rhs(xq, cfg):
    """4 pi rho_k at the quadrature points: four normalized Gaussians of
    width cfg['width'] per column, centres and charges of seed 100 + k."""
    w = cfg["width"]
    f = np.zeros((xq.shape[0], cfg["rhs"]))
    for k in range(cfg["rhs"]):
        cen, q = configs.poisson_rhs_charges(100 + k)
        for c, qq in zip(cen, q):
            r2 = ((xq - c) ** 2).sum(axis=1)
            f[:, k] += 4 * np.pi * qq * (2 * np.pi * w * w) ** -1.5 * np.exp(-r2 / (2 * w * w))
    return f


def test_c2_poisson_matches_oracle(ctx):
    """C2 at full size: the library's Poisson function (quadrature RHS, zero
    Dirichlet, Jacobi) and CG against the oracle's: right-hand side 1e-12,
    Jacobi diagonal 1e-13, iterations within one, solution 1e-8."""
    cfg = configs.config("c2")
    G = dfx.Problem(ctx, p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"], atoms=[],
                    orthogonalize=False)
    R = O.Problem.build(p=cfg["p"], n=cfg["n"], lo=cfg["lo"], hi=cfg["hi"], atoms=[],
                        keep_basis=True)
    assert G.nDofs == R.nDofs == 117649
    xq = G.get("quadXYZ").reshape(-1, 3)
    f = _c2_rhs(xq, cfg)
    H, M, K, Minv, Mlump, cons = G.operators(keep_fp32=False)
    fn = dfx.PoissonLinearSolverFunctionFE(G, K, cons, mv(ctx, f), dfx.PC_JACOBI)
    b_ref, diag_ref, uD_ref = R.poisson(np.zeros(R.nRows), f)
    b = fn.getRhs().to_numpy()
    assert rel(b, b_ref) < 1e-12
    ones = mv(ctx, np.ones((G.nDofs, 1)))
    dinv = dfx.MultiVector(ctx, G.nDofs, 1)
    fn.getPCContext().apply(ones, dinv)
    assert rel(dinv.to_numpy()[:, 0], 1.0 / diag_ref) < 1e-13
    cg = dfx.CGLinearSolver(ctx, 10000, 1e-30, cfg["rel_tol"], 1e10)
    rc = cg.solve(fn)
    sol = fn.getSolution().to_numpy()
    rco, x_ref, it_ref = R.cg(O.OP_K, 1.0 / diag_ref, b_ref, abs_tol=1e-30,
                              rel_tol=cfg["rel_tol"], max_iter=10000)
    print(f"C2: GPU {cg.iterations} iterations, oracle {it_ref}", flush=True)
    assert rc == 0 and rco == 0
    assert abs(cg.iterations - it_ref) <= 1
    assert rel(sol, x_ref + uD_ref) < 1e-8
    # the one-call path (CG + getSolution in the library) gives the same bits
    fn2 = dfx.PoissonLinearSolverFunctionFE(G, K, cons, mv(ctx, f), dfx.PC_JACOBI)
    out = dfx.MultiVector(ctx, G.nDofs, cfg["rhs"])
    rc2, it2, res2 = fn2.solve(10000, 1e-30, cfg["rel_tol"], 1e10, out)
    assert rc2 == 0 and it2 == cg.iterations
    assert np.array_equal(out.to_numpy(), sol)


def test_poisson_gaussian_charge_gives_the_erf_potential(ctx):
    """-lap(phi) = 4 pi rho for a normalized Gaussian charge (exponent 0.5)
    has phi = erf(sqrt(0.5) r) / r.  The library's Poisson function takes the
    Dirichlet data phi in the field constraints' inhomogeneities (lifted:
    b = integral of f - C2P(K u_D)), the right-hand side at the quadrature
    points, Jacobi PCG to 1e-12 on the GPU-built p = 4 mesh; the nodal error
    is the discretization error, as the oracle's own run of the same set-up
    gives.  Its right-hand side and diagonal match the oracle's."""
    from scipy.special import erf
    spec = dict(p=4, n=(8, 8, 8), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[])
    P = dfx.Problem(ctx, orthogonalize=False, **spec)
    N = P.nDofs
    a, c = 0.5, np.array([0.3, -0.2, 0.1])

    def phi(x):
        r = np.linalg.norm(x - c, axis=1)
        return np.where(r > 1e-12, erf(np.sqrt(a) * r) / np.maximum(r, 1e-300),
                        2 * np.sqrt(a / np.pi))

    xn = P.get("nodeXYZ").reshape(-1, 3)
    xq = P.get("quadXYZ").reshape(-1, 3)
    rq = np.linalg.norm(xq - c, axis=1)
    f = (4 * np.pi * (a / np.pi) ** 1.5 * np.exp(-a * rq * rq)).reshape(-1, 1)
    rows = P.get("rows")
    H, M, K, Minv, Mlump, cons = P.operators(keep_fp32=False)
    field = dfx.Constraints(ctx, N, rows, P.get("rowSizes"), P.get("rowOffsets"),
                            P.get("cols"), P.get("weights"), phi(xn)[rows])
    fn = dfx.PoissonLinearSolverFunctionFE(P, K, field, mv(ctx, f))
    R = O.Problem.build(keep_basis=True, **spec)
    b_ref, diag_ref, uD_ref = R.poisson(phi(xn)[rows], f)
    assert rel(fn.getRhs().to_numpy(), b_ref) < 1e-12
    cg = dfx.CGLinearSolver(ctx, 5000, 1e-14, 1e-12, 1e10)
    rc = cg.solve(fn)
    sol = fn.getSolution().to_numpy()[:, 0]
    err = np.abs(sol - phi(xn)).max() / np.abs(phi(xn)).max()
    print(f"Poisson Gaussian: {cg.iterations} iterations, max rel nodal error {err:.2e}",
          flush=True)
    assert rc == 0 and err < 1.5e-3
    assert np.array_equal(sol[rows], phi(xn)[rows])       # Dirichlet data kept exactly
    # the solution is the next initial guess (getSolution), so a second solve
    # from it converges at once
    first = cg.iterations
    rc = cg.solve(fn)
    assert rc == 0 and cg.iterations < first // 4
