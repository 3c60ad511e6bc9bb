"""Pins of the oracle's integration against the basis and its Poisson
set-up (PoissonLinearSolverFunctionFE), on the CPU.

integrate(f) is pinned by what the mathematics fixes: M = F^T W F with the
same basis values F and weights W, so integrating the interpolant of any
nodal vector y gives the assembled mass matrix times y; integrating the
constant 1 gives the row sums of M (partition of unity of the classical
basis); a linear function lies in the FE space and integrates to M times its
nodal values.  The Poisson set-up is pinned by its identities (b + C2P(K u_D)
equals the integral, u_D carries the Dirichlet data, the Jacobi diagonal is
e_i^T K e_i) and by the physics: a Gaussian charge gives the potential
erf(sqrt(a) r) / r to the discretization error."""
import numpy as np
import pytest

import oracle as O
from problems import assembled, quad_xyz


def c1_small(p=3, n=4):
    return O.Problem.build(p=p, n=(n, n, n), lo=(-10.,) * 3, hi=(10.,) * 3,
                           atoms=[(0.0, 0.0, 0.0, 1)], keep_basis=True)


def hanging_small():
    return O.Problem.build(p=3, n=(4, 4, 4), lo=(-8.,) * 3, hi=(8.,) * 3,
                           ref_lo=(1, 1, 1), ref_hi=(3, 3, 2), atoms=[(0.3, -0.2, 0.1, 1)],
                           keep_basis=True)


def enriched_small():
    return O.Problem.build(p=2, n=(6, 6, 6), lo=(-12.,) * 3, hi=(12.,) * 3,
                           atoms=[(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 6)], enrich=True,
                           orthogonalize=True, keep_basis=True)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_integral_of_one_is_the_mass_row_sum():
    P = c1_small()
    off, _ = P.quadrature()
    b = P.integrate(np.ones((int(off[-1]), 1)), c2p=False)[:, 0]
    M = assembled(P, "M")
    assert rel(b, np.asarray(M.sum(axis=1)).ravel()) < 1e-13


@pytest.mark.parametrize("make", [c1_small, hanging_small, enriched_small])
def test_integral_of_an_interpolant_is_mass_times_the_vector(make):
    P = make()
    y = O.rng_uniform(5, -1, 1, P.nDofs * 3).reshape(-1, 3)
    f = P.interpolate(y)
    b = P.integrate(f, c2p=False)
    ref = assembled(P, "M") @ y
    assert rel(b, ref) < 1e-12
    # the child -> parent step is the constraints' own
    assert np.array_equal(P.integrate(f, c2p=True), P.c2p(b.copy()))


def test_linear_function_integrates_to_mass_times_nodal_values():
    P = c1_small(p=2)
    xq = quad_xyz(P)
    xn = P.node_xyz()
    g = lambda x: 0.3 + 0.7 * x[:, 0] - 0.2 * x[:, 1] + 0.1 * x[:, 2]
    b = P.integrate(g(xq).reshape(-1, 1), c2p=False)[:, 0]
    assert rel(b, assembled(P, "M") @ g(xn)) < 1e-12


def test_poisson_setup_identities():
    P = c1_small()
    d = P.arrays()
    rows = d["rows"]
    xn = P.node_xyz()
    inhom = np.sin(xn[rows, 0]) + 0.5 * xn[rows, 2]
    off, _ = P.quadrature()
    f = O.rng_uniform(8, -1, 1, int(off[-1]) * 2).reshape(-1, 2)
    b, diag, uD = P.poisson(inhom, f)
    # u_D: the Dirichlet data on the constrained rows, zero elsewhere
    ref = np.zeros(P.nDofs)
    ref[rows] = inhom
    assert np.array_equal(uD[:, 0], ref) and np.array_equal(uD[:, 1], ref)
    # b + C2P(K u_D) = C2P(integral of f)
    Ku = P.c2p(assembled(P, "K") @ uD)
    assert rel(b + Ku, P.integrate(f, c2p=True)) < 1e-12
    # the constrained rows of b are zero (homogeneous child -> parent step)
    assert np.all(b[rows] == 0.0)
    # Jacobi diagonal: e_i^T K e_i on free rows, 1 on constrained rows
    assert np.all(diag[rows] == 1.0)
    free = np.setdiff1d(np.arange(P.nDofs), rows)
    for i in free[:: max(1, len(free) // 7)]:
        e = np.zeros((P.nDofs, 1))
        e[i] = 1.0
        assert abs(P.apply(O.OP_K, e)[i, 0] - diag[i]) < 1e-13 * abs(diag[i])


def test_gaussian_charge_gives_the_erf_potential():
    """-lap(phi) = 4 pi rho for a normalized Gaussian (exponent a) has
    phi = erf(sqrt(a) r) / r: Dirichlet data from phi, right-hand side the
    quadrature of 4 pi rho, Jacobi PCG; the nodal error is the
    discretization error of the p = 4 mesh."""
    from scipy.special import erf
    P = O.Problem.build(p=4, n=(8, 8, 8), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[],
                        keep_basis=True)
    a, c = 0.5, np.array([0.3, -0.2, 0.1])

    def phi(x):
        r = np.linalg.norm(x - c, axis=1)
        return np.where(r > 1e-12, erf(np.sqrt(a) * r) / np.maximum(r, 1e-300),
                        2 * np.sqrt(a / np.pi))

    xq, xn = quad_xyz(P), P.node_xyz()
    rq = np.linalg.norm(xq - c, axis=1)
    f = 4 * np.pi * (a / np.pi) ** 1.5 * np.exp(-a * rq * rq)
    rows = P.arrays()["rows"]
    b, diag, uD = P.poisson(phi(xn)[rows], f.reshape(-1, 1))
    rc, x, it = P.cg(O.OP_K, 1.0 / diag, b, abs_tol=1e-14, rel_tol=1e-12)
    assert rc == 0
    sol = (x + uD)[:, 0]
    err = np.abs(sol - phi(xn)).max() / np.abs(phi(xn)).max()
    assert err < 1.5e-3, err
