"""The GPU problem builder (product setup) against the oracle's FE setup on the
same specification: integer maps and constraint index sets bit-exact; cell
matrices, diagonals and the enrichment projection to FP64 tolerances."""
import numpy as np
import pytest

import dfx
import oracle as O

pytestmark = pytest.mark.gpu

SPECS = {
    "c1": dict(p=3, n=(4, 4, 4), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[(0., 0., 0., 1)]),
    "p4": dict(p=4, n=(3, 2, 2), lo=(-6., -4., -4.), hi=(6., 4., 4.),
               atoms=[(0.3, -0.2, 0.1, 1), (1.0, 0.5, -0.5, 6)]),
    "enr": dict(p=2, n=(6, 6, 6), lo=(-12.,) * 3, hi=(12.,) * 3,
                atoms=[(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 6)], enrich=True),
    "enr3": dict(p=3, n=(4, 4, 4), lo=(-12.,) * 3, hi=(12.,) * 3,
                 atoms=[(0.5, -0.5, 0.25, 6)], enrich=True),
}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(SPECS))
def test_problem_builder_matches_oracle(ctx, name):
    spec = SPECS[name]
    orth = spec.get("enrich", False)
    G = dfx.Problem(ctx, orthogonalize=orth, **spec)
    R = O.Problem.build(orthogonalize=orth, **spec)
    d = R.arrays()
    assert (G.nCells, G.nDofs, G.nClassical, G.nEnriched) == \
        (R.nCells, R.nDofs, R.nClassical, R.nEnriched)
    for key in ("cellNdofs", "cellMap", "rows", "rowSizes"):
        assert np.array_equal(G.get(key), d[key]), key
    assert G.nEntries == 0 and np.all(G.get("inhom") == 0.0)
    for key in ("M", "K", "H"):
        assert rel(G.get(key), d[key]) < 1e-12, key
    assert rel(G.get("dinv"), d["dinv"]) < 1e-14
    assert rel(G.get("mgll"), d["mgll"]) < 1e-14
    xyz = G.get("nodeXYZ").reshape(-1, 3)
    assert np.abs(xyz - R.node_xyz()).max() <= 1e-13 * np.abs(R.node_xyz()).max()
    if G.nEnriched:
        a_g, a_r = G.get("alpha").reshape(G.nEnriched, -1), R.alpha()
        assert rel(a_g, a_r) < 1e-9
        assert rel(G.get("enrMinv"), d["enrMinv"]) < 1e-8


def test_orthogonality_on_interior_nodes(ctx):
    """After the projection, int N_c phi~_g = 0 for every classical node c
    interior to atom g's support: check the assembled M^ce."""
    spec = SPECS["enr"]
    G = dfx.Problem(ctx, orthogonalize=True, **spec)
    nd, cmap, M = G.get("cellNdofs"), G.get("cellMap"), G.get("M")
    nC, nE = G.nClassical, G.nEnriched
    Mce = np.zeros((nC, nE))
    inc = np.zeros(nC, dtype=np.int64)
    ncl = (G.p + 1) ** 3
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
    rows = set(G.get("rows").tolist())
    worst, scale = 0.0, np.abs(Mce).max()
    for (a, c), k in cnt.items():
        if k // 5 == inc[c] and c not in rows:
            worst = max(worst, np.abs(Mce[c, 5 * a:5 * a + 5]).max())
    assert worst < 1e-10 * scale


def test_problem_operators_apply(ctx):
    """Operators built from the device-resident problem match operators built
    from the downloaded arrays (same kernels, same bits)."""
    G = dfx.Problem(ctx, orthogonalize=True, **SPECS["enr"])
    H, M, K, Minv, Mlump, cons = G.operators(keep_fp32=False)
    Hh = dfx.CellOperator(ctx, G.get("cellNdofs"), G.get("cellMap"), G.get("H"), G.nDofs,
                          G.nClassical, cons)
    X = O.rng_uniform(5, -1, 1, G.nDofs * 9).reshape(G.nDofs, 9)
    outs = []
    for op in (H, Hh):
        Xg = dfx.MultiVector.from_numpy(ctx, X)
        Y = dfx.MultiVector(ctx, G.nDofs, 9)
        op.apply(Xg, Y)
        outs.append(Y.to_numpy())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("name", list(SPECS))
def test_density_matches_oracle(ctx, name):
    """Electron density at the quadrature points (DensityCalculator::
    computeRho): the quadrature layout and weights as the oracle's, rho as the
    oracle's (enriched bases differ by the orthogonalization coefficients,
    which agree to 1e-9), and sum JxW rho = sum_i 2 f_i x_i^T M x_i with the
    GPU problem's own M."""
    spec = SPECS[name]
    orth = spec.get("enrich", False)
    G = dfx.Problem(ctx, orthogonalize=orth, **spec)
    R = O.Problem.build(orthogonalize=orth, keep_basis=True, **spec)
    off_r, w_r = R.quadrature()
    assert np.array_equal(G.get("quadOffsets"), off_r)
    w_g = G.get("quadW")
    assert np.abs(w_g - w_r).max() <= 1e-15 * w_r.max()
    X = O.rng_uniform(7, -1.0, 1.0, G.nDofs * 6).reshape(G.nDofs, 6)
    occ = np.array([1.0, 1.0, 0.9, 0.5, 0.1, 0.0])
    rho_g = G.density(occ, dfx.MultiVector.from_numpy(ctx, X)).to_numpy().ravel()
    rho_r = R.density(X, occ)
    assert rel(rho_g, rho_r) < (1e-8 if orth else 1e-12)
    # the identity with the GPU problem's own mass matrix
    nd, cmap, M = G.get("cellNdofs"), G.get("cellMap"), G.get("M")
    MX = np.zeros_like(X)
    off = moff = 0
    for n in nd:
        ids = cmap[off:off + n]
        MX[ids] += M[moff:moff + n * n].reshape(n, n, order="F") @ X[ids]
        off += n
        moff += n * n
    rhs = float(np.sum(2.0 * occ * np.sum(X * MX, axis=0)))
    assert abs(float(np.dot(w_g, rho_g)) - rhs) < 1e-11 * abs(rhs)
