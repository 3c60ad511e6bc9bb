"""Seeded problem instances shared by the tests (built by the oracle)."""
import numpy as np

import oracle as O


def c1_like(n=4, p=3, L=20.0):
    """Config-1 recipe (single smoothed Coulomb centre, Z=1, zero Dirichlet) on
    an n^3 mesh."""
    h = L / 2
    return O.Problem.build(p=p, n=(n, n, n), lo=(-h,) * 3, hi=(h,) * 3,
                           atoms=[(0.0, 0.0, 0.0, 1)])


def hanging(p=3, n=4):
    """A coarse n^3 mesh with a 2x2x1 block of cells refined once (hanging nodes)."""
    return O.Problem.build(p=p, n=(n, n, n), lo=(-8.,) * 3, hi=(8.,) * 3,
                           ref_lo=(1, 1, 1), ref_hi=(3, 3, 2),
                           atoms=[(0.3, -0.2, 0.1, 1)])


def enriched(p=2, n=4, L=24.0, nq_enr=6):
    """Two enriched atoms (1s, 2s, 2p Slater functions, support radius 6)."""
    h = L / 2
    atoms = [(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 6)]
    return O.Problem.build(p=p, n=(n, n, n), lo=(-h,) * 3, hi=(h,) * 3, atoms=atoms,
                           enrich=True, nq_enr=nq_enr)


def enriched_orth(p=2, n=6, L=24.0):
    """Two atoms with orthogonalized enrichment on a mesh where the supports
    have interior nodes (the KS driver test case)."""
    h = L / 2
    atoms = [(-1.5, 0.5, 0.2, 1), (2.0, -1.0, 0.7, 6)]
    return O.Problem.build(p=p, n=(n, n, n), lo=(-h,) * 3, hi=(h,) * 3, atoms=atoms,
                           enrich=True, orthogonalize=True)


def guess(P, nvec, seed=1):
    X = O.rng_uniform(seed, -1.0, 1.0, P.nDofs * nvec).reshape(P.nDofs, nvec)
    return P.p2c(X)


def quad_xyz(P):
    """Coordinates of the quadrature points of an oracle problem built with
    keep_basis, in its layout (cells in order, Gauss points x fastest within a
    cell): lo + h (x_q + 1) / 2 per direction."""
    off, _ = P.quadrature()
    lo, h = P.cell_boxes()
    out = np.zeros((int(off[-1]), 3))
    for e in range(P.nCells):
        nq = int(round((int(off[e + 1]) - int(off[e])) ** (1.0 / 3.0)))
        x, _ = O.gauss(nq)
        t = 0.5 * (x + 1.0)
        zz, yy, xx = np.meshgrid(t, t, t, indexing="ij")
        pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
        out[int(off[e]):int(off[e + 1])] = lo[e] + h[e] * pts
    return out


def assembled(P, key, d=None):
    """The raw assembled matrix sum_e Q_e^T A_e Q_e (no constraints) as a
    scipy sparse matrix, from the problem's cell matrices."""
    import scipy.sparse as sp
    d = d or P.arrays()
    rows, cols, vals = [], [], []
    off = moff = 0
    for n in d["cellNdofs"]:
        ids = d["cellMap"][off:off + n].astype(np.int64)
        A = d[key][moff:moff + n * n].reshape(n, n, order="F")
        rows.append(np.repeat(ids, n))
        cols.append(np.tile(ids, n))
        vals.append(A.ravel())
        off += n
        moff += n * n
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(P.nDofs, P.nDofs))
