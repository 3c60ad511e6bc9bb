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
