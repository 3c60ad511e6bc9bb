"""The benchmark workloads, restated with fixed seeds.

  c1  classical FE ChFSI: smoothed Coulomb centre (Z=1) in a 20 bohr cube,
      8^3 cells, p=3, zero Dirichlet, 10 states, block 10, degree 50
  c2  batched Poisson CG: 16^3 cells, p=3, 8 right-hand sides (4 Gaussians of
      width 0.3 bohr, charges 1-6, centres uniform in the cube), Jacobi PC,
      relative tolerance 1e-10
  c3  enriched FE ChFSI: 4 atoms (Z in {1,6}, min separation 2 bohr) in a
      24 bohr cube, 16^3 cells, p=3, 40 states, degree 40
  c4  large enriched ChFSI: 32 atoms (Z in {1,6,8}, min separation 1.8 bohr)
      in a 40 bohr cube, 32^3 cells, p=4, 200 states, block 64, degree 30
  c5  serving stream: 256 enriched eigenproblems (1-8 atoms, Z in {1,6},
      16-28 bohr cubes, 6^3-14^3 cells, p=3, 5-40 states), Poisson arrivals

Atoms of the enriched configs are drawn at least one enrichment support
radius (6 bohr) from the boundary, so the enrichment functions vanish on the
Dirichlet boundary.  Random numbers: SplitMix64 with the top 53 bits mapped
to [0, 1).
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_u01(seed, n, start=0):
    """Outputs start .. start+n-1 of SplitMix64(seed) mapped to [0, 1)."""
    with np.errstate(over="ignore"):
        t = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + t * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Stream:
    """Sequential SplitMix64 draws."""

    def __init__(self, seed):
        self.seed, self.pos = seed, 0

    def uniform(self, a=0.0, b=1.0, n=None):
        k = 1 if n is None else n
        u = splitmix_u01(self.seed, k, self.pos)
        self.pos += k
        v = a + (b - a) * u
        return float(v[0]) if n is None else v


def place_atoms(seed, n_atoms, lo, hi, zs, min_sep, margin=6.0):
    """Rejection sampling: uniform positions in [lo+margin, hi-margin]^3,
    Z uniform over `zs`, pairwise separation >= min_sep."""
    s = Stream(seed)
    atoms = []
    tries = 0
    while len(atoms) < n_atoms:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place atoms")
        x = [s.uniform(lo + margin, hi - margin) for _ in range(3)]
        z = zs[min(int(s.uniform() * len(zs)), len(zs) - 1)]
        if all(np.linalg.norm(np.subtract(x, a[:3])) >= min_sep for a in atoms):
            atoms.append((x[0], x[1], x[2], z))
    return atoms


def config(name):
    if name == "c1":
        return dict(p=3, n=(8, 8, 8), lo=(-10.,) * 3, hi=(10.,) * 3,
                    atoms=[(0.0, 0.0, 0.0, 1)], enrich=False, states=10, block=10,
                    degree=50, electrons=1)
    if name == "c2":
        return dict(p=3, n=(16, 16, 16), lo=(-10.,) * 3, hi=(10.,) * 3, atoms=[],
                    enrich=False, rhs=8, width=0.3, rel_tol=1e-10)
    if name == "c3":
        atoms = place_atoms(3, 4, -12.0, 12.0, (1, 6), 2.0)
        return dict(p=3, n=(16, 16, 16), lo=(-12.,) * 3, hi=(12.,) * 3, atoms=atoms,
                    enrich=True, states=40, block=40, degree=40,
                    electrons=sum(a[3] for a in atoms))
    if name == "c4":
        atoms = place_atoms(4, 32, -20.0, 20.0, (1, 6, 8), 1.8)
        return dict(p=4, n=(32, 32, 32), lo=(-20.,) * 3, hi=(20.,) * 3, atoms=atoms,
                    enrich=True, states=200, block=64, degree=30,
                    electrons=sum(a[3] for a in atoms))
    raise KeyError(name)


def serving_stream(seed=5, n_requests=256, rate=4.0, p=3):
    """The C5 request stream: Poisson arrivals at `rate` per second; each
    request has 1-8 atoms (Z in {1, 6}, minimum separation 2 bohr) in a
    16-28 bohr cube, a 6^3-14^3 mesh at order p, 5-40 wanted states (raised to
    hold 0.6 Ne + 2 when the atoms bring more electrons) and a 1e-6 residual
    tolerance.  All draws come from SplitMix64(seed)."""
    s = Stream(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += -np.log(1.0 - s.uniform()) / rate
        na = 1 + min(int(s.uniform() * 8), 7)
        L = 16.0 + 12.0 * s.uniform()
        nc = 6 + min(int(s.uniform() * 9), 8)
        k = 5 + min(int(s.uniform() * 36), 35)
        atoms = place_atoms(seed * 1000003 + i, na, -L / 2, L / 2, (1, 6), 2.0,
                            margin=min(6.0, L / 4))
        ne = sum(a[3] for a in atoms)
        k = min(40, max(k, int(np.ceil(0.6 * ne)) + 2))
        out.append(dict(p=p, n=(nc, nc, nc), lo=(-L / 2,) * 3, hi=(L / 2,) * 3, atoms=atoms,
                        enrich=True, states=k, electrons=float(ne), tol=1e-6, max_passes=100,
                        seed=seed * 7919 + i, arrival=t))
    return out


def poisson_rhs_charges(seed, n_centres=4, lo=-10.0, hi=10.0):
    """(centres, charges) of one Poisson right-hand side (config c2)."""
    s = Stream(seed)
    c = [[s.uniform(lo, hi) for _ in range(3)] for _ in range(n_centres)]
    q = [1 + min(int(s.uniform() * 6), 5) for _ in range(n_centres)]
    return np.array(c), np.array(q, dtype=float)
