"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OP_H, OP_M, OP_MINV, OP_MLUMP, OP_K = 0, 1, 2, 3, 4

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so with g++ (the oracle's own Makefile)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_error.restype = C.c_char_p
        L.orc_build.restype = C.c_void_p
        L.orc_build_ex.restype = C.c_void_p
        L.orc_build_affine.restype = C.c_void_p
        L.orc_from_arrays.restype = C.c_void_p
        L.orc_time_apply.restype = C.c_double
        for name in ("orc_free", "orc_sizes", "orc_get_u32", "orc_get_f64",
                     "orc_p2c", "orc_c2p", "orc_gather", "orc_scatter_add",
                     "orc_apply", "orc_chebyshev_filter", "orc_residual_filter",
                     "orc_cholgs", "orc_rr", "orc_chfsi", "orc_lanczos",
                     "orc_residual_norms", "orc_cg", "orc_time_apply", "orc_quadrature",
                     "orc_density", "orc_integrate", "orc_poisson", "orc_interpolate"):
            getattr(L, name).argtypes = None
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


_CHEBY_ORDER_LOOKUP = [(10, 6), (50, 9), (100, 12), (200, 16), (300, 19), (500, 24),
                       (750, 30), (1000, 39), (1500, 50), (2000, 53), (3000, 57),
                       (4000, 62), (5000, 69), (9000, 77), (14000, 104), (20000, 119),
                       (30000, 162), (50000, 300), (80000, 450), (100000, 550),
                       (200000, 700), (500000, 1000)]


def cheby_degree(upper):
    """CHEBY_ORDER_LOOKUP (src/ksdft/Defaults.cpp:51-58) with the lower_bound
    lookup on a size_type key (src/ksdft/KohnShamEigenSolver.t.cpp:36-46)."""
    key = int(upper) & 0xFFFFFFFF
    for k, v in _CHEBY_ORDER_LOOKUP:
        if k >= key:
            return v
    return 1250


K_BOLTZMANN_HARTREE = 3.166811429e-06  # src/ksdft/Defaults.cpp:69


def fermi_dirac(e, mu, kb, T):
    """src/ksdft/FractionalOccupancyFunction.cpp:9-18"""
    f = (e - mu) / (kb * T)
    return np.exp(-f) / (1.0 + np.exp(-f)) if f >= 0 else 1.0 / (1.0 + np.exp(f))


def fermi_dirac_der(e, mu, kb, T):
    """src/ksdft/FractionalOccupancyFunction.cpp:20-33"""
    f = (e - mu) / (kb * T)
    beta = 1.0 / (kb * T)
    if f >= 0:
        return beta * np.exp(-f) / (1.0 + np.exp(-f)) / (1.0 + np.exp(-f))
    return beta * np.exp(f) / (1.0 + np.exp(f)) / (1.0 + np.exp(f))


def fermi_level(eig, ne, T, tol):
    """NewtonRaphsonSolver on sum_i 2 f_i - Ne (src/linearAlgebra/
    NewtonRaphsonSolver.t.cpp, FractionalOccupancyFunction.cpp:49-71), started at
    eig[ceil(Ne/2) - 1] (src/ksdft/KohnShamEigenSolver.t.cpp:396-425)."""
    i0 = int(np.ceil(ne / 2.0))
    i0 = min(i0 - 1, len(eig) - 1) if i0 else 0
    x = float(eig[i0])
    for _ in range(100000):
        g = -ne + sum(2 * fermi_dirac(e, x, K_BOLTZMANN_HARTREE, T) for e in eig)
        gp = sum(2 * fermi_dirac_der(e, x, K_BOLTZMANN_HARTREE, T) for e in eig)
        if gp == 0.0:
            break
        x1 = x - g / gp
        if abs(x1 - x) < tol:
            return x1
        x = x1
    return x


def rng_uniform(seed, a, b, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_rng_uniform(C.c_uint64(seed), C.c_double(a), C.c_double(b),
                          C.c_uint64(n), _p(out))
    return out


def gauss(n):
    x = np.empty(n); w = np.empty(n)
    lib().orc_gauss(C.c_int(n), _p(x), _p(w))
    return x, w


def gll(n):
    x = np.empty(n); w = np.empty(n)
    lib().orc_gll(C.c_int(n), _p(x), _p(w))
    return x, w


def gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
    lib().orc_gemm(C.c_char(ta.encode()), C.c_char(tb.encode()), C.c_int(m),
                   C.c_int(n), C.c_int(k), C.c_double(alpha), _p(A), C.c_int(lda),
                   _p(B), C.c_int(ldb), C.c_double(beta), _p(Cm), C.c_int(ldc))


def syev(A):
    n = A.shape[0]
    a = np.array(A, dtype=np.float64, order="F")
    w = np.empty(n); Z = np.empty((n, n), order="F")
    lib().orc_syev(C.c_int(n), _p(a), _p(w), _p(Z))
    return w, Z


def enrichment(atom, Z, k, x, rcut=4.0, smooth=2.0):
    atom = np.ascontiguousarray(atom, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    val = C.c_double(0.0)
    grad = np.zeros(3)
    lib().orc_enrichment(_p(atom), C.c_int(Z), C.c_int(k), C.c_double(rcut),
                         C.c_double(smooth), _p(x), C.byref(val), _p(grad))
    return val.value, grad


class Problem:
    """An FE problem built by the oracle (mesh, maps, constraints, cell
    matrices), or wrapped around caller-provided arrays."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("oracle: " + lib().orc_error().decode())
        self.h = C.c_void_p(handle)
        s = np.zeros(8, dtype=np.uint64)
        lib().orc_sizes(self.h, _p(s))
        (self.nCells, self.nDofs, self.nClassical, self.nEnriched, self.mapLen,
         self.matLen, self.nRows, self.nEntries) = [int(v) for v in s]

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_free(self.h)
        except Exception:
            pass

    @classmethod
    def build(cls, p=3, n=(8, 8, 8), lo=(-10., -10., -10.), hi=(10., 10., 10.),
              ref_lo=(0, 0, 0), ref_hi=(0, 0, 0), nq=0, nq_enr=0, atoms=(),
              v_smear=0.5, enrich=False, rcut=4.0, smooth=2.0, orthogonalize=False,
              keep_basis=False):
        n = np.array(n, dtype=np.int32)
        lo = np.array(lo, dtype=np.float64); hi = np.array(hi, dtype=np.float64)
        rl = np.array(ref_lo, dtype=np.int32); rh = np.array(ref_hi, dtype=np.int32)
        xyz = np.array([a[:3] for a in atoms], dtype=np.float64).reshape(-1)
        Z = np.array([int(a[3]) for a in atoms], dtype=np.int32)
        if xyz.size == 0:
            xyz = np.zeros(3); Z = np.zeros(1, dtype=np.int32)
        h = lib().orc_build_ex(C.c_int(p), _p(n), _p(lo), _p(hi), _p(rl), _p(rh),
                               C.c_int(nq), C.c_int(nq_enr), C.c_int(len(atoms)),
                               _p(xyz), _p(Z), C.c_double(v_smear),
                               C.c_int(1 if enrich else 0), C.c_double(rcut),
                               C.c_double(smooth), C.c_int(1 if orthogonalize else 0),
                               C.c_int(1 if keep_basis else 0))
        P = cls(h)
        P.p = p
        return P

    @classmethod
    def build_affine(cls, v, p=3, n=(2, 2, 2), lo=(-1., -1., -1.), hi=(1., 1., 1.), nq=0):
        """A classical mesh whose H_e = K_e / 2 + int N_i V N_j carries the
        affine test potential V(x) = v[0] + v[1:4] . x."""
        n = np.array(n, dtype=np.int32)
        lo = np.array(lo, dtype=np.float64); hi = np.array(hi, dtype=np.float64)
        v = np.array(v, dtype=np.float64)
        h = lib().orc_build_affine(C.c_int(p), _p(n), _p(lo), _p(hi), C.c_int(nq), _p(v))
        P = cls(h)
        P.p = p
        return P

    @classmethod
    def from_arrays(cls, cellNdofs, cellMap, nDofs, nClassical, H=None, M=None,
                    K=None, rows=None, rowSizes=None, rowOffsets=None, cols=None,
                    weights=None, inhom=None, dinv=None, mgll=None, p=3):
        u = lambda a: np.ascontiguousarray(a if a is not None else [], dtype=np.uint32)
        f = lambda a: (np.ascontiguousarray(a, dtype=np.float64) if a is not None else None)
        cellNdofs, cellMap = u(cellNdofs), u(cellMap)
        rows, rowSizes, rowOffsets, cols = u(rows), u(rowSizes), u(rowOffsets), u(cols)
        weights = f(weights) if weights is not None else np.zeros(0)
        inhom = f(inhom) if inhom is not None else np.zeros(len(rows))
        H, M, K, dinv, mgll = f(H), f(M), f(K), f(dinv), f(mgll)
        keep = (cellNdofs, cellMap, rows, rowSizes, rowOffsets, cols, weights,
                inhom, H, M, K, dinv, mgll)
        nul = lambda a: _p(a) if a is not None else None
        h = lib().orc_from_arrays(
            C.c_uint32(len(cellNdofs)), C.c_uint32(nDofs), C.c_uint32(nClassical),
            _p(cellNdofs), _p(cellMap), nul(H), nul(M), nul(K),
            C.c_uint32(len(rows)), _p(rows), _p(rowSizes), _p(rowOffsets),
            _p(cols), _p(weights), _p(inhom), nul(dinv), nul(mgll), C.c_int(p))
        P = cls(h)
        P.p = p
        P._keep = keep
        return P

    # ------------------------------------------------------------ arrays
    def _u32(self, which, n):
        out = np.zeros(n, dtype=np.uint32)
        if n:
            lib().orc_get_u32(self.h, C.c_int(which), _p(out))
        return out

    def _f64(self, which, n):
        out = np.zeros(n, dtype=np.float64)
        if n:
            lib().orc_get_f64(self.h, C.c_int(which), _p(out))
        return out

    def arrays(self):
        d = {}
        d["cellNdofs"] = self._u32(0, self.nCells)
        d["cellMap"] = self._u32(1, self.mapLen)
        d["rows"] = self._u32(2, self.nRows)
        d["rowSizes"] = self._u32(3, self.nRows)
        d["rowOffsets"] = self._u32(4, self.nRows)
        d["cols"] = self._u32(5, self.nEntries)
        d["H"] = self._f64(0, self.matLen)
        d["M"] = self._f64(1, self.matLen)
        d["K"] = self._f64(2, self.matLen)
        d["weights"] = self._f64(3, self.nEntries)
        d["inhom"] = self._f64(4, self.nRows)
        d["dinv"] = self._f64(5, self.nDofs)
        d["mgll"] = self._f64(6, self.nDofs)
        d["enrMinv"] = self._f64(8, self.nEnriched * self.nEnriched)
        return d

    def alpha(self):
        return self._f64(11, self.nEnriched * self.nClassical).reshape(self.nEnriched, -1)

    def node_xyz(self):
        return self._f64(7, 3 * self.nClassical).reshape(-1, 3)

    def is_hanging(self):
        return self._u32(6, self.nClassical).astype(bool)

    def is_dirichlet(self):
        return self._u32(7, self.nClassical).astype(bool)

    def cell_level(self):
        return self._u32(8, self.nCells)

    def cell_boxes(self):
        return (self._f64(9, 3 * self.nCells).reshape(-1, 3),
                self._f64(10, 3 * self.nCells).reshape(-1, 3))

    # -------------------------------------------------------- operations
    def p2c(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        B = X.shape[1] if X.ndim == 2 else 1
        lib().orc_p2c(self.h, _p(X), C.c_uint32(B))
        return X

    def c2p(self, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        B = Y.shape[1] if Y.ndim == 2 else 1
        lib().orc_c2p(self.h, _p(Y), C.c_uint32(B))
        return Y

    def gather(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        B = X.shape[1]
        out = np.zeros((self.mapLen, B))
        lib().orc_gather(self.h, _p(X), C.c_uint32(B), _p(out))
        return out

    def apply(self, kind, X):
        """Y = Op X; X is modified in place (P2C), as in the reference."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        B = X.shape[1]
        Y = np.zeros_like(X)
        rc = lib().orc_apply(self.h, C.c_int(kind), _p(X), _p(Y), C.c_uint32(B))
        if rc != 0:
            raise RuntimeError(lib().orc_error().decode())
        return Y

    def chebyshev_filter(self, X, m, a, b, c):
        X = np.array(X, dtype=np.float64, order="C")
        Y = np.zeros_like(X)
        lib().orc_chebyshev_filter(self.h, _p(X), C.c_uint32(X.shape[1]), C.c_int(m),
                                   C.c_double(a), C.c_double(b), C.c_double(c), _p(Y))
        return Y

    def residual_filter(self, lam, X, m, a, b, c):
        X = np.array(X, dtype=np.float64, order="C")
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        Y = np.zeros_like(X)
        lib().orc_residual_filter(self.h, _p(lam), _p(X), C.c_uint32(X.shape[1]),
                                  C.c_int(m), C.c_double(a), C.c_double(b),
                                  C.c_double(c), _p(Y))
        return Y

    def cholgs(self, X):
        X = np.array(X, dtype=np.float64, order="C")
        Xo = np.zeros_like(X)
        rc = lib().orc_cholgs(self.h, _p(X), C.c_uint32(X.shape[1]), _p(Xo))
        return rc, Xo

    def rayleigh_ritz(self, X):
        X = np.array(X, dtype=np.float64, order="C")
        Xo = np.zeros_like(X)
        eig = np.zeros(X.shape[1])
        rc = lib().orc_rr(self.h, _p(X), C.c_uint32(X.shape[1]), _p(eig), _p(Xo))
        return rc, eig, Xo

    def chfsi(self, a, b, c, m, guess, eig, residual=True, batch=0):
        guess = np.array(guess, dtype=np.float64, order="C")
        eig = np.array(eig, dtype=np.float64)
        X = np.zeros_like(guess)
        n = guess.shape[1]
        rc = lib().orc_chfsi(self.h, C.c_double(a), C.c_double(b), C.c_double(c),
                             C.c_int(m), C.c_int(1 if residual else 0),
                             C.c_uint32(batch), _p(eig), _p(guess), _p(X),
                             C.c_uint32(n))
        return rc, eig, X, guess

    def lanczos(self, guess, max_krylov=20):
        g = np.ascontiguousarray(guess, dtype=np.float64)
        out = np.zeros(3)
        lib().orc_lanczos(self.h, _p(g), C.c_int(max_krylov), _p(out))
        return out  # lo, hi, beta_last

    def residual_norms(self, X, eig):
        X = np.ascontiguousarray(X, dtype=np.float64)
        eig = np.ascontiguousarray(eig, dtype=np.float64)
        res = np.zeros(X.shape[1])
        lib().orc_residual_norms(self.h, _p(X), _p(eig), C.c_uint32(X.shape[1]), _p(res))
        return res

    def cg(self, kind, diag_inv, b, x0=None, abs_tol=1e-10, rel_tol=1e-10,
           div_tol=1e10, max_iter=10000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, order="C")
        d = np.ascontiguousarray(diag_inv, dtype=np.float64)
        it = C.c_uint32(0)
        rc = lib().orc_cg(self.h, C.c_int(kind), _p(d), _p(b), _p(x),
                          C.c_uint32(b.shape[1]), C.c_double(abs_tol),
                          C.c_double(rel_tol), C.c_double(div_tol),
                          C.c_uint32(max_iter), C.byref(it))
        return rc, x, it.value

    def ks_eigensolve(self, X0, lanczos_guess, degree=0, tol=1e-6, max_pass=10,
                      residual=True, batch=0, max_krylov=20, num_electrons=0.0,
                      temperature=500.0, fermi_tol=1e-10, occ_tol=1e-3):
        """KohnShamEigenSolver::solve (src/ksdft/KohnShamEigenSolver.t.cpp:214-536)
        composed from the oracle's pieces: Lanczos bounds and the upper-bound
        heuristic (:264-297), degree lookup (:36-46), one degree-1 plain pass
        before residual-filter passes (src/ksdft/KohnShamDFT.t.cpp:462-487),
        then ChFSI passes until every occupied state (Fermi-Dirac occupancy
        above occ_tol at the Newton-Raphson Fermi level, :396-460) has residual
        norm within tol.  num_electrons == 0 counts every state as occupied."""
        n = X0.shape[1]
        N = self.nDofs
        lo, hi, beta = self.lanczos(lanczos_guess, max_krylov)
        r = beta / 10.0
        a = lo
        b = (hi + r - lo) * ((n * 200.0) / N) + lo
        if b >= hi + r:
            b = (hi + r + lo) * 0.5
        c = hi + r
        m = degree or cheby_degree(c)
        eig = np.zeros(n)
        G = np.array(X0, dtype=np.float64, order="C")
        if residual:
            rc, eig, X, G = self.chfsi(a, b, c, 1, G, eig, residual=False, batch=batch)
            G = X.copy()
        passes, res, rc, mu, nocc = 0, None, 6, 0.0, n
        for it in range(max_pass):
            rc, eig, X, G = self.chfsi(a, b, c, m, G, eig, residual=residual, batch=batch)
            res = self.residual_norms(X, eig)
            G = X.copy()
            passes = it + 1
            if rc != 0:
                rc = 7
                break
            if num_electrons > 0:
                mu = fermi_level(eig, num_electrons, temperature, fermi_tol)
                occ = np.array([fermi_dirac(e, mu, K_BOLTZMANN_HARTREE, temperature)
                                for e in eig]) > occ_tol
            else:
                occ = np.ones(n, dtype=bool)
            nocc = int(occ.sum())
            if np.all(res[occ] <= tol):
                rc = 0
                break
            a, b = eig[0], eig[-1]
            rc = 6
        return dict(rc=rc, eig=eig, X=X, passes=passes, res=res, fermi=mu, nocc=nocc,
                    bounds=(lo, hi, beta), a=a, b=b, c=c, degree=m)

    def quadrature(self):
        """(quadOff [nCells + 1], JxW [quadOff[-1]]) of a problem built with
        keep_basis."""
        off = np.zeros(self.nCells + 1, dtype=np.uint64)
        if lib().orc_quadrature(self.h, _p(off), None) != 0:
            raise RuntimeError("problem built without keep_basis")
        w = np.zeros(int(off[-1]))
        lib().orc_quadrature(self.h, None, _p(w))
        return off, w

    def interpolate(self, X):
        """FEBasisOperations::interpolate of nodal vectors X (nDofs x B) to
        the quadrature points (layout quadrature()), no constraint step."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        X2 = X.reshape(X.shape[0], -1)
        off, _ = self.quadrature()
        out = np.zeros((int(off[-1]), X2.shape[1]))
        if lib().orc_interpolate(self.h, _p(X2), C.c_uint32(X2.shape[1]), _p(out)) != 0:
            raise RuntimeError(lib().orc_error().decode())
        return out

    def integrate(self, f, c2p=True):
        """FEBasisOperations::integrateWithBasisValues (src/basis/
        FEBasisOperations.t.cpp:1846-2190) of quadrature values f
        (points x components, layout quadrature()) on a problem built with
        keep_basis; c2p: the constraints' child->parent step at the end."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        f2 = f.reshape(f.shape[0], -1)
        out = np.zeros((self.nDofs, f2.shape[1]))
        if lib().orc_integrate(self.h, _p(f2), C.c_uint32(f2.shape[1]), _p(out),
                               C.c_int(1 if c2p else 0)) != 0:
            raise RuntimeError(lib().orc_error().decode())
        return out

    def poisson(self, inhom, f):
        """PoissonLinearSolverFunctionFE's set-up with this problem's
        constraints as the homogeneous ones and `inhom` (one value per
        constraint row) as the field's Dirichlet data: returns (b, Jacobi
        diagonal, u_D) with b = integrate(f) - C2P(K u_D)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        f2 = f.reshape(f.shape[0], -1)
        B = f2.shape[1]
        inh = np.ascontiguousarray(inhom, dtype=np.float64)
        assert inh.size == self.nRows
        b = np.zeros((self.nDofs, B)); uD = np.zeros((self.nDofs, B))
        diag = np.zeros(self.nDofs)
        if lib().orc_poisson(self.h, _p(inh), _p(f2), C.c_uint32(B), _p(b), _p(diag),
                             _p(uD)) != 0:
            raise RuntimeError(lib().orc_error().decode())
        return b, diag, uD

    def density(self, X, occupation):
        """rho at every quadrature point (cell-major): sum_i 2 f_i psi_i^2."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        occ = np.ascontiguousarray(occupation, dtype=np.float64)
        off, _ = self.quadrature()
        rho = np.zeros(int(off[-1]))
        rc = lib().orc_density(self.h, _p(X), C.c_uint32(X.shape[1]), _p(occ),
                               _p(rho))
        if rc != 0:
            raise RuntimeError(lib().orc_error().decode())
        return rho

    def time_apply(self, kind, X, reps=3, threads=1):
        """Median seconds of the threaded CPU apply: (total, cell part, O(N B)
        vector part, Y)."""
        X = np.array(X, dtype=np.float64, order="C")
        Y = np.zeros_like(X)
        tc, tv = C.c_double(0), C.c_double(0)
        t = lib().orc_time_apply(self.h, C.c_int(kind), _p(X), _p(Y),
                                 C.c_uint32(X.shape[1]), C.c_int(reps), C.c_int(threads),
                                 C.byref(tc), C.byref(tv))
        return t, tc.value, tv.value, Y
