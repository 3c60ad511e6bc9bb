"""Python front end of libdfx, the MI355X-native Kohn-Sham eigensolver hot path.

Thin ctypes wrappers over the C ABI in include/dfx.h.  Class names follow the
reference's operator / solver interfaces (dft-efe):

  CellOperator                  KohnShamOperatorContextFE, CFEOverlapOperatorContext,
                                LaplaceOperatorContextFE   (cell-wise apply)
  DiagonalOperator              CFEOverlapInverseOpContextGLL,
                                OrthoEFEOverlapInverseOpContextGLL, PreconditionerJacobi
  ChebyshevFilteredEigenSolver  linearAlgebra::ChebyshevFilteredEigenSolver
  KohnShamEigenSolver           ksdft::KohnShamEigenSolver
  CGLinearSolver                linearAlgebra::CGLinearSolver

The compute path is the HIP library only: if libdfx.so is missing, importing
this module raises instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DFX_LIBRARY names another build of libdfx.so (e.g. one compiled with other
# tuning macros); the default is the in-tree build next to this file.
LIB_PATH = os.environ.get("DFX_LIBRARY") or os.path.join(_HERE, "libdfx.so")

F64, F32 = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
OP_KEEP_FP32 = 0x1
PC_NONE, PC_JACOBI = 0, 1   # linearAlgebra::PreconditionerType
NOALLOC_SCRATCH_BYTES = 64 << 20   # DFX_NOALLOC_SCRATCH_BYTES

EIG_CODES = {0: "SUCCESS", 1: "LAPACK_ERROR", 2: "LANCZOS_BETA_ZERO",
             3: "LANCZOS_SUBSPACE_INSUFFICIENT", 4: "CHFSI_ORTHONORMALIZATION_ERROR",
             5: "CHFSI_RAYLEIGH_RITZ_ERROR", 6: "KS_MAX_PASS_ERROR",
             7: "KS_CHFSI_ERROR", 8: "KS_LANCZOS_ERROR"}
LS_CODES = {0: "SUCCESS", 1: "FAILED_TO_CONVERGE", 2: "RESIDUAL_DIVERGENCE",
            3: "DIVISON_BY_ZERO", 4: "OTHER_ERROR"}


class DfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dfx error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libdfx.so not found at {LIB_PATH}; build it with "
            "`python dft-efe_amd/build.py` (or __graft_entry__.build())")
    return C.CDLL(LIB_PATH)


_lib = _load()
_vp = C.c_void_p
_u32 = C.c_uint32
_dbl = C.c_double
_pp = C.POINTER(C.c_void_p)


def _sig(name, *args, res=C.c_int):
    f = getattr(_lib, name)
    f.argtypes = list(args)
    f.restype = res
    return f


_sig("dfx_ctx_create", C.c_int, _pp)
_sig("dfx_ctx_destroy", _vp)
_sig("dfx_ctx_last_error", _vp, res=C.c_char_p)
_sig("dfx_ctx_synchronize", _vp)
_sig("dfx_ctx_stream", _vp, _pp)
_sig("dfx_ctx_graph_stats", _vp, _vp, _vp, _vp)
_sig("dfx_mv_create", _vp, _u32, _u32, C.c_int, _pp)
_sig("dfx_mv_wrap", _vp, _vp, _u32, _u32, C.c_int, _pp)
_sig("dfx_mv_destroy", _vp)
_sig("dfx_mv_upload", _vp, _vp)
_sig("dfx_mv_download", _vp, _vp)
_sig("dfx_mv_device_ptr", _vp, _pp)
_sig("dfx_mv_set", _vp, _dbl)
_sig("dfx_mv_fill_random", _vp, C.c_uint64, _dbl, _dbl)
_sig("dfx_mv_copy", _vp, _vp)
_sig("dfx_mv_axpby", _vp, _dbl, _vp, _dbl, _vp)
_sig("dfx_mv_axpby_blocked", _vp, _dbl, _vp, _vp, _dbl, _vp, _vp)
_sig("dfx_mv_l2norms", _vp, _vp)
_sig("dfx_mv_dots", _vp, _vp, _vp)
_sig("dfx_constraints_create", _vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _pp)
_sig("dfx_constraints_destroy", _vp)
_sig("dfx_constraints_distribute_parent_to_child", _vp, _vp)
_sig("dfx_constraints_distribute_child_to_parent", _vp, _vp)
_sig("dfx_constraints_set_constrained_nodes", _vp, _vp, _dbl)
_sig("dfx_constraints_sizes", _vp, _vp, _vp, _vp)
_sig("dfx_constraints_export", _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("dfx_op_create_cellwise", _vp, _u32, _vp, _vp, _vp, _u32, _u32, _vp, _u32, _pp)
_sig("dfx_op_update_cells", _vp, _vp)
_sig("dfx_op_create_diagonal", _vp, _u32, _vp, _u32, _vp, _vp, _pp)
_sig("dfx_op_apply", _vp, _vp, _vp)
_sig("dfx_op_apply_after_inverse", _vp, _vp, _vp, _vp)
_sig("dfx_op_apply_noalloc", _vp, _vp, _vp)
_sig("dfx_problem_integrate", _vp, _vp, _vp, _vp)
_sig("dfx_poisson_create", _vp, _vp, _vp, _vp, C.c_int, _pp)
_sig("dfx_poisson_reinit", _vp, _vp, _vp)
_sig("dfx_poisson_get_ax", _vp, _pp)
_sig("dfx_poisson_get_pc", _vp, _pp)
_sig("dfx_poisson_get_rhs", _vp, _vp)
_sig("dfx_poisson_get_initial_guess", _vp, _vp)
_sig("dfx_poisson_set_solution", _vp, _vp)
_sig("dfx_poisson_get_solution", _vp, _vp)
_sig("dfx_poisson_solve", _vp, _u32, _dbl, _dbl, _dbl, _vp, _vp, _vp)
_sig("dfx_poisson_destroy", _vp)
_sig("dfx_op_destroy", _vp)
_sig("dfx_op_get_info", _vp, _vp)
_sig("dfx_op_export_map", _vp, _vp, _vp)
_sig("dfx_op_gather", _vp, _vp, _vp)
_sig("dfx_op_scatter_add", _vp, _vp, _vp)
_sig("dfx_chfsi_create", _vp, _vp, _vp, _pp)
_sig("dfx_chfsi_reinit", _vp, _dbl, _dbl, _dbl, _u32, _dbl, _vp)
_sig("dfx_chfsi_solve", _vp, _vp, _vp, _vp, C.c_int, _vp, _vp)
_sig("dfx_chfsi_timings", _vp, _vp)
_sig("dfx_chfsi_get_filtered_subspace", _vp, C.c_int, _vp)
_sig("dfx_chfsi_destroy", _vp)
_sig("dfx_cholesky_gram_schmidt", _vp, _vp, _vp, _vp)
_sig("dfx_rayleigh_ritz", _vp, _vp, _vp, _vp, _vp)
_sig("dfx_lanczos_extreme", _vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp)
_sig("dfx_ks_eigensolve", _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("dfx_eigen_residual_norms", _vp, _vp, _vp, _vp, _vp, _vp)
_sig("dfx_chfsi_residual_norms", _vp, _vp, _vp, _vp, _vp, _vp)
_sig("dfx_cg_solve", _vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl, _u32, _vp, _vp)
_sig("dfx_op_set_profiling", _vp, C.c_int)
_sig("dfx_op_autotune", _vp, C.c_int, _u32, _u32, _u32, _vp, _vp, _vp)
_sig("dfx_op_get_profile", _vp, _vp, C.c_int)
_sig("dfx_problem_build", _vp, _vp, _pp)
_sig("dfx_problem_sizes", _vp, _vp)
_sig("dfx_problem_get", _vp, C.c_int, _vp)
_sig("dfx_problem_create_operators", _vp, _u32, _pp, _pp, _pp, _pp, _pp, _pp)
_sig("dfx_problem_density", _vp, _vp, _vp, _vp)
_sig("dfx_problem_destroy", _vp)
_sig("dfx_atom_read", C.c_char_p, _vp, _u32, _pp)
_sig("dfx_atom_read_upf", C.c_char_p, _vp, _u32, _pp)
_sig("dfx_atom_function_knots", _vp, C.c_char_p, _u32, _vp, _vp, _vp, _vp, _vp)
_sig("dfx_atom_metadata", _vp, C.c_char_p, C.c_char_p, C.c_uint64, _vp)
_sig("dfx_atom_radial_grid", _vp, _vp, _vp)
_sig("dfx_atom_num_functions", _vp, C.c_char_p, _vp)
_sig("dfx_atom_function", _vp, C.c_char_p, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("dfx_atom_evaluate", _vp, _vp, C.c_char_p, _u32, _vp, _vp, _vp, _vp)
_sig("dfx_atom_destroy", _vp)
_sig("dfx_request_footprint", _vp, _vp)
_sig("dfx_serve", _vp, _u32, _vp, _vp, _vp)


def exported_symbols():
    """Names this front end binds (all must exist in libdfx.so)."""
    return sorted(n for n in dir(_lib) if n.startswith("dfx_"))


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def _u32a(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64a(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One GPU and one HIP stream (LinAlgOpContext<DEVICE>)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = _lib.dfx_ctx_create(device, C.byref(h))
        if rc != 0:
            raise DfxError(rc, _lib.dfx_ctx_last_error(None).decode())
        self.h = h
        self.device = device

    def check(self, rc, allow_positive=False):
        if rc < 0 or (rc > 0 and not allow_positive):
            raise DfxError(rc, _lib.dfx_ctx_last_error(self.h).decode())
        return rc

    def synchronize(self):
        self.check(_lib.dfx_ctx_synchronize(self.h))

    @property
    def stream(self):
        s = C.c_void_p()
        _lib.dfx_ctx_stream(self.h, C.byref(s))
        return s.value or 0

    def graph_stats(self):
        """(captures, replays, cached graphs) of the captured degree loops."""
        c, r, n = C.c_uint64(), C.c_uint64(), C.c_uint32()
        self.check(_lib.dfx_ctx_graph_stats(self.h, C.byref(c), C.byref(r), C.byref(n)))
        return int(c.value), int(r.value), int(n.value)

    def close(self):
        if getattr(self, "h", None):
            _lib.dfx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiVector:
    """N x B device multivector, row-major, vector index fastest."""

    def __init__(self, ctx, N, B, dtype=F64):
        self.ctx, self.N, self.B, self.dtype = ctx, int(N), int(B), dtype
        h = C.c_void_p()
        ctx.check(_lib.dfx_mv_create(ctx.h, self.N, self.B, dtype, C.byref(h)))
        self.h = h

    @classmethod
    def from_numpy(cls, ctx, a, dtype=None):
        a = np.asarray(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        if dtype is None:
            dtype = F32 if a.dtype == np.float32 else F64
        mv = cls(ctx, a.shape[0], a.shape[1], dtype)
        mv.upload(a)
        return mv

    @property
    def np_dtype(self):
        return np.float64 if self.dtype == F64 else np.float32

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.np_dtype).reshape(self.N, self.B)
        self.ctx.check(_lib.dfx_mv_upload(self.h, _ptr(a)))

    def to_numpy(self):
        out = np.empty((self.N, self.B), dtype=self.np_dtype)
        self.ctx.check(_lib.dfx_mv_download(self.h, _ptr(out)))
        return out

    def set(self, v):
        self.ctx.check(_lib.dfx_mv_set(self.h, float(v)))

    def fill_random(self, seed, a=-1.0, b=1.0):
        """SplitMix64 uniform [a, b) fill, on the device."""
        self.ctx.check(_lib.dfx_mv_fill_random(self.h, C.c_uint64(seed), float(a), float(b)))

    def copy_from(self, other):
        self.ctx.check(_lib.dfx_mv_copy(self.h, other.h))

    def axpby(self, a, X, b, Y):
        """self = a X + b Y"""
        self.ctx.check(_lib.dfx_mv_axpby(self.h, float(a), X.h, float(b), Y.h))

    def axpby_blocked(self, a1, a, X, b1, b, Y):
        a = _f64a(a); b = _f64a(b)
        self.ctx.check(_lib.dfx_mv_axpby_blocked(self.h, float(a1), _ptr(a), X.h,
                                                 float(b1), _ptr(b), Y.h))

    def l2norms(self):
        out = np.zeros(self.B)
        self.ctx.check(_lib.dfx_mv_l2norms(self.h, _ptr(out)))
        return out

    def dots(self, Y):
        out = np.zeros(self.B)
        self.ctx.check(_lib.dfx_mv_dots(self.h, Y.h, _ptr(out)))
        return out

    @property
    def device_ptr(self):
        p = C.c_void_p()
        _lib.dfx_mv_device_ptr(self.h, C.byref(p))
        return p.value

    def close(self):
        if getattr(self, "h", None):
            _lib.dfx_mv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Constraints:
    """Hanging-node + Dirichlet constraints (basis::ConstraintsLocal), flattened."""

    def __init__(self, ctx, nDofs, rows, rowSizes, rowOffsets, cols, weights, inhom=None):
        self.ctx = ctx
        rows, rowSizes, rowOffsets = _u32a(rows), _u32a(rowSizes), _u32a(rowOffsets)
        cols, weights = _u32a(cols), _f64a(weights)
        inhom = np.zeros(len(rows)) if inhom is None else _f64a(inhom)
        self._keep = (rows, rowSizes, rowOffsets, cols, weights, inhom)
        h = C.c_void_p()
        ctx.check(_lib.dfx_constraints_create(
            ctx.h, int(nDofs), len(rows), _ptr(rows), _ptr(rowSizes), _ptr(rowOffsets),
            _ptr(cols), _ptr(weights), _ptr(inhom), C.byref(h)))
        self.h = h
        self.nDofs = int(nDofs)

    def distribute_parent_to_child(self, X):
        self.ctx.check(_lib.dfx_constraints_distribute_parent_to_child(self.h, X.h))

    def distribute_child_to_parent(self, Y):
        self.ctx.check(_lib.dfx_constraints_distribute_child_to_parent(self.h, Y.h))

    def set_constrained_nodes(self, X, v):
        self.ctx.check(_lib.dfx_constraints_set_constrained_nodes(self.h, X.h, float(v)))

    def export(self):
        nr, ne, npar = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _lib.dfx_constraints_sizes(self.h, C.byref(nr), C.byref(ne), C.byref(npar))
        nr, ne, npar = nr.value, ne.value, npar.value
        d = dict(rows=np.zeros(nr, np.uint32), rowOffsets=np.zeros(nr + 1, np.uint32),
                 cols=np.zeros(ne, np.uint32), parents=np.zeros(npar, np.uint32),
                 parentOffsets=np.zeros(npar + 1, np.uint32),
                 parentRows=np.zeros(ne, np.uint32), parentEntries=np.zeros(ne, np.uint32))
        self.ctx.check(_lib.dfx_constraints_export(
            self.h, _ptr(d["rows"]), _ptr(d["rowOffsets"]), _ptr(d["cols"]),
            _ptr(d["parents"]), _ptr(d["parentOffsets"]), _ptr(d["parentRows"]),
            _ptr(d["parentEntries"])))
        return d

    def close(self):
        if getattr(self, "h", None):
            _lib.dfx_constraints_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _OpInfo(C.Structure):
    _fields_ = [("nCells", C.c_uint32), ("nDofs", C.c_uint32), ("nClassical", C.c_uint32),
                ("nColors", C.c_uint32), ("maxCellDofs", C.c_uint32),
                ("sumNe", C.c_uint64), ("sumNe2", C.c_uint64),
                ("matBytesF64", C.c_uint64), ("matBytesF32", C.c_uint64)]


class _OpProfile(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("kernelMs", C.c_double), ("flops", C.c_double),
                ("bytes", C.c_double)]


class _Operator:
    def apply(self, X, Y):
        """Y = Op X (X's constrained rows are set by parent->child, as in the
        reference OperatorContext::apply)."""
        self.ctx.check(_lib.dfx_op_apply(self.h, X.h, Y.h))

    def apply_noalloc(self, X, Y):
        """apply() with no device allocation (library-owned per-GPU scratch of
        NOALLOC_SCRATCH_BYTES for the enriched partial rows; larger applies
        raise).  Overlapping calls on one GPU race."""
        self.ctx.check(_lib.dfx_op_apply_noalloc(self.h, X.h, Y.h))

    def apply_after_inverse(self, BInv, X, Y):
        """Y = Op (BInv X), the residual filter's BInv.apply then A.apply
        (BInv folded into this operator's gather when eligible)."""
        self.ctx.check(_lib.dfx_op_apply_after_inverse(self.h, BInv.h, X.h, Y.h))

    def info(self):
        i = _OpInfo()
        _lib.dfx_op_get_info(self.h, C.byref(i))
        return {k: getattr(i, k) for k, _ in _OpInfo._fields_}

    def set_profiling(self, enable=True, precision=None, every=1):
        """Time this operator's cell-apply kernel launches; precision F64 / F32
        restricts to launches of that value type; every = k times the colour
        steps of every k-th such apply only."""
        mode = 0 if not enable else (3 if precision is None else (1 if precision == F64 else 2))
        if mode and int(every) > 1:
            mode |= (int(every) & 0xffff) << 8
        _lib.dfx_op_set_profiling(self.h, mode)

    def autotune(self, dtype, B, reps=3):
        """Pick the fastest column-chunk width (tiles of 16 columns) of every
        size class for applies of dtype with exactly B columns (B = 0 forgets
        the tuned widths of dtype).  Returns {class (M tiles per wave): width}."""
        cls = np.zeros(16, dtype=np.uint32)
        wid = np.zeros(16, dtype=np.uint32)
        n = C.c_uint32(0)
        self.ctx.check(_lib.dfx_op_autotune(self.h, int(dtype), int(B), int(reps), 16,
                                            _ptr(cls), _ptr(wid), C.byref(n)))
        return {int(c): int(w) for c, w in zip(cls[:n.value], wid[:n.value])}

    def profile(self, reset=True):
        """launches, summed kernel ms, algorithmic flops and bytes of the
        profiled cell-apply kernel launches."""
        p = _OpProfile()
        self.ctx.check(_lib.dfx_op_get_profile(self.h, C.byref(p), 1 if reset else 0))
        return {k: getattr(p, k) for k, _ in _OpProfile._fields_}

    def close(self):
        if getattr(self, "h", None):
            _lib.dfx_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CellOperator(_Operator):
    """Cell-wise operator Y = C^T (sum_e Q_e^T A_e Q_e) C X
    (KohnShamOperatorContextFE / CFEOverlapOperatorContext / LaplaceOperatorContextFE)."""

    def __init__(self, ctx, cellNdofs, cellMap, cellMatrices, nDofs, nClassical=None,
                 constraints=None, keep_fp32=False):
        self.ctx = ctx
        cellNdofs, cellMap = _u32a(cellNdofs), _u32a(cellMap)
        mats = _f64a(cellMatrices)
        nClassical = int(nDofs if nClassical is None else nClassical)
        h = C.c_void_p()
        ctx.check(_lib.dfx_op_create_cellwise(
            ctx.h, len(cellNdofs), _ptr(cellNdofs), _ptr(cellMap), _ptr(mats), int(nDofs),
            nClassical, constraints.h if constraints else None,
            OP_KEEP_FP32 if keep_fp32 else 0, C.byref(h)))
        self.h = h
        self.constraints = constraints
        self.nDofs = int(nDofs)
        self.sumNe = int(cellNdofs.astype(np.uint64).sum())
        self.nCells = len(cellNdofs)

    def update_cells(self, cellMatrices):
        mats = _f64a(cellMatrices)
        self.ctx.check(_lib.dfx_op_update_cells(self.h, _ptr(mats)))

    def export_map(self):
        n = np.zeros(self.nCells, np.uint32)
        m = np.zeros(self.sumNe, np.uint32)
        self.ctx.check(_lib.dfx_op_export_map(self.h, _ptr(n), _ptr(m)))
        return n, m

    def gather(self, X, Xcell):
        self.ctx.check(_lib.dfx_op_gather(self.h, X.h, Xcell.h))

    def scatter_add(self, Ycell, Y):
        self.ctx.check(_lib.dfx_op_scatter_add(self.h, Ycell.h, Y.h))


class DiagonalOperator(_Operator):
    """Y = C2P(diag o P2C(X) + [0; E X_enr])  (CFEOverlapInverseOpContextGLL,
    OrthoEFEOverlapInverseOpContextGLL, PreconditionerJacobi when constraints is None)."""

    def __init__(self, ctx, diag, nEnriched=0, enrBlock=None, constraints=None):
        self.ctx = ctx
        diag = _f64a(diag)
        blk = _f64a(enrBlock) if nEnriched else None
        h = C.c_void_p()
        ctx.check(_lib.dfx_op_create_diagonal(
            ctx.h, len(diag), _ptr(diag), int(nEnriched), _ptr(blk) if blk is not None else None,
            constraints.h if constraints else None, C.byref(h)))
        self.h = h
        self.nDofs = len(diag)
        self.constraints = constraints


class _ChfsiParams(C.Structure):
    _fields_ = [("wantedSpectrumLowerBound", C.c_double),
                ("wantedSpectrumUpperBound", C.c_double),
                ("unWantedSpectrumUpperBound", C.c_double),
                ("polynomialDegree", C.c_uint32), ("illConditionTolerance", C.c_double),
                ("isResidualChebyshevFilter", C.c_int),
                ("eigenVectorBatchSize", C.c_uint32), ("precision", C.c_int),
                ("useGraph", C.c_int), ("orthogonalizationType", C.c_int),
                ("storeIntermediateSubspaces", C.c_int),
                ("noProductReuse", C.c_int)]


# linearAlgebra::OrthogonalizationType (ChebyshevFilteredEigenSolver.h:45-50)
ORTHO_CHOLESKY_GRAMSCHMIDT, ORTHO_MULTIPASS_LOWDIN, ORTHO_MULTIPASS_CGS = 0, 1, 2


class ChebyshevFilteredEigenSolver:
    """linearAlgebra::ChebyshevFilteredEigenSolver on one GPU."""

    def __init__(self, ctx, wantedSpectrumLowerBound, wantedSpectrumUpperBound,
                 unWantedSpectrumUpperBound, polynomialDegree, illConditionTolerance,
                 eigenSubspaceGuess, isResidualChebyshevFilter=True,
                 eigenVectorBatchSize=0, precision=PREC_FP64, useGraph=False,
                 orthoType=ORTHO_CHOLESKY_GRAMSCHMIDT, storeIntermediateSubspaces=False,
                 productReuse=True):
        self.ctx = ctx
        p = _ChfsiParams(wantedSpectrumLowerBound, wantedSpectrumUpperBound,
                         unWantedSpectrumUpperBound, int(polynomialDegree),
                         illConditionTolerance, 1 if isResidualChebyshevFilter else 0,
                         int(eigenVectorBatchSize), int(precision), 1 if useGraph else 0,
                         int(orthoType), 1 if storeIntermediateSubspaces else 0,
                         0 if productReuse else 1)
        h = C.c_void_p()
        ctx.check(_lib.dfx_chfsi_create(ctx.h, C.byref(p), eigenSubspaceGuess.h, C.byref(h)))
        self.h = h
        self.guess = eigenSubspaceGuess

    def reinit(self, a, b, c, degree, tol, guess=None):
        self.ctx.check(_lib.dfx_chfsi_reinit(self.h, a, b, c, int(degree), tol,
                                             guess.h if guess is not None else None))
        if guess is not None:
            self.guess = guess

    def solve(self, A, eigenValues, eigenVectors, computeEigenVectors=True, B=None, BInv=None):
        """Returns the EigenSolverErrorCode; eigenValues (numpy, in/out) updated in place."""
        ev = eigenValues
        if not (isinstance(ev, np.ndarray) and ev.dtype == np.float64 and ev.flags.c_contiguous
                and ev.ndim == 1):
            raise TypeError("eigenValues must be a contiguous 1-D float64 numpy array")
        if ev.size != eigenVectors.B:
            raise ValueError(f"eigenValues holds {ev.size} values for {eigenVectors.B} "
                             "eigenvectors")
        if B is None or BInv is None:
            raise ValueError("B (overlap) and BInv (its inverse) operators are required")
        rc = _lib.dfx_chfsi_solve(self.h, A.h, _ptr(ev), eigenVectors.h,
                                  1 if computeEigenVectors else 0, B.h, BInv.h)
        return self.ctx.check(rc, allow_positive=True)

    def _subspace(self, ortho):
        out = MultiVector(self.ctx, self.guess.N, self.guess.B)
        self.ctx.check(_lib.dfx_chfsi_get_filtered_subspace(self.h, ortho, out.h))
        return out

    def getFilteredSubspace(self):
        return self._subspace(0)

    def getOrthogonalizedFilteredSubspace(self):
        return self._subspace(1)

    def residual_norms(self, H, M, X, eigenValues):
        """||H x_j - lambda_j M x_j|| per column of the last solve's
        eigenvectors X: from the solver's kept H X and M X when X still holds
        that output, else by applying H and M."""
        ev = _f64a(eigenValues)
        out = np.zeros(X.B)
        self.ctx.check(_lib.dfx_chfsi_residual_norms(self.h, H.h, M.h, X.h, _ptr(ev),
                                                     _ptr(out)))
        return out

    def timings(self):
        out = np.zeros(3)
        _lib.dfx_chfsi_timings(self.h, _ptr(out))
        return {"filter_ms": out[0], "ortho_ms": out[1], "rr_ms": out[2]}

    def close(self):
        if getattr(self, "h", None):
            _lib.dfx_chfsi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cholesky_gram_schmidt(ctx, X, Xo, B):
    return ctx.check(_lib.dfx_cholesky_gram_schmidt(ctx.h, X.h, Xo.h, B.h), allow_positive=True)


def rayleigh_ritz(ctx, A, X, Xo):
    eig = np.zeros(X.B)
    rc = ctx.check(_lib.dfx_rayleigh_ritz(ctx.h, A.h, X.h, _ptr(eig), Xo.h), allow_positive=True)
    return rc, eig


def eigen_residual_norms(ctx, H, M, X, eigenValues):
    """||H x_j - lambda_j M x_j|| per column (KohnShamEigenSolver::
    getLinearEigenSolveResidual)."""
    ev = _f64a(eigenValues)
    if ev.size != X.B:
        raise ValueError(f"{ev.size} eigenvalues for {X.B} vectors")
    out = np.zeros(X.B)
    ctx.check(_lib.dfx_eigen_residual_norms(ctx.h, H.h, M.h, X.h, _ptr(ev), _ptr(out)))
    return out


def lanczos_extreme(ctx, A, B, BInv, guess, max_krylov=20):
    lo, hi, beta = C.c_double(), C.c_double(), C.c_double()
    ctx.check(_lib.dfx_lanczos_extreme(ctx.h, A.h, B.h, BInv.h, guess.h, int(max_krylov),
                                       C.byref(lo), C.byref(hi), C.byref(beta)))
    return lo.value, hi.value, beta.value


class _KsParams(C.Structure):
    _fields_ = [("numWantedEigenvalues", C.c_uint32), ("maxChebyshevFilterPass", C.c_uint32),
                ("polynomialDegree", C.c_uint32), ("eigenSolveResidualTolerance", C.c_double),
                ("waveFunctionBatchSize", C.c_uint32), ("isResidualChebyshevFilter", C.c_int),
                ("precision", C.c_int), ("lanczosMaxKrylov", C.c_uint32),
                ("numElectrons", C.c_double), ("smearingTemperature", C.c_double),
                ("fermiEnergyTolerance", C.c_double), ("fracOccupancyTolerance", C.c_double),
                ("useGraph", C.c_int)]


class _KsResult(C.Structure):
    _fields_ = [("passes", C.c_uint32), ("lanczosLow", C.c_double),
                ("lanczosHigh", C.c_double), ("lanczosBeta", C.c_double),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("degree", C.c_uint32), ("filterMs", C.c_double), ("orthoMs", C.c_double),
                ("rrMs", C.c_double), ("residualMs", C.c_double), ("lanczosMs", C.c_double),
                ("fermiEnergy", C.c_double), ("numOccupied", C.c_uint32)]


class KohnShamEigenSolver:
    """ksdft::KohnShamEigenSolver (Lanczos bounds, degree lookup, ChFSI passes,
    Fermi level and occupied-state residual convergence)."""

    def __init__(self, ctx, numWantedEigenvalues, eigenSolveResidualTolerance,
                 maxChebyshevFilterPass, waveFunctionSubspaceGuess, lanczosGuess,
                 polynomialDegree=0, waveFunctionBatchSize=0,
                 isResidualChebyshevFilter=True, precision=PREC_FP64, lanczosMaxKrylov=20,
                 numElectrons=0.0, smearingTemperature=500.0, fermiEnergyTolerance=1e-10,
                 fracOccupancyTolerance=1e-3, useGraph=False):
        self.ctx = ctx
        self.p = _KsParams(int(numWantedEigenvalues), int(maxChebyshevFilterPass),
                           int(polynomialDegree), float(eigenSolveResidualTolerance),
                           int(waveFunctionBatchSize), 1 if isResidualChebyshevFilter else 0,
                           int(precision), int(lanczosMaxKrylov), float(numElectrons),
                           float(smearingTemperature), float(fermiEnergyTolerance),
                           float(fracOccupancyTolerance), 1 if useGraph else 0)
        self.guess = waveFunctionSubspaceGuess
        self.lanczosGuess = lanczosGuess
        self.result = None
        self.residualNorms = None

    def solve(self, H, energies, waveFunctions, M, MInv, MLanczos):
        res = np.zeros(self.p.numWantedEigenvalues)
        r = _KsResult()
        rc = _lib.dfx_ks_eigensolve(self.ctx.h, C.byref(self.p), H.h, M.h, MInv.h, MLanczos.h,
                                    self.guess.h, self.lanczosGuess.h, _ptr(energies),
                                    waveFunctions.h, _ptr(res), C.byref(r))
        self.ctx.check(rc, allow_positive=True)
        self.residualNorms = res
        self.result = {k: getattr(r, k) for k, _ in _KsResult._fields_}
        return rc


class CGLinearSolver:
    """linearAlgebra::CGLinearSolver (multi-RHS PCG, per-column convergence)."""

    def __init__(self, ctx, maxIter, absoluteTol, relativeTol, divergenceTol):
        self.ctx, self.maxIter = ctx, int(maxIter)
        self.absTol, self.relTol, self.divTol = absoluteTol, relativeTol, divergenceTol
        self.iterations = 0
        self.residualNorms = None

    def solve(self, *args):
        """solve(linearSolverFunction) as the reference (CGLinearSolver.t.cpp:
        68-313: the function's Ax / PC contexts, right-hand side and initial
        guess; setSolution at the end), or solve(A, pc, b, x) with x the
        initial guess on entry and the solution on return.  Returns the
        LinearSolverErrorCode."""
        if len(args) == 1:
            fn = args[0]
            b = fn.getRhs()
            x = fn.getInitialGuess()
            rc = self.solve(fn.getAxContext(), fn.getPCContext(), b, x)
            fn.setSolution(x)
            return rc
        A, pc, b, x = args
        it = C.c_uint32()
        res = np.zeros(b.B)
        rc = _lib.dfx_cg_solve(self.ctx.h, A.h, pc.h, b.h, x.h, self.absTol, self.relTol,
                               self.divTol, self.maxIter, C.byref(it), _ptr(res))
        self.ctx.check(rc, allow_positive=True)
        self.iterations = it.value
        self.residualNorms = res
        return rc


class _Borrowed(_Operator):
    """An operator handle owned by another object (never destroyed here)."""

    def __init__(self, ctx, h, owner):
        self.ctx, self._hb, self._owner = ctx, h, owner

    @property
    def h(self):
        return self._hb

    def close(self):
        pass


class PoissonLinearSolverFunctionFE:
    """electrostatics::PoissonLinearSolverFunctionFE on the GPU
    (src/electrostatics/PoissonLinearSolverFunctionFE.t.cpp:38-660):
    right-hand side = integral of the quadrature values against the basis
    minus the stiffness applied to the Dirichlet lift, Jacobi diagonal
    assembled from the stiffness cell matrices, solution = x + u_D.

    problem: the Problem whose basis and quadrature the values live on;
    stiffness: its cell-wise K operator (homogeneous constraints);
    fieldConstraints: the field's constraints (Dirichlet values in the
    inhomogeneities); rhs: MultiVector (quadrature points x components)."""

    def __init__(self, problem, stiffness, fieldConstraints, rhs, pcType=PC_JACOBI):
        self.ctx = problem.ctx
        self._keep = [problem, stiffness, fieldConstraints]
        h = C.c_void_p()
        self.ctx.check(_lib.dfx_poisson_create(problem.h, stiffness.h, fieldConstraints.h,
                                               rhs.h, int(pcType), C.byref(h)))
        self.h = h
        self.nDofs, self.nComponents = problem.nDofs, rhs.B
        a, p = C.c_void_p(), C.c_void_p()
        _lib.dfx_poisson_get_ax(self.h, C.byref(a))
        _lib.dfx_poisson_get_pc(self.h, C.byref(p))
        self._ax = _Borrowed(self.ctx, a, self)
        self._pc = _Borrowed(self.ctx, p, self)

    def reinit(self, fieldConstraints, rhs):
        if fieldConstraints is not None:
            self._keep.append(fieldConstraints)
        self.ctx.check(_lib.dfx_poisson_reinit(
            self.h, fieldConstraints.h if fieldConstraints is not None else None, rhs.h))

    def getAxContext(self):
        return self._ax

    def getPCContext(self):
        return self._pc

    def _copy(self, fn):
        out = MultiVector(self.ctx, self.nDofs, self.nComponents)
        self.ctx.check(fn(self.h, out.h))
        return out

    def getRhs(self):
        return self._copy(_lib.dfx_poisson_get_rhs)

    def getInitialGuess(self):
        return self._copy(_lib.dfx_poisson_get_initial_guess)

    def setSolution(self, x):
        self.ctx.check(_lib.dfx_poisson_set_solution(self.h, x.h))

    def getSolution(self, solution=None):
        out = solution or MultiVector(self.ctx, self.nDofs, self.nComponents)
        self.ctx.check(_lib.dfx_poisson_get_solution(self.h, out.h))
        return out

    def solve(self, maxIter, absTol, relTol, divTol, solution):
        """CG solve and getSolution in one library call; returns
        (LinearSolverErrorCode, iterations, residual norms)."""
        it = C.c_uint32()
        res = np.zeros(self.nComponents)
        rc = _lib.dfx_poisson_solve(self.h, int(maxIter), absTol, relTol, divTol, solution.h,
                                    C.byref(it), _ptr(res))
        self.ctx.check(rc, allow_positive=True)
        return rc, it.value, res

    def close(self):
        if getattr(self, "h", None):
            _lib.dfx_poisson_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- problems
class _ProblemSpec(C.Structure):
    _fields_ = [("p", C.c_uint32), ("n", C.c_uint32 * 3), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("nq", C.c_uint32), ("nqEnr", C.c_uint32),
                ("nAtoms", C.c_uint32), ("atomXYZ", C.c_void_p), ("atomZ", C.c_void_p),
                ("vSmear", C.c_double), ("enrich", C.c_int), ("rcut", C.c_double),
                ("smooth", C.c_double), ("orthogonalize", C.c_int),
                ("atomData", C.c_void_p), ("enrichField", C.c_char_p),
                ("localPotential", C.c_void_p)]


_PROB_ARRAYS = {  # name: (id, dtype, size key)
    "cellNdofs": (0, np.uint32, "nCells"), "cellMap": (1, np.uint32, "mapLen"),
    "rows": (2, np.uint32, "nRows"), "rowSizes": (3, np.uint32, "nRows"),
    "rowOffsets": (4, np.uint32, "nRows"), "cols": (5, np.uint32, "nEntries"),
    "weights": (6, np.float64, "nEntries"), "inhom": (7, np.float64, "nRows"),
    "H": (8, np.float64, "matLen"), "M": (9, np.float64, "matLen"),
    "K": (10, np.float64, "matLen"), "dinv": (11, np.float64, "nDofs"),
    "mgll": (12, np.float64, "nDofs"), "enrMinv": (13, np.float64, "nE2"),
    "enrM": (14, np.float64, "nE2"), "alpha": (15, np.float64, "nEC"),
    "nodeXYZ": (16, np.float64, "nC3"),
    "quadOffsets": (17, np.uint64, "nCells1"), "quadW": (18, np.float64, "nQuad"),
    "quadXYZ": (19, np.float64, "nQuad3"),
}


def _host_check(rc):
    if rc < 0:
        raise DfxError(rc, _lib.dfx_ctx_last_error(None).decode())
    return rc


class SphericalDataNumerical:
    """One radial function of an atom file: R(r) Y_lm f_c(r) with R the
    zero-end-slope cubic spline through the file's values
    (atoms::SphericalDataNumerical, src/atoms/SphericalDataNumerical.h:40-140)."""

    def __init__(self, atom, field, k):
        self.atom, self.field, self.k = atom, field, k
        nk, mixed = C.c_uint32(), C.c_int()
        tail, cl = C.c_double(), C.c_double()
        _host_check(_lib.dfx_atom_function_knots(atom.h, field.encode(), k, C.byref(nk), None,
                                                 C.byref(mixed), C.byref(tail), C.byref(cl)))
        NR = nk.value
        self.knots = np.zeros(NR)
        _host_check(_lib.dfx_atom_function_knots(atom.h, field.encode(), k, None,
                                                 _ptr(self.knots), None, None, None))
        self.mixed, self.tail, self.cLeft = bool(mixed.value), tail.value, cl.value
        q = np.zeros(3, dtype=np.int32)
        cut, sm = C.c_double(), C.c_double()
        self.values = np.zeros(NR)
        self.spline_b, self.spline_c, self.spline_d = np.zeros(NR), np.zeros(NR), np.zeros(NR)
        _host_check(_lib.dfx_atom_function(atom.h, field.encode(), k, _ptr(q), C.byref(cut),
                                           C.byref(sm), _ptr(self.values), _ptr(self.spline_b),
                                           _ptr(self.spline_c), _ptr(self.spline_d)))
        self.qNumbers = [int(v) for v in q]
        self.cutoff, self.smoothness = cut.value, sm.value

    def getQNumbers(self):
        return list(self.qNumbers)

    def evaluate(self, ctx, points, origin=(0.0, 0.0, 0.0), gradient=True):
        """Values (and gradients) at points (n x 3) on the device; returns
        numpy arrays (n,) and (n, 3)."""
        P = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        n = P.shape[0]
        pm = MultiVector(ctx, n, 3)
        if n:
            pm.upload(P)
        vm = MultiVector(ctx, n, 1)
        gm = MultiVector(ctx, n, 3) if gradient else None
        o = np.ascontiguousarray(origin, dtype=np.float64)
        ctx.check(_lib.dfx_atom_evaluate(ctx.h, self.atom.h, self.field.encode(), self.k,
                                         _ptr(o), pm.h, vm.h, gm.h if gm else None))
        v = vm.to_numpy().reshape(n) if n else np.zeros(0)
        g = (gm.to_numpy() if n else np.zeros((0, 3))) if gm else None
        return (v, g) if gradient else v


class AtomSphericalDataEnrichment:
    """An atom file of the reference's enrichment XML format
    (atoms::AtomSphericalDataEnrichment(fileName, fieldNames, metadataNames),
    src/atoms/AtomSphericalDataEnrichment.h:41-89): metadata elements, the
    radial grid and, per field, its numerical spherical functions."""

    _reader = "dfx_atom_read"

    def __init__(self, fileName, fieldNames, metadataNames=()):
        names = [f.encode() for f in fieldNames]
        arr = (C.c_char_p * max(1, len(names)))(*names)
        h = C.c_void_p()
        _host_check(getattr(_lib, self._reader)(os.fspath(fileName).encode(), arr, len(names),
                                                C.byref(h)))
        self.h = h
        self.fileName = os.fspath(fileName)
        nr = C.c_uint32()
        _lib.dfx_atom_radial_grid(self.h, C.byref(nr), None)
        self.NR = nr.value
        self._metadata = {m: self.getMetadata(m) for m in metadataNames}
        self._data = {}
        for f in fieldNames:
            cnt = C.c_uint32()
            _host_check(_lib.dfx_atom_num_functions(self.h, f.encode(), C.byref(cnt)))
            self._data[f] = [SphericalDataNumerical(self, f, k) for k in range(cnt.value)]

    def __del__(self):
        if getattr(self, "h", None):
            _lib.dfx_atom_destroy(self.h)
            self.h = None

    def getFileName(self):
        return self.fileName

    def getMetadata(self, name):
        need = C.c_uint64()
        _host_check(_lib.dfx_atom_metadata(self.h, name.encode(), None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _host_check(_lib.dfx_atom_metadata(self.h, name.encode(), buf, need.value, None))
        return buf.value.decode()

    def radialGrid(self):
        r = np.zeros(self.NR)
        _host_check(_lib.dfx_atom_radial_grid(self.h, None, _ptr(r)))
        return r

    def getSphericalData(self, fieldName, qNumbers=None):
        funs = self._data[fieldName]
        if qNumbers is None:
            return funs
        for f in funs:
            if f.qNumbers == list(qNumbers):
                return f
        raise KeyError(f"no function with quantum numbers {qNumbers} in {fieldName}")


class AtomSphericalDataPSP(AtomSphericalDataEnrichment):
    """A UPF v2 pseudopotential file (atoms::AtomSphericalDataPSP(fileName,
    fieldNames, metadataNames), src/atoms/AtomSphericalDataPSP.h:47-140):
    PP_HEADER metadata and the fields vlocal (mixed), beta, pswfc, nlcc,
    rhoatom."""
    _reader = "dfx_atom_read_upf"


class Problem:
    """The structured-mesh FE problem (maps, Dirichlet rows, cell matrices)
    built on the GPU: the stand-in for the reference's deal.II setup."""

    def __init__(self, ctx, p=3, n=(8, 8, 8), lo=(-10., -10., -10.), hi=(10., 10., 10.),
                 atoms=(), nq=0, nq_enr=0, v_smear=0.5, enrich=False, rcut=4.0,
                 smooth=2.0, orthogonalize=True, atom_data=None, enrich_field="orbital",
                 local_potential=None):
        """atom_data: per atom an AtomSphericalDataEnrichment whose field
        `enrich_field` supplies that atom's enrichment functions (numerical
        radial data, each with its own cutoff); None: the analytic Slater set
        with cutoff rcut and smoothness smooth.  local_potential: per atom an
        AtomSphericalDataPSP read with "vlocal" (its local part replaces the
        smeared Coulomb of that atom) or None."""
        self.ctx = ctx
        xyz = np.ascontiguousarray([a[:3] for a in atoms], dtype=np.float64).reshape(-1)
        Z = np.ascontiguousarray([int(a[3]) for a in atoms], dtype=np.int32)
        handles = None
        if atom_data is not None:
            if len(atom_data) != len(atoms):
                raise ValueError("atom_data needs one entry per atom")
            handles = (C.c_void_p * max(1, len(atoms)))(*[d.h.value for d in atom_data])
        pots = None
        if local_potential is not None:
            if len(local_potential) != len(atoms):
                raise ValueError("local_potential needs one entry per atom")
            pots = (C.c_void_p * max(1, len(atoms)))(
                *[d.h.value if d is not None else None for d in local_potential])
        self._keep = (xyz, Z, handles, list(atom_data or ()), pots, list(local_potential or ()))
        s = _ProblemSpec()
        s.p = p
        s.n = (C.c_uint32 * 3)(*[int(v) for v in n])
        s.lo = (C.c_double * 3)(*[float(v) for v in lo])
        s.hi = (C.c_double * 3)(*[float(v) for v in hi])
        s.nq, s.nqEnr, s.nAtoms = nq, nq_enr, len(atoms)
        s.atomXYZ = xyz.ctypes.data if len(atoms) else None
        s.atomZ = Z.ctypes.data if len(atoms) else None
        s.vSmear, s.enrich, s.rcut, s.smooth = v_smear, 1 if enrich else 0, rcut, smooth
        s.orthogonalize = 1 if orthogonalize else 0
        s.atomData = C.cast(handles, C.c_void_p) if handles is not None else None
        s.enrichField = enrich_field.encode() if handles is not None else None
        s.localPotential = C.cast(pots, C.c_void_p) if pots is not None else None
        h = C.c_void_p()
        ctx.check(_lib.dfx_problem_build(ctx.h, C.byref(s), C.byref(h)))
        self.h = h
        sz = np.zeros(8, dtype=np.uint64)
        _lib.dfx_problem_sizes(self.h, _ptr(sz))
        (self.nCells, self.nDofs, self.nClassical, self.nEnriched, self.mapLen,
         self.matLen, self.nRows, self.nEntries) = [int(v) for v in sz]
        self.p = p

    def get(self, name):
        wid, dt, key = _PROB_ARRAYS[name]
        sizes = {"nE2": self.nEnriched ** 2, "nEC": self.nEnriched * self.nClassical,
                 "nC3": 3 * self.nClassical, "nCells1": self.nCells + 1}
        if key in ("nQuad", "nQuad3"):
            sizes["nQuad"] = int(self.get("quadOffsets")[-1])
            sizes["nQuad3"] = 3 * sizes["nQuad"]
        n = sizes.get(key, getattr(self, key, 0))
        out = np.zeros(n, dtype=dt)
        if n:
            self.ctx.check(_lib.dfx_problem_get(self.h, wid, _ptr(out)))
        return out

    def density(self, occupation, X):
        """Electron density at every quadrature point (layout: get("quadOffsets"))
        from FP64 wavefunctions X (nDofs x nStates): sum_i 2 f_i psi_i^2.
        Returns a device MultiVector (points x 1)."""
        occ = np.ascontiguousarray(occupation, dtype=np.float64)
        nq = int(self.get("quadOffsets")[-1])
        rho = MultiVector(self.ctx, nq, 1)
        self.ctx.check(_lib.dfx_problem_density(self.h, _ptr(occ), X.h, rho.h))
        return rho

    def integrate(self, f, constraints=None):
        """FEBasisOperations::integrateWithBasisValues: the integrals of the
        quadrature values f (MultiVector, points x components) against every
        basis function, child rows distributed to their parents by
        `constraints` (None: no distribution).  Returns a MultiVector
        (nDofs x components)."""
        out = MultiVector(self.ctx, self.nDofs, f.B)
        self.ctx.check(_lib.dfx_problem_integrate(
            self.h, constraints.h if constraints is not None else None, f.h, out.h))
        return out

    def operators(self, keep_fp32=True):
        """(H, M, K, MInv, MLump, constraints) as operator objects."""
        hs = [C.c_void_p() for _ in range(6)]
        self.ctx.check(_lib.dfx_problem_create_operators(
            self.h, OP_KEEP_FP32 if keep_fp32 else 0, *[C.byref(x) for x in hs]))
        ops = []
        for k, x in enumerate(hs[:5]):
            o = CellOperator.__new__(CellOperator) if k < 3 else DiagonalOperator.__new__(DiagonalOperator)
            o.ctx, o.h, o.nDofs = self.ctx, x, self.nDofs
            o.nCells, o.sumNe = self.nCells, self.mapLen
            ops.append(o)
        cons = Constraints.__new__(Constraints)
        cons.ctx, cons.h, cons.nDofs = self.ctx, hs[5], self.nDofs
        for o in ops:
            o.constraints = cons
        return tuple(ops) + (cons,)

    def close(self):
        if getattr(self, "h", None):
            _lib.dfx_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving
class _Request(C.Structure):
    _fields_ = [("spec", _ProblemSpec), ("numStates", C.c_uint32), ("tolerance", C.c_double),
                ("numElectrons", C.c_double), ("maxPasses", C.c_uint32),
                ("guessSeed", C.c_uint64), ("arrivalSeconds", C.c_double)]


class _RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("passes", C.c_uint32), ("admitSeconds", C.c_double),
                ("doneSeconds", C.c_double), ("footprintBytes", C.c_uint64),
                ("slot", C.c_uint32)]


class _ServeParams(C.Structure):
    _fields_ = [("device", C.c_int), ("maxSlots", C.c_uint32),
                ("memoryBudgetBytes", C.c_uint64), ("precision", C.c_int),
                ("useGraph", C.c_int)]


def _request_struct(r, keep):
    """r: dict with p, n, lo, hi, atoms [(x, y, z, Z)], states, tol, electrons,
    max_passes, seed, arrival (and optional enrich / rcut / smooth / v_smear)."""
    atoms = r["atoms"]
    xyz = np.ascontiguousarray([a[:3] for a in atoms], dtype=np.float64).reshape(-1)
    Z = np.ascontiguousarray([int(a[3]) for a in atoms], dtype=np.int32)
    keep.append((xyz, Z))
    q = _Request()
    s = q.spec
    s.p = int(r["p"])
    s.n = (C.c_uint32 * 3)(*[int(v) for v in r["n"]])
    s.lo = (C.c_double * 3)(*[float(v) for v in r["lo"]])
    s.hi = (C.c_double * 3)(*[float(v) for v in r["hi"]])
    s.nq, s.nqEnr, s.nAtoms = 0, 0, len(atoms)
    s.atomXYZ = xyz.ctypes.data if len(atoms) else None
    s.atomZ = Z.ctypes.data if len(atoms) else None
    s.vSmear = float(r.get("v_smear", 0.5))
    s.enrich = 1 if r.get("enrich", True) else 0
    s.rcut, s.smooth = float(r.get("rcut", 4.0)), float(r.get("smooth", 2.0))
    s.orthogonalize = 1 if r.get("orthogonalize", True) else 0
    files = r.get("atom_files")
    if files:
        # one atom file per element symbol or per atom; the handles live in
        # `keep` for the duration of the call
        per = [files[a[4]] if len(a) > 4 else files[i] for i, a in enumerate(atoms)] \
            if isinstance(files, dict) else list(files)
        if len(per) != len(atoms):
            raise ValueError("atom_files: one file per atom (or per symbol)")
        cache = {}
        data = [cache.setdefault(f, AtomSphericalDataEnrichment(f, [r.get("enrich_field", "orbital")]))
                for f in per]
        handles = (C.c_void_p * len(atoms))(*[d.h.value for d in data])
        field = r.get("enrich_field", "orbital").encode()
        keep.append((data, handles, field))
        s.atomData = C.cast(handles, C.c_void_p)
        s.enrichField = field
    q.numStates = int(r["states"])
    q.tolerance = float(r.get("tol", 1e-6))
    q.numElectrons = float(r["electrons"])
    q.maxPasses = int(r.get("max_passes", 50))
    q.guessSeed = int(r.get("seed", 1))
    q.arrivalSeconds = float(r.get("arrival", 0.0))
    return q


# ------------------------------------------------------------ wire format
# One JSON object per request and per response (JSON lines).
REQUEST_SCHEMA = {
    # key: (type, required, meaning)
    "id": (str, True, "caller's request id, echoed in the response"),
    "p": (int, True, "FE order"),
    "n": (list, True, "cells per direction [nx, ny, nz]"),
    "lo": (list, True, "domain lower corner [x, y, z] (bohr)"),
    "hi": (list, True, "domain upper corner"),
    "atoms": (list, True, "[[x, y, z, Z], ...] or [[x, y, z, Z, symbol], ...]"),
    "states": (int, True, "eigenpairs wanted"),
    "electrons": (float, True, "electron count (occupations, convergence set)"),
    "tol": (float, False, "residual tolerance of the occupied states, default 1e-6"),
    "max_passes": (int, False, "ChFSI pass limit, default 50"),
    "seed": (int, False, "initial-subspace seed, default 1"),
    "arrival": (float, False, "arrival time (s) in the stream, default 0"),
    "enrich": (bool, False, "enrichment on, default true"),
    "rcut": (float, False, "analytic enrichment cutoff, default 4"),
    "smooth": (float, False, "analytic enrichment smoothness, default 2"),
    "v_smear": (float, False, "nuclear smearing radius, default 0.5"),
    "orthogonalize": (bool, False, "orthogonalized enrichment, default true"),
    "atom_files": ((dict, list), False, "atom XML file per symbol (dict) or per atom (list)"),
    "enrich_field": (str, False, "field of the atom files, default \"orbital\""),
}
RESPONSE_KEYS = ("id", "status", "passes", "eigenvalues", "arrival", "admit_seconds",
                 "done_seconds", "latency_seconds", "footprint_bytes", "slot")


def parse_request(line):
    """A request dict from one JSON line, checked against REQUEST_SCHEMA
    (ValueError names the offending key)."""
    import json
    r = json.loads(line) if isinstance(line, (str, bytes)) else dict(line)
    if not isinstance(r, dict):
        raise ValueError("a request is a JSON object")
    for k in r:
        if k not in REQUEST_SCHEMA:
            raise ValueError(f"unknown request key {k!r}")
    for k, (t, req, _) in REQUEST_SCHEMA.items():
        if k not in r:
            if req:
                raise ValueError(f"missing request key {k!r}")
            continue
        v = r[k]
        ok = isinstance(v, t) and not (t is int and isinstance(v, bool))
        if t is float:
            ok = isinstance(v, (int, float)) and not isinstance(v, bool)
        if not ok:
            raise ValueError(f"request key {k!r}: expected {t}")
    for k, m in (("n", 3), ("lo", 3), ("hi", 3)):
        if len(r[k]) != m:
            raise ValueError(f"request key {k!r}: expected {m} values")
    for a in r["atoms"]:
        if not (isinstance(a, list) and len(a) in (4, 5)):
            raise ValueError("request key 'atoms': entries [x, y, z, Z] or [x, y, z, Z, symbol]")
    if isinstance(r.get("atom_files"), dict) and any(len(a) != 5 for a in r["atoms"]):
        raise ValueError("atom_files by symbol needs [x, y, z, Z, symbol] atoms")
    return r


def serve_json(lines, **kw):
    """Serve JSON-line requests; returns one JSON line per request, in input
    order: id, status (0 or the library / solver code), passes, eigenvalues,
    arrival, admit / done / latency seconds, footprint bytes, slot."""
    import json
    reqs = [parse_request(l) for l in lines if (l.strip() if isinstance(l, str) else l)]
    energies, res = serve(reqs, **kw)
    out = []
    for r, e, x in zip(reqs, energies, res):
        arr = float(r.get("arrival", 0.0))
        out.append(json.dumps({
            "id": r["id"], "status": int(x["status"]), "passes": int(x["passes"]),
            "eigenvalues": [float(v) for v in e], "arrival": arr,
            "admit_seconds": float(x["admitSeconds"]), "done_seconds": float(x["doneSeconds"]),
            "latency_seconds": float(x["doneSeconds"]) - arr,
            "footprint_bytes": int(x["footprintBytes"]), "slot": int(x["slot"])}))
    return out


def request_footprint(r):
    """Estimated peak device bytes of one request (what admission counts)."""
    keep = []
    q = _request_struct(r, keep)
    out = C.c_uint64(0)
    rc = _lib.dfx_request_footprint(C.byref(q), C.byref(out))
    if rc != 0:
        raise DfxError(rc, "dfx_request_footprint")
    return int(out.value)


def serve(requests, max_slots=4, memory_budget=200 << 30, precision=PREC_MIXED, device=0,
          use_graph=True):
    """Run a stream of independent Kohn-Sham eigenproblems with continuous
    batching (dfx_serve).  use_graph: each request replays its captured
    degree loop on every pass.  Returns (energies per request, result dicts)."""
    keep = []
    arr = (_Request * len(requests))(*[_request_struct(r, keep) for r in requests])
    res = (_RequestResult * len(requests))()
    total = sum(int(r["states"]) for r in requests)
    en = np.zeros(max(total, 1))
    p = _ServeParams(int(device), int(max_slots), int(memory_budget), int(precision),
                     1 if use_graph else 0)
    rc = _lib.dfx_serve(C.byref(p), len(requests), arr, _ptr(en), res)
    if rc < 0:
        raise DfxError(rc, "dfx_serve failed (a request hit a library error)")
    out, off = [], 0
    energies = []
    for r, x in zip(requests, res):
        k = int(r["states"])
        energies.append(en[off:off + k].copy())
        off += k
        out.append({f: getattr(x, f) for f, _ in _RequestResult._fields_})
    return energies, out
