// CPU oracle, part 2: constraints, gather / scatter, the cell-wise apply,
// overlap inverses, Chebyshev filters, Cholesky-Gram-Schmidt, Rayleigh-Ritz,
// ChFSI, Lanczos bounds, eigen-residual norms and multi-RHS PCG.
//
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).  Every routine is the reference's
// sequential host algorithm restated in plain loops; citations per function.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace oracle
{
  // --------------------------------------------------------- constraints
  // src/basis/ConstraintsInternal.cpp:38-107.  newValues = inhom, then an
  // axpy per parent (x[col] * w accumulated with fma), then stored in the row.
  void distributeParentToChild(const Constraints &c, double *x, u32 B)
  {
    std::vector<double> nv(B);
    for (size_t i = 0; i < c.rows.size(); ++i)
      {
        std::fill(nv.begin(), nv.end(), c.inhom[i]);
        const u32 off = c.rowOffsets[i];
        for (u32 j = 0; j < c.rowSizes[i]; ++j)
          {
            const double  w   = c.weights[off + j];
            const double *src = x + (size_t)c.cols[off + j] * B;
            for (u32 b = 0; b < B; ++b)
              nv[b] = std::fma(w, src[b], nv[b]);
          }
        std::copy(nv.begin(), nv.end(), x + (size_t)c.rows[i] * B);
      }
  }

  // src/basis/ConstraintsInternal.cpp:113-170: sequential in row order,
  // y[col] += w * y[row] (axpy, fma), then y[row] = 0.
  void distributeChildToParent(const Constraints &c, double *y, u32 B)
  {
    for (size_t i = 0; i < c.rows.size(); ++i)
      {
        double   *row = y + (size_t)c.rows[i] * B;
        const u32 off = c.rowOffsets[i];
        for (u32 j = 0; j < c.rowSizes[i]; ++j)
          {
            const double w   = c.weights[off + j];
            double      *dst = y + (size_t)c.cols[off + j] * B;
            for (u32 b = 0; b < B; ++b)
              dst[b] = std::fma(w, row[b], dst[b]);
          }
        std::fill(row, row + B, 0.0);
      }
  }

  // src/basis/ConstraintsInternal.cpp:176-217
  void setConstrainedNodes(const Constraints &c, double *x, u32 B, double v)
  {
    for (u32 r : c.rows)
      std::fill(x + (size_t)r * B, x + (size_t)r * B + B, v);
  }

  // ------------------------------------------------------ gather / scatter
  // src/basis/FECellWiseDataOperations.t.cpp:31-84
  void gatherCells(const Problem &P, const double *x, u32 B, double *xCell)
  {
    for (u32 e = 0; e < P.nCells; ++e)
      for (u32 i = 0; i < P.cellNdofs[e]; ++i)
        {
          const size_t slot = P.cellMapOffset[e] + i;
          std::memcpy(xCell + slot * B, x + (size_t)P.cellMap[slot] * B,
                      sizeof(double) * B);
        }
  }

  // src/basis/FECellWiseDataOperations.t.cpp:89-151 (sequential, cell order)
  void scatterAddCells(const Problem &P, const double *yCell, u32 B, double *y)
  {
    for (u32 e = 0; e < P.nCells; ++e)
      for (u32 i = 0; i < P.cellNdofs[e]; ++i)
        {
          const size_t slot = P.cellMapOffset[e] + i;
          double      *dst  = y + (size_t)P.cellMap[slot] * B;
          for (u32 b = 0; b < B; ++b)
            dst[b] += yCell[slot * B + b];
        }
  }

  // --------------------------------------------------------- dense LA
  // Column-major GEMM, C = alpha op(A) op(B) + beta C; reduction over k in
  // ascending order (BLAS semantics of src/linearAlgebra/BlasLapack.t.cpp:
  // 391-436 -> dgemm_).
  void gemm(char ta, char tb, int m, int n, int k, double alpha, const double *A,
            int lda, const double *B, int ldb, double beta, double *C, int ldc)
  {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i)
        {
          double acc = 0.0;
          for (int l = 0; l < k; ++l)
            {
              const double a = (ta == 'N') ? A[i + (size_t)lda * l] : A[l + (size_t)lda * i];
              const double b = (tb == 'N') ? B[l + (size_t)ldb * j] : B[j + (size_t)ldb * l];
              acc += a * b;
            }
          double &c = C[i + (size_t)ldc * j];
          c         = (beta == 0.0 ? 0.0 : beta * c) + alpha * acc;
        }
  }

  // LAPACK dpotrf('L') semantics: lower triangle overwritten with L.
  int potrfLower(int n, double *A, int lda)
  {
    for (int j = 0; j < n; ++j)
      {
        double d = A[j + (size_t)lda * j];
        for (int k = 0; k < j; ++k)
          d -= A[j + (size_t)lda * k] * A[j + (size_t)lda * k];
        if (!(d > 0.0))
          return j + 1;
        d                      = std::sqrt(d);
        A[j + (size_t)lda * j] = d;
        for (int i = j + 1; i < n; ++i)
          {
            double s = A[i + (size_t)lda * j];
            for (int k = 0; k < j; ++k)
              s -= A[i + (size_t)lda * k] * A[j + (size_t)lda * k];
            A[i + (size_t)lda * j] = s / d;
          }
      }
    return 0;
  }

  // LAPACK dtrtri('L','N'): in-place inverse of the lower triangle.
  int trtriLower(int n, double *A, int lda)
  {
    for (int j = 0; j < n; ++j)
      if (A[j + (size_t)lda * j] == 0.0)
        return j + 1;
    std::vector<double> inv((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
      {
        inv[j + (size_t)n * j] = 1.0 / A[j + (size_t)lda * j];
        for (int i = j + 1; i < n; ++i)
          {
            double s = 0.0;
            for (int k = j; k < i; ++k)
              s += A[i + (size_t)lda * k] * inv[k + (size_t)n * j];
            inv[i + (size_t)n * j] = -s / A[i + (size_t)lda * i];
          }
      }
    for (int j = 0; j < n; ++j)
      for (int i = j; i < n; ++i)
        A[i + (size_t)lda * j] = inv[i + (size_t)n * j];
    return 0;
  }

  // Symmetric eigensolver (cyclic Jacobi), eigenvalues ascending, eigenvectors
  // in the columns of Z.  Stands in for heevd / steqr (LAPACK) in the oracle.
  int syevJacobi(int n, double *Ain, int lda, double *w, double *Z, int ldz)
  {
    std::vector<double> A((size_t)n * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        A[i + (size_t)n * j] = (i >= j) ? Ain[i + (size_t)lda * j] : Ain[j + (size_t)lda * i];
    std::vector<double> V((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
      V[i + (size_t)n * i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep)
      {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i)
            {
              tot += A[i + (size_t)n * j] * A[i + (size_t)n * j];
              if (i != j)
                off += A[i + (size_t)n * j] * A[i + (size_t)n * j];
            }
        if (off <= 1e-30 * tot || off == 0.0)
          break;
        for (int p = 0; p < n - 1; ++p)
          for (int q = p + 1; q < n; ++q)
            {
              const double apq = A[p + (size_t)n * q];
              if (std::fabs(apq) < 1e-300)
                continue;
              const double app = A[p + (size_t)n * p], aqq = A[q + (size_t)n * q];
              const double theta = (aqq - app) / (2.0 * apq);
              const double t = (theta >= 0 ? 1.0 : -1.0) /
                               (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
              const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
              for (int k = 0; k < n; ++k)
                {
                  const double akp = A[k + (size_t)n * p], akq = A[k + (size_t)n * q];
                  A[k + (size_t)n * p] = c * akp - s * akq;
                  A[k + (size_t)n * q] = s * akp + c * akq;
                }
              for (int k = 0; k < n; ++k)
                {
                  const double apk = A[p + (size_t)n * k], aqk = A[q + (size_t)n * k];
                  A[p + (size_t)n * k] = c * apk - s * aqk;
                  A[q + (size_t)n * k] = s * apk + c * aqk;
                }
              for (int k = 0; k < n; ++k)
                {
                  const double vkp = V[k + (size_t)n * p], vkq = V[k + (size_t)n * q];
                  V[k + (size_t)n * p] = c * vkp - s * vkq;
                  V[k + (size_t)n * q] = s * vkp + c * vkq;
                }
            }
      }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      return A[a + (size_t)n * a] < A[b + (size_t)n * b];
    });
    for (int j = 0; j < n; ++j)
      {
        w[j] = A[idx[j] + (size_t)n * idx[j]];
        for (int i = 0; i < n; ++i)
          Z[i + (size_t)ldz * j] = V[i + (size_t)n * idx[j]];
      }
    return 0;
  }

  // ---------------------------------------------------------- operators
  // KohnShamOperatorContextFE::apply, src/ksdft/KohnShamOperatorContextFE.t.cpp:
  // 1326-1415 with computeAxCellWiseLocal :766-949: P2C on X, Y = 0, gather,
  // per-cell column-major GEMM (m = B, n = k = n_e, lda = B, ldb = n_e),
  // scatter-add, C2P on Y.
  void cellApply(const Problem &P, const double *mats, double *X, double *Y, u32 B)
  {
    distributeParentToChild(P.cons, X, B);
    std::fill(Y, Y + (size_t)P.nDofs * B, 0.0);
    std::vector<double> xc, yc;
    for (u32 e = 0; e < P.nCells; ++e)
      {
        const u32 ne = P.cellNdofs[e];
        xc.resize((size_t)ne * B);
        yc.resize((size_t)ne * B);
        for (u32 i = 0; i < ne; ++i)
          std::memcpy(&xc[(size_t)i * B],
                      X + (size_t)P.cellMap[P.cellMapOffset[e] + i] * B,
                      sizeof(double) * B);
        gemm('N', 'N', (int)B, (int)ne, (int)ne, 1.0, xc.data(), (int)B,
             mats + P.cellMatOffset[e], (int)ne, 0.0, yc.data(), (int)B);
        for (u32 i = 0; i < ne; ++i)
          {
            double *dst = Y + (size_t)P.cellMap[P.cellMapOffset[e] + i] * B;
            for (u32 b = 0; b < B; ++b)
              dst[b] += yc[(size_t)i * B + b];
          }
      }
    distributeChildToParent(P.cons, Y, B);
  }

  // CFEOverlapInverseOpContextGLL::apply (src/basis/CFEOverlapInverseOpContextGLL.t.cpp:
  // 525-556) and its enriched form OrthoEFEOverlapInverseOpContextGLL::apply
  // (src/basis/OrthoEFEOverlapInverseOpContextGLL.t.cpp:1188-1273):
  // P2C(X); Y = dinv o X; Y_enr = X_enr Minv_ee^T; C2P(Y).
  void diagInvApply(const Problem &P, double *X, double *Y, u32 B)
  {
    distributeParentToChild(P.cons, X, B);
    for (u32 i = 0; i < P.nDofs; ++i)
      for (u32 b = 0; b < B; ++b)
        Y[(size_t)i * B + b] = P.dinv[i] * X[(size_t)i * B + b];
    if (P.nEnriched)
      gemm('N', 'T', (int)B, (int)P.nEnriched, (int)P.nEnriched, 1.0,
           X + (size_t)P.nClassical * B, (int)B, P.enrMinv.data(),
           (int)P.nEnriched, 0.0, Y + (size_t)P.nClassical * B, (int)B);
    distributeChildToParent(P.cons, Y, B);
  }

  // The Lanczos B operator (the GLL-quadrature overlap, diagonal per cell):
  // P2C(X); Y = mgll o X (classical), Y_enr = Mee X_enr; C2P(Y).
  void lumpedMassApply(const Problem &P, double *X, double *Y, u32 B)
  {
    distributeParentToChild(P.cons, X, B);
    for (u32 i = 0; i < P.nDofs; ++i)
      for (u32 b = 0; b < B; ++b)
        Y[(size_t)i * B + b] = P.mgll[i] * X[(size_t)i * B + b];
    if (P.nEnriched)
      {
        // Mee = inverse of enrMinv, rebuilt from the cell blocks on demand
        const u32           nE = P.nEnriched;
        const int           nl = (P.p + 1) * (P.p + 1) * (P.p + 1);
        std::vector<double> Mee((size_t)nE * nE, 0.0);
        for (u32 e = 0; e < P.nCells; ++e)
          {
            const u32     ne = P.cellNdofs[e];
            const double *Me = &P.M[P.cellMatOffset[e]];
            for (u32 j = nl; j < ne; ++j)
              for (u32 i = nl; i < ne; ++i)
                Mee[(P.cellMap[P.cellMapOffset[e] + i] - P.nClassical) +
                    (size_t)nE * (P.cellMap[P.cellMapOffset[e] + j] - P.nClassical)] +=
                  Me[i + (size_t)ne * j];
          }
        gemm('N', 'T', (int)B, (int)nE, (int)nE, 1.0, X + (size_t)P.nClassical * B,
             (int)B, Mee.data(), (int)nE, 0.0, Y + (size_t)P.nClassical * B, (int)B);
      }
    distributeChildToParent(P.cons, Y, B);
  }

  void applyOp(const Problem &P, int kind, double *X, double *Y, u32 B)
  {
    switch (kind)
      {
        case OP_H:
          cellApply(P, P.H.data(), X, Y, B);
          break;
        case OP_M:
          cellApply(P, P.M.data(), X, Y, B);
          break;
        case OP_K:
          cellApply(P, P.K.data(), X, Y, B);
          break;
        case OP_MINV:
          diagInvApply(P, X, Y, B);
          break;
        case OP_MLUMP:
          lumpedMassApply(P, X, Y, B);
          break;
        default:
          throw std::invalid_argument("unknown operator kind");
      }
  }

  // blasLapack::axpby (src/linearAlgebra/BlasLapackKernels.cpp): z = a x + b y
  static void axpby(size_t n, double a, const double *x, double b, const double *y,
                    double *z)
  {
    for (size_t i = 0; i < n; ++i)
      z[i] = a * x[i] + b * y[i];
  }

  // blasLapack::axpbyBlocked (BlasLapackKernels.cpp:477-502):
  // z[i,j] = a1 a[j] x[i,j] + b1 b[j] y[i,j]
  static void axpbyBlocked(size_t n, u32 B, double a1, const double *a,
                           const double *x, double b1, const double *b,
                           const double *y, double *z)
  {
    for (size_t i = 0; i < n; ++i)
      for (u32 j = 0; j < B; ++j)
        z[i * B + j] = a1 * a[j] * x[i * B + j] + b1 * b[j] * y[i * B + j];
  }

  // ------------------------------------------------------- Chebyshev
  // linearAlgebra::ChebyshevFilter, src/linearAlgebra/ChebyshevFilter.t.cpp:39-135.
  // On return X and Y both hold the filtered block.
  void chebyshevFilter(const Problem &P, double *X, u32 B, int m, double a,
                       double b, double c, double *Y)
  {
    const size_t n  = (size_t)P.nDofs * B;
    const double e  = 0.5 * (c - b);
    const double c0 = 0.5 * (c + b);
    double       sigma  = e / (a - c0);
    const double sigma1 = sigma;
    const double gamma  = 2.0 / sigma1;
    std::vector<double> s1(n, 0.0), s2(n, 0.0), xk(X, X + n), yk(n);
    applyOp(P, OP_H, xk.data(), s1.data(), B);
    applyOp(P, OP_MINV, s1.data(), s2.data(), B);
    axpby(n, sigma1 / e, s2.data(), -sigma1 / e * c0, xk.data(), yk.data());
    for (int deg = 2; deg <= m; ++deg)
      {
        const double sigma2 = 1.0 / (gamma - sigma);
        applyOp(P, OP_H, yk.data(), s1.data(), B);
        applyOp(P, OP_MINV, s1.data(), s2.data(), B);
        axpby(n, 2.0 * sigma2 / e, s2.data(), -2.0 * sigma2 / e * c0, yk.data(),
              s1.data());
        axpby(n, 1.0, s1.data(), -sigma * sigma2, xk.data(), xk.data());
        std::swap(xk, yk);
        sigma = sigma2;
      }
    std::copy(yk.begin(), yk.end(), X);
    std::copy(yk.begin(), yk.end(), Y);
  }

  // ResidualChebyshevFilterGEP, src/linearAlgebra/ChebyshevFilter.t.cpp:242-444.
  // X is read (after the P2C the applies perform on it); Y receives the filtered
  // block X_m = M^-1 R_m + mu_m o X.
  void residualChebyshevFilter(const Problem &P, const double *lambda, double *X,
                               u32 B, int m, double a, double b, double c,
                               double *Y)
  {
    const size_t n  = (size_t)P.nDofs * B;
    const double e  = 0.5 * (c - b);
    const double c0 = 0.5 * (c + b);
    double       sigma  = e / (a - c0);
    const double sigma1 = sigma;
    const double gamma  = 2.0 / sigma1;
    std::vector<double> s1(n, 0.0), s2(n, 0.0), s3(n, 0.0), res(n, 0.0),
      resNew(n, 0.0);
    std::vector<double> lam(lambda, lambda + B), l1(B, 1.0), l2(lambda, lambda + B),
      ones(B, 1.0);
    double alpha1 = sigma1 / e, alpha2 = -c0;

    applyOp(P, OP_M, X, Y, B);
    applyOp(P, OP_H, X, s3.data(), B);
    axpbyBlocked(P.nDofs, B, 1.0, ones.data(), s3.data(), -1.0, lam.data(), Y, Y);
    resNew.assign(Y, Y + n);
    std::fill(res.begin(), res.end(), 0.0);
    std::fill(l2.begin(), l2.end(), alpha1 * alpha2);
    axpbyBlocked(1, B, 1.0, ones.data(), l2.data(), alpha1, lam.data(), l1.data(),
                 l2.data());
    for (size_t i = 0; i < n; ++i)
      resNew[i] = alpha1 * resNew[i];

    for (int deg = 2; deg <= m; ++deg)
      {
        const double sigma2 = 1.0 / (gamma - sigma);
        alpha1              = 2.0 * sigma2 / e;
        alpha2              = -(sigma * sigma2);
        applyOp(P, OP_MINV, resNew.data(), s1.data(), B);
        applyOp(P, OP_H, s1.data(), s2.data(), B);
        axpby(n, alpha1, s2.data(), -c0 * alpha1, resNew.data(), s1.data());
        axpby(n, 1.0, s1.data(), alpha2, res.data(), res.data());
        axpbyBlocked(P.nDofs, B, 1.0, ones.data(), res.data(), alpha1, l2.data(), Y,
                     res.data());
        axpby(B, -c0 * alpha1, l2.data(), alpha2, l1.data(), l1.data());
        axpbyBlocked(1, B, 1.0, ones.data(), l1.data(), alpha1, lam.data(), l2.data(),
                     l1.data());
        std::swap(resNew, res);
        std::swap(l1, l2);
        sigma = sigma2;
      }
    applyOp(P, OP_MINV, resNew.data(), res.data(), B);
    axpbyBlocked(P.nDofs, B, 1.0, ones.data(), res.data(), 1.0, l2.data(), X, Y);
  }

  // ------------------------------------------------- CholGS and RR
  // S = X^T (Op X), column-major n x n (computeXTransOpX,
  // src/linearAlgebra/OrthonormalizationFunctions.t.cpp:1322-1380): the GEMM
  // ('N','C') over the dof index, ascending.
  static void xtOpx(const Problem &P, int kind, double *X, u32 n, double *S)
  {
    std::vector<double> OX((size_t)P.nDofs * n);
    applyOp(P, kind, X, OX.data(), n);
    gemm('N', 'T', (int)n, (int)n, (int)P.nDofs, 1.0, X, (int)n, OX.data(),
         (int)n, 0.0, S, (int)n);
  }

  // CholeskyGramSchmidt, LAPACK branch (OrthonormalizationFunctions.t.cpp:
  // 356-415): S = X^T M X (lower), potrf('L'), trtri('L'), Xo^T = L^-1 X^T
  // with the strictly-upper part of the factor zero (as in the ScaLAPACK
  // branch :262-273).  Returns 0, or a reference OrthonormalizationErrorCode.
  int choleskyGramSchmidt(const Problem &P, double *X, u32 n, double *Xo)
  {
    if (P.nDofs < n)
      return 2; // NON_ORTHONORMALIZABLE_MULTIVECTOR
    std::vector<double> S((size_t)n * n);
    xtOpx(P, OP_M, X, n, S.data());
    if (potrfLower((int)n, S.data(), (int)n) != 0)
      return 1; // LAPACK_ERROR
    for (u32 i = 0; i < n; ++i)
      if (std::fabs(S[i + (size_t)n * i]) < 1e-14)
        return 1;
    for (u32 j = 0; j < n; ++j)
      for (u32 i = 0; i < j; ++i)
        S[i + (size_t)n * j] = 0.0;
    if (trtriLower((int)n, S.data(), (int)n) != 0)
      return 1;
    gemm('N', 'N', (int)n, (int)P.nDofs, (int)n, 1.0, S.data(), (int)n, X, (int)n,
         0.0, Xo, (int)n);
    return 0;
  }

  // RayleighRitzEigenSolver::solve(A, X, ...), LAPACK branch
  // (src/linearAlgebra/RayleighRitzEigenSolver.t.cpp:233-301): Hp = X^T H X,
  // heevd('V','L'), Xo^T = Q^T X^T.
  int rayleighRitz(const Problem &P, double *X, u32 n, double *eig, double *Xo)
  {
    std::vector<double> Hp((size_t)n * n), Q((size_t)n * n);
    xtOpx(P, OP_H, X, n, Hp.data());
    syevJacobi((int)n, Hp.data(), (int)n, eig, Q.data(), (int)n);
    gemm('T', 'N', (int)n, (int)P.nDofs, (int)n, 1.0, Q.data(), (int)n, X, (int)n,
         0.0, Xo, (int)n);
    return 0;
  }

  // ChebyshevFilteredEigenSolver::solve (src/linearAlgebra/
  // ChebyshevFilteredEigenSolver.t.cpp:194-397): block loop over `batch`
  // columns (filtered block -> X, filter input -> guess), CholGS(X -> guess, M),
  // RR(H, guess -> eig, X).  Returns a reference EigenSolverErrorCode.
  int chfsiSolve(const Problem &P, double a, double b, double c, int m,
                 bool residualFilter, u32 batch, double *eig, double *guess,
                 double *X, u32 n)
  {
    if (batch == 0 || batch > n)
      batch = n;
    const size_t N = P.nDofs;
    std::vector<double> xin, xout;
    for (u32 s = 0; s < n; s += batch)
      {
        const u32 nb = std::min(batch, n - s);
        xin.assign(N * nb, 0.0);
        xout.assign(N * nb, 0.0);
        for (size_t i = 0; i < N; ++i)
          std::memcpy(&xin[i * nb], guess + i * n + s, sizeof(double) * nb);
        if (residualFilter)
          residualChebyshevFilter(P, eig + s, xin.data(), nb, m, a, b, c, xout.data());
        else
          chebyshevFilter(P, xin.data(), nb, m, a, b, c, xout.data());
        for (size_t i = 0; i < N; ++i)
          {
            std::memcpy(X + i * n + s, &xout[i * nb], sizeof(double) * nb);
            std::memcpy(guess + i * n + s, &xin[i * nb], sizeof(double) * nb);
          }
      }
    if (choleskyGramSchmidt(P, X, n, guess) != 0)
      return 4; // CHFSI_ORTHONORMALIZATION_ERROR
    if (rayleighRitz(P, guess, n, eig, X) != 0)
      return 5; // CHFSI_RAYLEIGH_RITZ_ERROR
    return 0;
  }

  // ---------------------------------------------------------- Lanczos
  // LanczosExtremeEigenSolver::solve, non-adaptive (src/linearAlgebra/
  // LanczosExtremeEigenSolver.t.cpp:224-540) with A = H, B = lumped GLL
  // overlap, BInv = GLL overlap inverse, as KohnShamEigenSolver uses it
  // (src/ksdft/KohnShamEigenSolver.t.cpp:240-262).
  int lanczosExtreme(const Problem &P, const double *guess, int maxKrylov,
                     double *lo, double *hi, double *betaLast)
  {
    const size_t        N = P.nDofs;
    std::vector<double> g(guess, guess + N), temp(N), v(N), q(N), qPrev(N, 0.0);
    std::vector<double> alphaVec, betaVec;
    auto dot = [&](const std::vector<double> &x, const std::vector<double> &y) {
      double s = 0.0;
      for (size_t i = 0; i < N; ++i)
        s += x[i] * y[i];
      return s;
    };
    applyOp(P, OP_MLUMP, g.data(), temp.data(), 1);
    double alpha = std::sqrt(dot(g, temp));
    for (size_t i = 0; i < N; ++i)
      q[i] = (1.0 / alpha) * g[i];
    double beta = 0.0;
    for (int iter = 1; iter <= maxKrylov; ++iter)
      {
        applyOp(P, OP_H, q.data(), temp.data(), 1);
        applyOp(P, OP_MINV, temp.data(), v.data(), 1);
        alpha = dot(q, temp);
        alphaVec.push_back(alpha);
        const double nAlpha = -alpha, nBeta = -beta;
        for (size_t i = 0; i < N; ++i)
          v[i] = nAlpha * q[i] + v[i];
        for (size_t i = 0; i < N; ++i)
          v[i] = nBeta * qPrev[i] + v[i];
        applyOp(P, OP_MLUMP, v.data(), temp.data(), 1);
        beta = std::sqrt(dot(v, temp));
        betaVec.push_back(beta);
        qPrev = q;
        for (size_t i = 0; i < N; ++i)
          q[i] = (1.0 / beta) * v[i];
      }
    const int           k = (int)alphaVec.size();
    std::vector<double> T((size_t)k * k, 0.0), w(k), Z((size_t)k * k);
    for (int i = 0; i < k; ++i)
      {
        T[i + (size_t)k * i] = alphaVec[i];
        if (i + 1 < k)
          {
            T[i + 1 + (size_t)k * i] = betaVec[i];
            T[i + (size_t)k * (i + 1)] = betaVec[i];
          }
      }
    syevJacobi(k, T.data(), k, w.data(), Z.data(), k);
    *lo       = w[0];
    *hi       = w[k - 1];
    *betaLast = betaVec.back();
    return 0;
  }

  // KohnShamEigenSolver::getLinearEigenSolveResidual
  // (src/ksdft/KohnShamEigenSolver.t.cpp:574-684): || H x_i - eps_i M x_i ||_2.
  void eigenResidualNorms(const Problem &P, const double *Xin, const double *eig,
                          u32 n, double *res)
  {
    const size_t        N = P.nDofs;
    std::vector<double> X(Xin, Xin + N * n), HX(N * n), MX(N * n);
    applyOp(P, OP_H, X.data(), HX.data(), n);
    applyOp(P, OP_M, X.data(), MX.data(), n);
    for (u32 j = 0; j < n; ++j)
      {
        double s = 0.0;
        for (size_t i = 0; i < N; ++i)
          {
            const double r = -1.0 * HX[i * n + j] + eig[j] * MX[i * n + j];
            s += r * r;
          }
        res[j] = std::sqrt(s);
      }
  }

  // ------------------------------------------------------------- CG
  // CGLinearSolver::solve (src/linearAlgebra/CGLinearSolver.t.cpp:68-313)
  // with PreconditionerJacobi (src/linearAlgebra/PreconditionerJacobi.t.cpp:
  // 56-80).  Per-column alpha / beta; converged columns are frozen into the
  // returned solution.  Returns a reference LinearSolverErrorCode.
  int cgSolve(const Problem &P, int opKind, const double *diagInv,
              const double *b, double *xio, u32 B, double absTol, double relTol,
              double divTol, u32 maxIter, u32 *iters)
  {
    const size_t        N = P.nDofs, n = N * B;
    std::vector<double> x(xio, xio + n), xc(xio, xio + n), r(n), w(n), z(n), p(n);
    auto norms = [&](const std::vector<double> &v, std::vector<double> &out) {
      out.assign(B, 0.0);
      for (size_t i = 0; i < N; ++i)
        for (u32 j = 0; j < B; ++j)
          out[j] += v[i * B + j] * v[i * B + j];
      for (auto &o : out)
        o = std::sqrt(o);
    };
    auto dots = [&](const std::vector<double> &u, const std::vector<double> &v,
                    std::vector<double> &out) {
      out.assign(B, 0.0);
      for (size_t i = 0; i < N; ++i)
        for (u32 j = 0; j < B; ++j)
          out[j] += u[i * B + j] * v[i * B + j];
    };
    auto pc = [&](const std::vector<double> &in, std::vector<double> &out) {
      for (size_t i = 0; i < N; ++i)
        for (u32 j = 0; j < B; ++j)
          out[i * B + j] = diagInv[i] * in[i * B + j];
    };
    std::vector<double> bNorm, rNorm, zr, pw, zrNew;
    std::vector<double> bv(b, b + n);
    norms(bv, bNorm);
    std::vector<char> conv(B, 0);
    bool              diverged = false, allConv = false;
    int               err      = 4; // OTHER_ERROR
    u32               iter     = 0;
    for (; iter <= maxIter; ++iter)
      {
        if (iter == 0)
          {
            applyOp(P, opKind, x.data(), w.data(), B);
            for (size_t i = 0; i < n; ++i)
              r[i] = 1.0 * b[i] + (-1.0) * w[i];
            pc(r, z);
            p = z;
          }
        else
          {
            applyOp(P, opKind, p.data(), w.data(), B);
            dots(z, r, zr);
            dots(p, w, pw);
            for (size_t i = 0; i < N; ++i)
              for (u32 j = 0; j < B; ++j)
                x[i * B + j] = 1.0 * x[i * B + j] + (zr[j] / pw[j]) * p[i * B + j];
            for (size_t i = 0; i < N; ++i)
              for (u32 j = 0; j < B; ++j)
                r[i * B + j] = 1.0 * r[i * B + j] + (-zr[j] / pw[j]) * w[i * B + j];
            pc(r, z);
            dots(z, r, zrNew);
            for (size_t i = 0; i < N; ++i)
              for (u32 j = 0; j < B; ++j)
                p[i * B + j] = 1.0 * z[i * B + j] + (zrNew[j] / zr[j]) * p[i * B + j];
          }
        norms(r, rNorm);
        for (u32 j = 0; j < B; ++j)
          {
            if (rNorm[j] < std::max(absTol, bNorm[j] * relTol) && !conv[j])
              {
                err     = 0;
                conv[j] = 1;
                for (size_t i = 0; i < N; ++i)
                  xc[i * B + j] = x[i * B + j];
              }
            if (rNorm[j] > divTol && !diverged)
              {
                err      = 2; // RESIDUAL_DIVERGENCE
                diverged = true;
              }
          }
        allConv = std::all_of(conv.begin(), conv.end(), [](char c) { return c != 0; });
        if (diverged || allConv)
          break;
      }
    std::copy(xc.begin(), xc.end(), xio);
    if (iter > maxIter)
      err = 1; // FAILED_TO_CONVERGE
    *iters = iter;
    return err;
  }
} // namespace oracle
