// Host-side launchers of the libdfx kernels (gfx950).  All launches are
// asynchronous on the given stream.
#pragma once
#include "common.hpp"

namespace dfx
{
  namespace k
  {
    // ---------------------------------------------------------- vectors
    template <class T>
    void fill(hipStream_t s, View<T> x, double v);
    // x[t] = a + (b - a) * u53(splitmix64_t(seed)),  t = i * B + j
    template <class T>
    void fillRandom(hipStream_t s, View<T> x, u64 seed, double a, double b);
    template <class To, class Ti>
    void convert(hipStream_t s, View<To> dst, View<const Ti> src);
    // z = a x + b y   (z may alias x or y)
    template <class T>
    void axpby(hipStream_t s, View<T> z, double a, View<const T> x, double b,
               View<const T> y);
    // z[i,j] = a1 a[j] x[i,j] + b1 b[j] y[i,j]; a, b device arrays of length B
    template <class T>
    void axpbyBlocked(hipStream_t s, View<T> z, double a1, const double *a,
                      View<const T> x, double b1, const double *b, View<const T> y);
    // y[i,j] = d[i] x[i,j] for rows i < nRowsDiag (rows beyond are left alone)
    template <class T>
    void diagScale(hipStream_t s, View<T> y, const double *d, View<const T> x,
                   u32 nRowsDiag);
    // y_E = E x_E for the enriched rows [nC, nC + nE): E column-major nE x nE
    template <class T>
    void enrichedBlock(hipStream_t s, View<T> y, const double *E, View<const T> x,
                       u32 nC, u32 nE);
    // column dots (x_j . y_j) in FP64, deterministic two-pass; result to
    // `out` (device, length B).  `partials` must hold >= 1024 * B doubles.
    template <class T>
    void colDots(hipStream_t s, View<const T> x, View<const T> y, double *partials,
                 double *out);

    // out[j] = sum_i (a1 a[j] x_ij + b1 b[j] y_ij)^2 without forming the
    // combination: the bits of the blocked axpby followed by colDots
    // (partials: 1024 B doubles; out on the device)
    void combinationNorms(hipStream_t s, double a1, const double *a, View<const double> x,
                          double b1, const double *b, View<const double> y, double *partials,
                          double *out);
    // PCG iteration with a Jacobi (diagonal d) preconditioner, fused
    // (CGLinearSolver.t.cpp:68-313): alpha = zr / pw per column; x += alpha
    // p; r -= alpha w; z = d o r; zrrr[0, B) = z.r, zrrr[B, 2B) = r.r, the
    // bits of the separate launches.  partials: 2048 B doubles; xs (may be
    // null, contiguous like x) receives a copy of the new x
    void cgUpdate(hipStream_t s, View<double> x, View<double> r, View<double> z,
                  View<const double> p, View<const double> w, const double *d,
                  const double *zr, const double *pw, double *partials, double *zrrr,
                  double *xs = nullptr);
    // p = z + (zrNew / zr) p per column, one launch (contiguous p, z)
    void cgDirection(hipStream_t s, View<double> p, View<const double> z, const double *zrNew,
                     const double *zr);
    // Plain Chebyshev degree step, fused (ChebyshevFilter.t.cpp:96-129):
    //   t  = Minv hx  (diag rows < nC: dinv o hx; enriched rows: given in tE)
    //   s1 = alpha t + beta xk;  xprev = s1 + gamma xprev
    // xnext aliases xprev.
    template <class T>
    void chebStep(hipStream_t s, View<T> xprev, View<const T> hx, const double *dinv,
                  View<const T> tE, u32 nC, View<const T> xk, double alpha,
                  double beta, double gamma);
    // Residual Chebyshev degree step, fused (ChebyshevFilter.t.cpp:360-397):
    //   s1  = alpha1 s2 + (-c alpha1) rn
    //   res = s1 + alpha2 res
    //   res = res + alpha1 mu[j] r0
    // coef: device array {alpha1, alpha2, c} of the degree (read by the
    // kernel, so a captured launch stays valid when the filter bounds change)
    template <class T>
    void resStep(hipStream_t s, View<T> res, View<const T> s2, View<const T> rn,
                 View<const T> r0, const double *mu, const double *coef);
    // content hash of nWords 64-bit words (synchronizes s); partials holds
    // HASH_BLOCKS words.  Equal contents give equal hashes; a change of any
    // word changes it but for a 2^-64 chance
    constexpr u64 HASH_BLOCKS = 2048;
    u64 hashWords(hipStream_t s, const void *x, u64 nWords, u64 *partials);
    // resStep on the listed rows only (res, rn, r0 share one leading
    // dimension): the rows a cell apply with the fused step leaves over
    template <class T>
    void resFix(hipStream_t s, View<T> res, View<const T> s2, View<const T> rn,
                View<const T> r0, const u32 *rows, u32 nRows, const double *mu,
                const double *coef);
    // residual filter set-up from the FP64 seed Y: r0 = (T) Y, rn = (T)(alpha1 Y
    // + 0 Y), res = 0 (r0, rn, res share one leading dimension)
    template <class T>
    void resSeed(hipStream_t s, View<T> r0, View<T> rn, View<T> res, View<const double> Y,
                 double alpha1);
    // resSeed from AX and BX (kept products): Y = AX - lambda o BX formed in
    // registers (the blocked axpby's element), the same bits as the blocked
    // axpby into Y followed by resSeed
    template <class T>
    void resSeedFrom(hipStream_t s, View<T> r0, View<T> rn, View<T> res, View<const double> AX,
                     View<const double> BX, const double *ones, const double *lam,
                     double alpha1);
    // z[i,j] = a1 a[j] (double) x[i,j] + b1 b[j] y[i,j] (x in T; a, b device
    // arrays of length B): the residual filter's result
    template <class T>
    void resFinal(hipStream_t s, View<double> z, double a1, const double *a, View<const T> x,
                  double b1, const double *b, View<const double> y);
    // the same with x = D r on the first nC rows (D diagonal, the product
    // formed in T as the diagonal apply does) and x = xE on the others
    template <class T>
    void resFinalDiag(hipStream_t s, View<double> z, double a1, const double *a,
                      View<const T> r, const double *d, u32 nC, View<const T> xE, double b1,
                      const double *b, View<const double> y);

    // -------------------------------------------------------- constraints
    template <class T>
    void p2c(hipStream_t s, const dfx_constraints *c, View<T> x);
    template <class T>
    void c2p(hipStream_t s, const dfx_constraints *c, View<T> y);
    template <class T>
    void setRows(hipStream_t s, const dfx_constraints *c, View<T> x, double v);

    // -------------------------------------------------- gather / scatter
    // xCell[slot, j] = x[map[slot], j]
    template <class T>
    void gather(hipStream_t s, const u32 *map, u64 nSlots, View<const T> x, View<T> xCell);
    // y[d, j] += sum over the dof's slots (cell order) of yCell[slot, j]
    template <class T>
    void scatterAdd(hipStream_t s, const u32 *dofSlotOffsets, const u32 *dofSlots,
                    u32 nDofs, View<const T> yCell, View<T> y);

    // ------------------------------------------------------- cell apply
    // one colour / size-class launch of the fused gather -> MFMA -> scatter
    // dsc (optional): the launch gathers diag o X for the classical rows
    // (M^-1 folded into the gather) and the enriched rows from xe
    // nt (optional): the column-chunk width in 16-column tiles (1, 2, 4, 5 or
    // 8) instead of the tuned or built-in choice
    // the residual Chebyshev step fused into a folded cell apply: on the
    // classical rows whose last touch is the launch's (and which are not
    // constrained), res = resStep(y, x, res, r0) is stored instead of y; the
    // other rows of y are stored as usual and finished by resFix
    template <class T>
    struct ResFuse
    {
      T            *res;  // leading dimension that of x
      const T      *r0;   // leading dimension that of x
      const double *mu;   // one per column
      const double *coef; // {alpha1, alpha2, c}
    };
    // rf (optional, with dsc only): the residual step fused into the launch
    template <class T>
    void cellApply(hipStream_t s, const dfx_op *op, const Launch &L, const T *frag,
                   View<const T> x, View<T> y, T *yE, const double *dsc = nullptr,
                   const T *xe = nullptr, int nt = 0, const ResFuse<T> *rf = nullptr);
    // the built-in column-chunk width for B columns and a class of padded K
    // up to kpadMax
    template <class T>
    int builtinNT(u32 B, u32 kpadMax);
    // out[s] = d[map[s]] (map[s] < nC) else 1, for every slot of a cell->dof map
    void gatherDiag(hipStream_t s, const double *d, const u32 *map, u64 n, u32 nC,
                    double *out);
    // whether launches on (x, y) take the buffer-addressed kernel
    template <class T>
    bool cellBufferPath(const dfx_op *op, View<const T> x, View<const T> y);
    // y[nC + g, j] = sum over the enriched dof's slots of yE[slot, j]
    template <class T>
    void enrichedReduce(hipStream_t s, const dfx_op *op, const T *yE, View<T> y);
    // out[mapOff_e + i] = A_e(i, i) from the FP64 fragment-ordered matrices
    // (caller slot order)
    void cellDiag(hipStream_t s, const dfx_op *op, double *out);
    // repack raw column-major cell matrices (device) into the fragment layout
    template <class T>
    void repackCells(hipStream_t s, const dfx_op *op, const double *raw,
                     const u64 *rawOffsets, T *frag);

    // --------------------------------------------------------- dense
    // S (n x n, column-major, FP64) = X^T Y over all N rows; deterministic
    template <class T>
    void gram(hipStream_t s, View<const T> X, View<const T> Y, double *partials,
              size_t partialsCap, double *S);
    // Xo = X Q  (Q n x n column-major FP64, device); Xo must not alias X;
    // qrScratch holds n * 16 * ceil(n / 16) doubles.  With a rocBLAS handle
    // (bound to stream s) an FP64 rotation is one library GEMM
    template <class T>
    void rotate(hipStream_t s, View<const T> X, const double *Q, View<T> Xo,
                double *qrScratch, rocblas_handle blas = nullptr);
    size_t gramPartialsNeeded(u32 N, u32 n);
  } // namespace k
} // namespace dfx
