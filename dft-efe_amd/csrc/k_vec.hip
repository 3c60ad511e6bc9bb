// Streaming kernels of libdfx: vector updates of the Chebyshev recurrences,
// column reductions, constraint distribution and the stand-alone cell
// gather / scatter.  All HBM-bound; written as grid-stride loops over
// N x B views with a leading dimension (column blocks of a wider multivector
// are addressed in place, without the reference's per-dof copy loops,
// src/linearAlgebra/ChebyshevFilteredEigenSolver.t.cpp:246-333).
#include <vector>

#include "kernels.hpp"
#include "res_update.hpp"

namespace dfx
{
  namespace k
  {
    namespace
    {
      constexpr int TPB = 256;

      inline unsigned gridFor(u64 n)
      {
        u64 g = (n + TPB - 1) / TPB;
        if (g > 65536ull * 8)
          g = 65536ull * 8;
        return (unsigned)(g ? g : 1);
      }

      // flat index -> (row, col) of an N x B view
      struct RC
      {
        u32 i, j;
      };
      __device__ inline RC rc(u64 t, u32 B)
      {
        if ((t >> 32) == 0) // 32-bit division while the flat index fits
          {
            const u32 tt = (u32)t, i = tt / B;
            return {i, tt - i * B};
          }
        u32 i = (u32)(t / B);
        return {i, (u32)(t - (u64)i * B)};
      }

      template <class T>
      __global__ void fillK(T *x, u32 N, u32 B, u32 ld, T v)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r = rc(t, B);
            x[(u64)r.i * ld + r.j] = v;
          }
      }

      // SplitMix64 (Steele, Lea, Flood 2014): output t is the mix of
      // seed + (t + 1) * golden gamma, so every element is independent
      template <class T>
      __global__ void fillRandomK(T *x, u32 N, u32 B, u32 ld, u64 seed, double a, double b)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            u64 z = seed + (t + 1) * 0x9E3779B97F4A7C15ULL;
            z     = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z     = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z     = z ^ (z >> 31);
            const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
            RC r = rc(t, B);
            x[(u64)r.i * ld + r.j] = (T)(a + (b - a) * u);
          }
      }

      template <class To, class Ti>
      __global__ void convertK(To *y, u32 ldy, const Ti *x, u32 ldx, u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r = rc(t, B);
            y[(u64)r.i * ldy + r.j] = (To)x[(u64)r.i * ldx + r.j];
          }
      }

      template <class T>
      __global__ void axpbyK(T *z, u32 ldz, T a, const T *x, u32 ldx, T b,
                             const T *y, u32 ldy, u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r  = rc(t, B);
            T  xv = x[(u64)r.i * ldx + r.j], yv = y[(u64)r.i * ldy + r.j];
            z[(u64)r.i * ldz + r.j] = a * xv + b * yv;
          }
      }

      // one element of the blocked axpby (shared with the residual filter's
      // result and the fused CG update): the multiply-add written out, so
      // that the rounding does not depend on which operands the compiler
      // sees as constants
      template <class T>
      __device__ inline T axpbyOne(T a1, double a, T xv, T b1, double b, T yv)
      {
        return fma(b1 * (T)b, yv, a1 * (T)a * xv);
      }

      template <class T>
      __global__ void axpbyBlockedK(T *z, u32 ldz, T a1, const double *a,
                                    const T *x, u32 ldx, T b1, const double *b,
                                    const T *y, u32 ldy, u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r  = rc(t, B);
            T  xv = x[(u64)r.i * ldx + r.j], yv = y[(u64)r.i * ldy + r.j];
            z[(u64)r.i * ldz + r.j] = axpbyOne<T>(a1, a[r.j], xv, b1, b[r.j], yv);
          }
      }

      // PCG search direction p = z + beta p with beta = zrNew / zr per column
      // (the division of the separate beta launch, then the blocked axpby's
      // element with unit coefficients)
      __global__ void cgDirectionK(double *p, const double *z, const double *zrNew,
                                   const double *zr, u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC           r    = rc(t, B);
            const double beta = zrNew[r.j] / zr[r.j];
            const u64    o    = (u64)r.i * B + r.j;
            p[o]              = axpbyOne<double>(1.0, 1.0, z[o], 1.0, beta, p[o]);
          }
      }

      template <class T>
      __global__ void diagScaleK(T *y, u32 ldy, const double *d, const T *x,
                                 u32 ldx, u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r = rc(t, B);
            y[(u64)r.i * ldy + r.j] = (T)d[r.i] * x[(u64)r.i * ldx + r.j];
          }
      }

      // y_E[i, j] = sum_k E[i + nE k] x_E[k, j]: one workgroup per row i and
      // 64 columns; four k lanes per column each add k = lane, lane + 4, ...
      // in ascending order, combined in lane order (fixed, deterministic)
      template <class T>
      __global__ __launch_bounds__(256) void enrBlockK(T *y, u32 ldy, const double *E,
                                                       const T *x, u32 ldx, u32 nC, u32 nE,
                                                       u32 B)
      {
        __shared__ T part[4][64];
        const u32 i = blockIdx.x, cj = threadIdx.x % 64, kl = threadIdx.x / 64;
        const u32 j = blockIdx.y * 64 + cj;
        T         acc = 0;
        if (j < B)
          {
            // eight k of a lane loaded before they are added (in the same
            // ascending order), so that the loads are in flight together
            u32 kk = kl;
            for (; kk + 28 < nE; kk += 32)
              {
                T ev[8], xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                  {
                    const u32 k = kk + 4u * u;
                    ev[u]       = (T)E[i + (u64)nE * k];
                    xv[u]       = x[(u64)(nC + k) * ldx + j];
                  }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                  acc += ev[u] * xv[u];
              }
            for (; kk < nE; kk += 4)
              acc += (T)E[i + (u64)nE * kk] * x[(u64)(nC + kk) * ldx + j];
          }
        part[kl][cj] = acc;
        __syncthreads();
        if (kl == 0 && j < B)
          y[(u64)(nC + i) * ldy + j] = ((part[0][cj] + part[1][cj]) + part[2][cj]) + part[3][cj];
      }

      // pass 1 of the column dots: block b owns rows [b*rows, (b+1)*rows);
      // threads are C columns x R row-lanes
      template <class T, int C>
      __global__ void colDotsK(const T *x, u32 ldx, const T *y, u32 ldy, u32 N,
                               u32 B, u32 rowsPerBlock, double *partials)
      {
        constexpr int R  = TPB / C;
        const int     tx = threadIdx.x % C, ty = threadIdx.x / C;
        __shared__ double red[TPB];
        const u32 r0 = blockIdx.x * rowsPerBlock;
        const u32 r1 = min(N, r0 + rowsPerBlock);
        for (u32 j0 = 0; j0 < B; j0 += C)
          {
            const u32 j   = j0 + tx;
            double    acc = 0.0;
            if (j < B)
              for (u32 i = r0 + ty; i < r1; i += R)
                acc = fma((double)x[(u64)i * ldx + j], (double)y[(u64)i * ldy + j], acc);
            red[threadIdx.x] = acc;
            __syncthreads();
            if (ty == 0 && j < B)
              {
                double s = 0.0;
                for (int q = 0; q < R; ++q)
                  s += red[q * C + tx];
                partials[(u64)blockIdx.x * B + j] = s;
              }
            __syncthreads();
          }
      }

      // The PCG update of one iteration with a Jacobi preconditioner, fused:
      //   x += alpha p, r -= alpha w, z = d o r, and the partial sums of z.r
      //   and r.r (alpha = zr / pw per column, formed here as cgAlphaK does).
      // Rows, lanes and sums follow colDotsK exactly (same blocks of rows,
      // same lane order, same combine), and each element is the blocked
      // axpby's / the diagonal scaling's expression, so every value is the
      // separate launches' bit for bit.  partials: nBlocks x 2B (z.r in
      // columns [0, B), r.r in [B, 2B)); N x B arrays with leading dimension B
      template <int C>
      __global__ void cgUpdateK(double *x, double *r, double *z, const double *p,
                                const double *w, const double *d, const double *zr,
                                const double *pw, u32 N, u32 B, u32 rowsPerBlock,
                                double *partials, double *xs)
      {
        constexpr int R  = TPB / C;
        const int     tx = threadIdx.x % C, ty = threadIdx.x / C;
        __shared__ double red[2][TPB];
        const u32 r0 = blockIdx.x * rowsPerBlock;
        const u32 r1 = min(N, r0 + rowsPerBlock);
        for (u32 j0 = 0; j0 < B; j0 += C)
          {
            const u32 j     = j0 + tx;
            double    accZ  = 0.0, accR = 0.0;
            if (j < B)
              {
                const double alpha = zr[j] / pw[j], nalpha = -zr[j] / pw[j];
                for (u32 i = r0 + ty; i < r1; i += R)
                  {
                    const u64    o  = (u64)i * B + j;
                    const double xn = axpbyOne<double>(1.0, 1.0, x[o], 1.0, alpha, p[o]);
                    const double rn = axpbyOne<double>(1.0, 1.0, r[o], 1.0, nalpha, w[o]);
                    const double zn = d[i] * rn;
                    x[o]            = xn;
                    if (xs) // a copy of the new x (launch-uniform)
                      xs[o] = xn;
                    r[o]            = rn;
                    z[o]            = zn;
                    accZ = fma(zn, rn, accZ);
                    accR = fma(rn, rn, accR);
                  }
              }
            red[0][threadIdx.x] = accZ;
            red[1][threadIdx.x] = accR;
            __syncthreads();
            if (ty == 0 && j < B)
              {
                double sz = 0.0, sr = 0.0;
                for (int q = 0; q < R; ++q)
                  {
                    sz += red[0][q * C + tx];
                    sr += red[1][q * C + tx];
                  }
                partials[(u64)blockIdx.x * 2 * B + j]     = sz;
                partials[(u64)blockIdx.x * 2 * B + B + j] = sr;
              }
            __syncthreads();
          }
      }

      // partial sums of |a1 a_j x_ij + b1 b_j y_ij|^2 per column without
      // writing the combination: colDotsK's rows, lanes and combine and the
      // blocked axpby's element, so the norms are those of the separate
      // launches bit for bit (N x B arrays with leading dimension B)
      template <int C>
      __global__ void combNormsK(double a1, const double *a, const double *x, double b1,
                                 const double *b, const double *y, u32 N, u32 B,
                                 u32 rowsPerBlock, double *partials)
      {
        constexpr int R  = TPB / C;
        const int     tx = threadIdx.x % C, ty = threadIdx.x / C;
        __shared__ double red[TPB];
        const u32 r0 = blockIdx.x * rowsPerBlock;
        const u32 r1 = min(N, r0 + rowsPerBlock);
        for (u32 j0 = 0; j0 < B; j0 += C)
          {
            const u32 j   = j0 + tx;
            double    acc = 0.0;
            if (j < B)
              {
                const double aj = a[j], bj = b[j];
                for (u32 i = r0 + ty; i < r1; i += R)
                  {
                    const u64    o = (u64)i * B + j;
                    const double v = axpbyOne<double>(a1, aj, x[o], b1, bj, y[o]);
                    acc            = fma(v, v, acc);
                  }
              }
            red[threadIdx.x] = acc;
            __syncthreads();
            if (ty == 0 && j < B)
              {
                double s = 0.0;
                for (int q = 0; q < R; ++q)
                  s += red[q * C + tx];
                partials[(u64)blockIdx.x * B + j] = s;
              }
            __syncthreads();
          }
      }

      // one workgroup per column: thread t adds partials t, t + 256, ... in
      // order, then a fixed-shape tree over the 256 sums (deterministic; the
      // partial count depends on N only, so a column's result does not
      // depend on how many columns share the call)
      __global__ __launch_bounds__(256) void colDotsFinalK(const double *partials, u32 nBlocks,
                                                           u32 B, double *out)
      {
        __shared__ double red[256];
        const u32 j = blockIdx.x;
        double    s = 0.0;
        for (u32 b = threadIdx.x; b < nBlocks; b += 256)
          s += partials[(u64)b * B + j];
        red[threadIdx.x] = s;
        __syncthreads();
        for (u32 w = 128; w > 0; w >>= 1)
          {
            if (threadIdx.x < w)
              red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
          }
        if (threadIdx.x == 0)
          out[j] = red[0];
      }

      template <class T>
      __global__ void chebStepK(T *xp, u32 ldp, const T *hx, u32 ldh,
                                const double *dinv, const T *tE, u32 ldt, u32 nC,
                                const T *xk, u32 ldk, T alpha, T beta, T gamma,
                                u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r = rc(t, B);
            T  tv;
            if (r.i < nC)
              tv = (T)dinv[r.i] * hx[(u64)r.i * ldh + r.j];
            else
              tv = tE[(u64)r.i * ldt + r.j];
            const T s1 = alpha * tv + beta * xk[(u64)r.i * ldk + r.j];
            T      &o  = xp[(u64)r.i * ldp + r.j];
            o          = (T)1 * s1 + gamma * o;
          }
      }

      template <class T>
      __global__ void resStepK(T *res, u32 ldr, const T *s2, u32 lds, const T *rn,
                               u32 ldn, const T *r0, u32 ld0, const double *mu,
                               const double *coef, u32 N, u32 B)
      {
        const u64 n      = (u64)N * B;
        const T   alpha1 = (T)coef[0], alpha2 = (T)coef[1], c = (T)coef[2];
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r = rc(t, B);
            T &o = res[(u64)r.i * ldr + r.j];
            o    = resUpdate<T>(s2[(u64)r.i * lds + r.j], rn[(u64)r.i * ldn + r.j], o,
                                r0[(u64)r.i * ld0 + r.j], mu[r.j], alpha1, alpha2, c);
          }
      }

      // resStepK on contiguous N x B arrays (every leading dimension B) with
      // B a multiple of the 16-byte vector: one vector of V columns per
      // thread, the same arithmetic per element
      template <class T>
      __global__ void resStepVecK(T *__restrict__ res, const T *__restrict__ s2,
                                  const T *__restrict__ rn, const T *__restrict__ r0,
                                  const double *__restrict__ mu,
                                  const double *__restrict__ coef, u32 nVec, u32 B)
      {
        constexpr int V = 16 / sizeof(T);
        typedef T vt __attribute__((ext_vector_type(V)));
        const T alpha1 = (T)coef[0], alpha2 = (T)coef[1], c = (T)coef[2];
        for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < nVec;
             t += gridDim.x * blockDim.x)
          {
            const u32 j0 = (t * V) % B;
            const vt  a  = reinterpret_cast<const vt *>(s2)[t];
            const vt  b  = reinterpret_cast<const vt *>(rn)[t];
            const vt  z  = reinterpret_cast<const vt *>(r0)[t];
            vt        o  = reinterpret_cast<const vt *>(res)[t];
#pragma unroll
            for (int e = 0; e < V; ++e)
              o[e] = resUpdate<T>(a[e], b[e], o[e], z[e], mu[j0 + e], alpha1, alpha2, c);
            reinterpret_cast<vt *>(res)[t] = o;
          }
      }

      // content hash of n 64-bit words: the wrapping sum over i of
      // mix(x[i] + (i + 1) golden), mix the splitmix64 finalizer; a sum, so
      // the value does not depend on the order the blocks add up in
      __device__ inline u64 mix64(u64 z)
      {
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
      }
      __global__ __launch_bounds__(256) void hashWordsK(const u64 *__restrict__ x, u64 n,
                                                        u64 *__restrict__ partials)
      {
        __shared__ u64 sm[256];
        u64            acc = 0;
        for (u64 i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (u64)gridDim.x * 256)
          acc += mix64(x[i] + (i + 1) * 0x9e3779b97f4a7c15ull);
        sm[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1)
          {
            if ((int)threadIdx.x < w)
              sm[threadIdx.x] += sm[threadIdx.x + w];
            __syncthreads();
          }
        if (threadIdx.x == 0)
          partials[blockIdx.x] = sm[0];
      }

      // the residual step on the rows a fused cell apply leaves to it (listed
      // in `rows`: constrained, enriched and cell-less rows), after the
      // enriched reduction and the child->parent step have finished y there
      template <class T>
      __global__ void resFixK(T *res, const T *s2, u32 lds, const T *rn, const T *r0, u32 ld,
                              const u32 *__restrict__ rows, u32 nRows,
                              const double *mu, const double *coef, u32 B)
      {
        const u64 n      = (u64)nRows * B;
        const T   alpha1 = (T)coef[0], alpha2 = (T)coef[1], c = (T)coef[2];
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            const RC  r = rc(t, B);
            const u32 i = rows[r.i];
            const u64 o = (u64)i * ld + r.j;
            res[o]      = resUpdate<T>(s2[(u64)i * lds + r.j], rn[o], res[o], r0[o], mu[r.j],
                                       alpha1, alpha2, c);
          }
      }

      // residual filter set-up from the FP64 seed Y (ChebyshevFilter.t.cpp:
      // 330-359): r0 = Y, rn = a Y + b Y (a = alpha1, b = 0, the reference's
      // axpby), res = 0, in T; the same arithmetic as the separate
      // convert / axpby / convert / fill launches
      template <class T>
      __device__ inline void resSeedOne(double y, double a, double b, T &r0, T &rn, T &res)
      {
        r0  = (T)y;
        rn  = (T)fma(a, y, b * y);
        res = (T)0.0;
      }

      template <class T>
      __global__ void resSeedK(T *r0, T *rn, T *res, u32 ld, const double *Y, u32 ldy,
                               double a, double b, u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC        r = rc(t, B);
            const u64 o = (u64)r.i * ld + r.j;
            resSeedOne<T>(Y[(u64)r.i * ldy + r.j], a, b, r0[o], rn[o], res[o]);
          }
      }

      // the seed from kept products: Y = AX - lambda o BX (the blocked
      // axpby's element) is formed in registers and converted as resSeedK
      // does, without being written
      template <class T>
      __global__ void resSeedFromK(T *r0, T *rn, T *res, u32 ld, const double *AX,
                                   const double *BX, u32 ldp, const double *ones,
                                   const double *lam, double a, double b, u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC           r = rc(t, B);
            const u64    p = (u64)r.i * ldp + r.j, o = (u64)r.i * ld + r.j;
            const double y = axpbyOne<double>(1.0, ones[r.j], AX[p], -1.0, lam[r.j], BX[p]);
            resSeedOne<T>(y, a, b, r0[o], rn[o], res[o]);
          }
      }

      // residual filter result (ChebyshevFilter.t.cpp:424-441): z = a1 a[j]
      // (double) x + b1 b[j] y with x the last residual in T; the arithmetic of
      // a conversion followed by the FP64 blocked axpby
      template <class T>
      __global__ void resFinalK(double *z, u32 ldz, double a1, const double *a, const T *x,
                                u32 ldx, double b1, const double *b, const double *y, u32 ldy,
                                u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC           r  = rc(t, B);
            const double xv = (double)x[(u64)r.i * ldx + r.j], yv = y[(u64)r.i * ldy + r.j];
            z[(u64)r.i * ldz + r.j] = axpbyOne<double>(a1, a[r.j], xv, b1, b[r.j], yv);
          }
      }

      // resFinalK with the diagonal apply folded in: x[i, j] = (T)(d[i] r[i, j])
      // on the first nC rows (diagScaleK's product), xE[i, j] below them
      template <class T>
      __global__ void resFinalDiagK(double *z, u32 ldz, double a1, const double *a,
                                    const T *r, const double *d, u32 nC, const T *xE, u32 ldx,
                                    double b1, const double *b, const double *y, u32 ldy,
                                    u32 N, u32 B)
      {
        const u64 n = (u64)N * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC        rr = rc(t, B);
            const u64 o  = (u64)rr.i * ldx + rr.j;
            const T   xt = rr.i < nC ? (T)d[rr.i] * r[o] : xE[o];
            const double xv = (double)xt, yv = y[(u64)rr.i * ldy + rr.j];
            z[(u64)rr.i * ldz + rr.j] = axpbyOne<double>(a1, a[rr.j], xv, b1, b[rr.j], yv);
          }
      }

      // ------------------------------------------------ constraints
      // ConstraintsInternal.cpp:38-107 restated per (row, column): the row's
      // value is inhom + sum_j w_j x[col_j] accumulated with fma in entry order.
      template <class T>
      __global__ void p2cK(T *x, u32 ld, const u32 *rows, const u32 *rowOff,
                           const u32 *cols, const double *w, const double *inhom,
                           u32 nRows, u32 B)
      {
        const u64 n = (u64)nRows * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC  r  = rc(t, B);
            T   v  = (T)inhom[r.i];
            u32 e0 = rowOff[r.i], e1 = rowOff[r.i + 1];
            for (u32 e = e0; e < e1; ++e)
              v = fma((T)w[e], x[(u64)cols[e] * ld + r.j], v);
            x[(u64)rows[r.i] * ld + r.j] = v;
          }
      }

      // ConstraintsInternal.cpp:113-170 without atomics: each parent gathers
      // w * x[row] from its children in the rows' order, which is the order
      // of the reference's sequential loop, so the result is identical.
      template <class T>
      __global__ void c2pParentsK(T *y, u32 ld, const u32 *par, const u32 *parOff,
                                  const u32 *parRowDof, const double *parW,
                                  u32 nPar, u32 B)
      {
        const u64 n = (u64)nPar * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC  r   = rc(t, B);
            T  &dst = y[(u64)par[r.i] * ld + r.j];
            T   acc = dst;
            u32 e0 = parOff[r.i], e1 = parOff[r.i + 1];
            for (u32 e = e0; e < e1; ++e)
              acc = fma((T)parW[e], y[(u64)parRowDof[e] * ld + r.j], acc);
            dst = acc;
          }
      }

      template <class T>
      __global__ void setRowsK(T *x, u32 ld, const u32 *rows, u32 nRows, u32 B, T v)
      {
        const u64 n = (u64)nRows * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r = rc(t, B);
            x[(u64)rows[r.i] * ld + r.j] = v;
          }
      }

      // ------------------------------------------------ gather / scatter
      template <class T>
      __global__ void gatherK(const u32 *map, u64 nSlots, const T *x, u32 ldx,
                              T *xc, u32 ldc, u32 B)
      {
        const u64 n = nSlots * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            const u64 s = t / B, j = t - s * B;
            xc[s * ldc + j] = x[(u64)map[s] * ldx + j];
          }
      }

      template <class T>
      __global__ void scatterAddK(const u32 *off, const u32 *slots, u32 nDofs,
                                  const T *yc, u32 ldc, T *y, u32 ldy, u32 B)
      {
        const u64 n = (u64)nDofs * B;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            RC r   = rc(t, B);
            T &dst = y[(u64)r.i * ldy + r.j];
            T  acc = dst;
            for (u32 e = off[r.i]; e < off[r.i + 1]; ++e)
              acc += yc[(u64)slots[e] * ldc + r.j];
            dst = acc;
          }
      }
    } // namespace

    // ==================================================== launchers
    template <class T>
    void fill(hipStream_t s, View<T> x, double v)
    {
      const u64 n = (u64)x.N * x.B;
      if (!n)
        return;
      fillK<T><<<gridFor(n), TPB, 0, s>>>(x.p, x.N, x.B, x.ld, (T)v);
      checkLaunch("fill");
    }

    template <class T>
    void fillRandom(hipStream_t s, View<T> x, u64 seed, double a, double b)
    {
      const u64 n = (u64)x.N * x.B;
      if (!n)
        return;
      fillRandomK<T><<<gridFor(n), TPB, 0, s>>>(x.p, x.N, x.B, x.ld, seed, a, b);
      checkLaunch("fillRandom");
    }

    template <class To, class Ti>
    void convert(hipStream_t s, View<To> d, View<const Ti> x)
    {
      const u64 n = (u64)x.N * x.B;
      if (!n)
        return;
      convertK<To, Ti><<<gridFor(n), TPB, 0, s>>>(d.p, d.ld, x.p, x.ld, x.N, x.B);
      checkLaunch("convert");
    }

    template <class T>
    void axpby(hipStream_t s, View<T> z, double a, View<const T> x, double b,
               View<const T> y)
    {
      const u64 n = (u64)z.N * z.B;
      if (!n)
        return;
      axpbyK<T><<<gridFor(n), TPB, 0, s>>>(z.p, z.ld, (T)a, x.p, x.ld, (T)b, y.p,
                                           y.ld, z.N, z.B);
      checkLaunch("axpby");
    }

    template <class T>
    void axpbyBlocked(hipStream_t s, View<T> z, double a1, const double *a,
                      View<const T> x, double b1, const double *b, View<const T> y)
    {
      const u64 n = (u64)z.N * z.B;
      if (!n)
        return;
      axpbyBlockedK<T><<<gridFor(n), TPB, 0, s>>>(z.p, z.ld, (T)a1, a, x.p, x.ld,
                                                  (T)b1, b, y.p, y.ld, z.N, z.B);
      checkLaunch("axpbyBlocked");
    }

    template <class T>
    void diagScale(hipStream_t s, View<T> y, const double *d, View<const T> x,
                   u32 nRowsDiag)
    {
      const u64 n = (u64)nRowsDiag * x.B;
      if (!n)
        return;
      diagScaleK<T><<<gridFor(n), TPB, 0, s>>>(y.p, y.ld, d, x.p, x.ld, nRowsDiag, x.B);
      checkLaunch("diagScale");
    }

    template <class T>
    void enrichedBlock(hipStream_t s, View<T> y, const double *E, View<const T> x,
                       u32 nC, u32 nE)
    {
      const u64 n = (u64)nE * x.B;
      if (!n)
        return;
      enrBlockK<T><<<dim3(nE, (x.B + 63) / 64), 256, 0, s>>>(y.p, y.ld, E, x.p, x.ld, nC, nE,
                                                             x.B);
      checkLaunch("enrichedBlock");
    }

    template <class T>
    void colDots(hipStream_t s, View<const T> x, View<const T> y, double *partials,
                 double *out)
    {
      const u32 B = x.B;
      if (!B)
        return;
      u32 nBlocks = (x.N + 255) / 256;
      if (nBlocks > 1024)
        nBlocks = 1024;
      if (nBlocks == 0)
        nBlocks = 1;
      const u32 rows = (x.N + nBlocks - 1) / nBlocks;
      if (B <= 16)
        colDotsK<T, 16><<<nBlocks, TPB, 0, s>>>(x.p, x.ld, y.p, y.ld, x.N, B, rows, partials);
      else if (B <= 32)
        colDotsK<T, 32><<<nBlocks, TPB, 0, s>>>(x.p, x.ld, y.p, y.ld, x.N, B, rows, partials);
      else
        colDotsK<T, 64><<<nBlocks, TPB, 0, s>>>(x.p, x.ld, y.p, y.ld, x.N, B, rows, partials);
      checkLaunch("colDots");
      colDotsFinalK<<<B, 256, 0, s>>>(partials, nBlocks, B, out);
      checkLaunch("colDotsFinal");
    }

    void combinationNorms(hipStream_t s, double a1, const double *a, View<const double> x,
                          double b1, const double *b, View<const double> y, double *partials,
                          double *out)
    {
      const u32 B = x.B, N = x.N;
      if (!B)
        return;
      DFX_REQUIRE(x.ld == B && y.ld == B && y.N == N && y.B == B,
                  "combination norms: contiguous multivectors of one shape expected");
      u32 nBlocks = (N + 255) / 256;
      if (nBlocks > 1024)
        nBlocks = 1024;
      if (nBlocks == 0)
        nBlocks = 1;
      const u32 rows = (N + nBlocks - 1) / nBlocks;
      if (B <= 16)
        combNormsK<16><<<nBlocks, TPB, 0, s>>>(a1, a, x.p, b1, b, y.p, N, B, rows, partials);
      else if (B <= 32)
        combNormsK<32><<<nBlocks, TPB, 0, s>>>(a1, a, x.p, b1, b, y.p, N, B, rows, partials);
      else
        combNormsK<64><<<nBlocks, TPB, 0, s>>>(a1, a, x.p, b1, b, y.p, N, B, rows, partials);
      checkLaunch("combNorms");
      colDotsFinalK<<<B, 256, 0, s>>>(partials, nBlocks, B, out);
      checkLaunch("colDotsFinal");
    }

    void cgUpdate(hipStream_t s, View<double> x, View<double> r, View<double> z,
                  View<const double> p, View<const double> w, const double *d,
                  const double *zr, const double *pw, double *partials, double *zrrr,
                  double *xs)
    {
      const u32 B = x.B, N = x.N;
      if (!B)
        return;
      DFX_REQUIRE(x.ld == B && r.ld == B && z.ld == B && p.ld == B && w.ld == B,
                  "CG update: contiguous multivectors expected");
      // the partition of colDots
      u32 nBlocks = (N + 255) / 256;
      if (nBlocks > 1024)
        nBlocks = 1024;
      if (nBlocks == 0)
        nBlocks = 1;
      const u32 rows = (N + nBlocks - 1) / nBlocks;
      if (B <= 16)
        cgUpdateK<16><<<nBlocks, TPB, 0, s>>>(x.p, r.p, z.p, p.p, w.p, d, zr, pw, N, B, rows,
                                              partials, xs);
      else if (B <= 32)
        cgUpdateK<32><<<nBlocks, TPB, 0, s>>>(x.p, r.p, z.p, p.p, w.p, d, zr, pw, N, B, rows,
                                              partials, xs);
      else
        cgUpdateK<64><<<nBlocks, TPB, 0, s>>>(x.p, r.p, z.p, p.p, w.p, d, zr, pw, N, B, rows,
                                              partials, xs);
      checkLaunch("cgUpdate");
      colDotsFinalK<<<2 * B, 256, 0, s>>>(partials, nBlocks, 2 * B, zrrr);
      checkLaunch("colDotsFinal");
    }

    void cgDirection(hipStream_t s, View<double> p, View<const double> z, const double *zrNew,
                     const double *zr)
    {
      const u64 n = (u64)p.N * p.B;
      if (!n)
        return;
      DFX_REQUIRE(p.ld == p.B && z.ld == z.B && z.N == p.N && z.B == p.B,
                  "CG direction: contiguous multivectors of one shape expected");
      cgDirectionK<<<gridFor(n), TPB, 0, s>>>(p.p, z.p, zrNew, zr, p.N, p.B);
      checkLaunch("cgDirection");
    }

    template <class T>
    void chebStep(hipStream_t s, View<T> xp, View<const T> hx, const double *dinv,
                  View<const T> tE, u32 nC, View<const T> xk, double alpha,
                  double beta, double gamma)
    {
      const u64 n = (u64)xp.N * xp.B;
      if (!n)
        return;
      chebStepK<T><<<gridFor(n), TPB, 0, s>>>(xp.p, xp.ld, hx.p, hx.ld, dinv, tE.p,
                                              tE.ld, nC, xk.p, xk.ld, (T)alpha,
                                              (T)beta, (T)gamma, xp.N, xp.B);
      checkLaunch("chebStep");
    }

    template <class T>
    void resStep(hipStream_t s, View<T> res, View<const T> s2, View<const T> rn,
                 View<const T> r0, const double *mu, const double *coef)
    {
      const u64 n = (u64)res.N * res.B;
      if (!n)
        return;
      constexpr u32 V       = 16 / sizeof(T);
      auto          aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
      if (res.ld == res.B && s2.ld == res.B && rn.ld == res.B && r0.ld == res.B &&
          res.B % V == 0 && n / V < (1ull << 32) && aligned(res.p) && aligned(s2.p) &&
          aligned(rn.p) && aligned(r0.p))
        {
          const u64 nv = n / V;
          resStepVecK<T><<<gridFor(nv), TPB, 0, s>>>(res.p, s2.p, rn.p, r0.p, mu, coef,
                                                     (u32)nv, res.B);
          checkLaunch("resStepVec");
          return;
        }
      resStepK<T><<<gridFor(n), TPB, 0, s>>>(res.p, res.ld, s2.p, s2.ld, rn.p, rn.ld,
                                             r0.p, r0.ld, mu, coef, res.N, res.B);
      checkLaunch("resStep");
    }

    u64 hashWords(hipStream_t s, const void *x, u64 nWords, u64 *partials)
    {
      if (!nWords)
        return 0;
      u64 blocks = (nWords + 255) / 256;
      if (blocks > HASH_BLOCKS)
        blocks = HASH_BLOCKS;
      hashWordsK<<<(unsigned)blocks, 256, 0, s>>>(static_cast<const u64 *>(x), nWords,
                                                  partials);
      checkLaunch("hashWords");
      std::vector<u64> h(blocks);
      DFX_HIP(hipMemcpyAsync(h.data(), partials, blocks * sizeof(u64), hipMemcpyDeviceToHost,
                             s));
      DFX_HIP(hipStreamSynchronize(s));
      u64 sum = 0;
      for (u64 v : h)
        sum += v;
      return sum;
    }

    template <class T>
    void resFix(hipStream_t s, View<T> res, View<const T> s2, View<const T> rn,
                View<const T> r0, const u32 *rows, u32 nRows, const double *mu,
                const double *coef)
    {
      const u64 n = (u64)nRows * res.B;
      if (!n)
        return;
      DFX_REQUIRE(res.ld == rn.ld && rn.ld == r0.ld, "residual fix-up: one leading dimension");
      resFixK<T><<<gridFor(n), TPB, 0, s>>>(res.p, s2.p, s2.ld, rn.p, r0.p, res.ld, rows, nRows,
                                            mu, coef, res.B);
      checkLaunch("resFix");
    }

    template <class T>
    void resSeed(hipStream_t s, View<T> r0, View<T> rn, View<T> res, View<const double> Y,
                 double alpha1)
    {
      const u64 n = (u64)Y.N * Y.B;
      if (!n)
        return;
      DFX_REQUIRE(r0.ld == rn.ld && rn.ld == res.ld, "residual seed: one leading dimension");
      resSeedK<T><<<gridFor(n), TPB, 0, s>>>(r0.p, rn.p, res.p, r0.ld, Y.p, Y.ld, alpha1, 0.0,
                                             Y.N, Y.B);
      checkLaunch("resSeed");
    }

    template <class T>
    void resSeedFrom(hipStream_t s, View<T> r0, View<T> rn, View<T> res, View<const double> AX,
                     View<const double> BX, const double *ones, const double *lam,
                     double alpha1)
    {
      const u64 n = (u64)AX.N * AX.B;
      if (!n)
        return;
      DFX_REQUIRE(r0.ld == rn.ld && rn.ld == res.ld && AX.ld == BX.ld,
                  "residual seed: one leading dimension");
      resSeedFromK<T><<<gridFor(n), TPB, 0, s>>>(r0.p, rn.p, res.p, r0.ld, AX.p, BX.p, AX.ld,
                                                 ones, lam, alpha1, 0.0, AX.N, AX.B);
      checkLaunch("resSeedFrom");
    }

    template <class T>
    void resFinal(hipStream_t s, View<double> z, double a1, const double *a, View<const T> x,
                  double b1, const double *b, View<const double> y)
    {
      const u64 n = (u64)z.N * z.B;
      if (!n)
        return;
      resFinalK<T><<<gridFor(n), TPB, 0, s>>>(z.p, z.ld, a1, a, x.p, x.ld, b1, b, y.p, y.ld,
                                              z.N, z.B);
      checkLaunch("resFinal");
    }

    template <class T>
    void resFinalDiag(hipStream_t s, View<double> z, double a1, const double *a,
                      View<const T> r, const double *d, u32 nC, View<const T> xE, double b1,
                      const double *b, View<const double> y)
    {
      const u64 n = (u64)z.N * z.B;
      if (!n)
        return;
      DFX_REQUIRE(r.ld == xE.ld && r.N == z.N && xE.N == z.N && nC <= z.N,
                  "residual result: shapes do not match");
      resFinalDiagK<T><<<gridFor(n), TPB, 0, s>>>(z.p, z.ld, a1, a, r.p, d, nC, xE.p, r.ld, b1,
                                                  b, y.p, y.ld, z.N, z.B);
      checkLaunch("resFinalDiag");
    }

    template <class T>
    void p2c(hipStream_t s, const dfx_constraints *c, View<T> x)
    {
      if (!c || !c->nRows || !x.B)
        return;
      const u64 n = (u64)c->nRows * x.B;
      p2cK<T><<<gridFor(n), TPB, 0, s>>>(x.p, x.ld, c->rows.as<u32>(),
                                         c->rowOffsets.as<u32>(), c->cols.as<u32>(),
                                         c->weights.as<double>(), c->inhom.as<double>(),
                                         c->nRows, x.B);
      checkLaunch("p2c");
    }

    template <class T>
    void c2p(hipStream_t s, const dfx_constraints *c, View<T> y)
    {
      if (!c || !c->nRows || !y.B)
        return;
      if (c->nParents)
        {
          const u64 n = (u64)c->nParents * y.B;
          c2pParentsK<T><<<gridFor(n), TPB, 0, s>>>(
            y.p, y.ld, c->parents.as<u32>(), c->parentOffsets.as<u32>(),
            c->parentRowDofs.as<u32>(), c->parentWeights.as<double>(), c->nParents,
            y.B);
          checkLaunch("c2pParents");
        }
      const u64 n = (u64)c->nRows * y.B;
      setRowsK<T><<<gridFor(n), TPB, 0, s>>>(y.p, y.ld, c->rows.as<u32>(), c->nRows,
                                             y.B, (T)0);
      checkLaunch("c2pZero");
    }

    template <class T>
    void setRows(hipStream_t s, const dfx_constraints *c, View<T> x, double v)
    {
      if (!c || !c->nRows || !x.B)
        return;
      const u64 n = (u64)c->nRows * x.B;
      setRowsK<T><<<gridFor(n), TPB, 0, s>>>(x.p, x.ld, c->rows.as<u32>(), c->nRows,
                                             x.B, (T)v);
      checkLaunch("setRows");
    }

    template <class T>
    void gather(hipStream_t s, const u32 *map, u64 nSlots, View<const T> x, View<T> xc)
    {
      const u64 n = nSlots * x.B;
      if (!n)
        return;
      gatherK<T><<<gridFor(n), TPB, 0, s>>>(map, nSlots, x.p, x.ld, xc.p, xc.ld, x.B);
      checkLaunch("gather");
    }

    template <class T>
    void scatterAdd(hipStream_t s, const u32 *off, const u32 *slots, u32 nDofs,
                    View<const T> yc, View<T> y)
    {
      const u64 n = (u64)nDofs * y.B;
      if (!n)
        return;
      scatterAddK<T><<<gridFor(n), TPB, 0, s>>>(off, slots, nDofs, yc.p, yc.ld, y.p,
                                                y.ld, y.B);
      checkLaunch("scatterAdd");
    }

#define DFX_INST_T(T)                                                              \
  template void fill<T>(hipStream_t, View<T>, double);                             \
  template void fillRandom<T>(hipStream_t, View<T>, u64, double, double);          \
  template void axpby<T>(hipStream_t, View<T>, double, View<const T>, double,      \
                         View<const T>);                                           \
  template void axpbyBlocked<T>(hipStream_t, View<T>, double, const double *,      \
                                View<const T>, double, const double *,             \
                                View<const T>);                                    \
  template void diagScale<T>(hipStream_t, View<T>, const double *, View<const T>,  \
                             u32);                                                 \
  template void enrichedBlock<T>(hipStream_t, View<T>, const double *,             \
                                 View<const T>, u32, u32);                         \
  template void colDots<T>(hipStream_t, View<const T>, View<const T>, double *,    \
                           double *);                                              \
  template void chebStep<T>(hipStream_t, View<T>, View<const T>, const double *,   \
                            View<const T>, u32, View<const T>, double, double,     \
                            double);                                               \
  template void resStep<T>(hipStream_t, View<T>, View<const T>, View<const T>,     \
                           View<const T>, const double *, const double *);         \
  template void resFix<T>(hipStream_t, View<T>, View<const T>, View<const T>,      \
                          View<const T>, const u32 *, u32, const double *,         \
                          const double *);                                         \
  template void resSeedFrom<T>(hipStream_t, View<T>, View<T>, View<T>,             \
                               View<const double>, View<const double>,             \
                               const double *, const double *, double);            \
  template void resSeed<T>(hipStream_t, View<T>, View<T>, View<T>,                 \
                           View<const double>, double);                            \
  template void resFinal<T>(hipStream_t, View<double>, double, const double *,     \
                            View<const T>, double, const double *,                 \
                            View<const double>);                                   \
  template void resFinalDiag<T>(hipStream_t, View<double>, double, const double *, \
                                View<const T>, const double *, u32, View<const T>, \
                                double, const double *, View<const double>);       \
  template void p2c<T>(hipStream_t, const dfx_constraints *, View<T>);             \
  template void c2p<T>(hipStream_t, const dfx_constraints *, View<T>);             \
  template void setRows<T>(hipStream_t, const dfx_constraints *, View<T>, double); \
  template void gather<T>(hipStream_t, const u32 *, u64, View<const T>, View<T>);  \
  template void scatterAdd<T>(hipStream_t, const u32 *, const u32 *, u32,          \
                              View<const T>, View<T>);

    DFX_INST_T(double)
    DFX_INST_T(float)
    template void convert<double, float>(hipStream_t, View<double>, View<const float>);
    template void convert<float, double>(hipStream_t, View<float>, View<const double>);
    template void convert<double, double>(hipStream_t, View<double>, View<const double>);
    template void convert<float, float>(hipStream_t, View<float>, View<const float>);
  } // namespace k
} // namespace dfx
