// Tall-skinny dense kernels of the orthonormalization and Rayleigh-Ritz
// steps, FP64 MFMA (v_mfma_f64_16x16x4f64):
//   gram:   S = X^T Y      (computeXTransOpX, src/linearAlgebra/
//                           OrthonormalizationFunctions.t.cpp:1178-1380 and
//                           RayleighRitzEigenSolver.t.cpp:488-844; only the
//                           lower block triangle is formed, as there)
//   rotate: Xo = X Q       (the subspace rotations, OrthonormalizationFunctions.t.cpp:
//                           380-394 and RayleighRitzEigenSolver.t.cpp:272-286)
// The N-long reduction of gram is split over row chunks whose partial n x n
// blocks are summed in chunk order by a second kernel: deterministic, no
// atomics.
#include "kernels.hpp"

#include <algorithm>

namespace dfx
{
  namespace k
  {
    namespace
    {
      typedef double d4_t __attribute__((ext_vector_type(4)));
      constexpr int GB  = 64; // S block edge handled by one workgroup
      constexpr int KS  = 32; // rows per LDS slab
      constexpr int MAXCHUNKS = 192;

      __device__ inline d4_t mfma64(double a, double b, d4_t c)
      {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
      }

      // grid.x: lower-triangular block pair index, grid.y: row chunk
      template <class T>
      __global__ __launch_bounds__(256) void gramK(const T *__restrict__ X, u32 ldx,
                                                   const T *__restrict__ Y, u32 ldy,
                                                   u32 N, u32 n, u32 rowsPerChunk,
                                                   double *__restrict__ partials)
      {
        __shared__ double sX[KS][GB];
        __shared__ double sY[KS][GB];
        // decode block pair (ab >= bb)
        u32 p = blockIdx.x, ab = 0;
        while (p > ab)
          {
            p -= ab + 1;
            ++ab;
          }
        const u32 bb = p;
        const u32 a0 = ab * GB, b0 = bb * GB;
        const u32 r0 = blockIdx.y * rowsPerChunk;
        const u32 r1 = min(N, r0 + rowsPerChunk);
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;

        d4_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[t] = d4_t{0, 0, 0, 0};

        for (u32 i0 = r0; i0 < r1; i0 += KS)
          {
            for (int t = threadIdx.x; t < KS * GB; t += 256)
              {
                const int kk = t / GB, c = t % GB;
                const u32 i  = i0 + kk;
                double    xv = 0.0, yv = 0.0;
                if (i < r1)
                  {
                    if (a0 + c < n)
                      xv = (double)X[(u64)i * ldx + a0 + c];
                    if (b0 + c < n)
                      yv = (double)Y[(u64)i * ldy + b0 + c];
                  }
                sX[kk][c] = xv;
                sY[kk][c] = yv;
              }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < KS; kk += 4)
              {
                const double a = sX[kk + (l >> 4)][w * 16 + (l & 15)];
#pragma unroll
                for (int t = 0; t < 4; ++t)
                  acc[t] = mfma64(a, sY[kk + (l >> 4)][t * 16 + (l & 15)], acc[t]);
              }
            __syncthreads();
          }
        // partial block -> partials[chunk][a + n b] (column-major n x n)
        double *P = partials + (u64)blockIdx.y * n * n;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            {
              const u32 a = a0 + w * 16 + (l >> 4) + 4 * v;
              const u32 b = b0 + t * 16 + (l & 15);
              if (a < n && b < n)
                P[a + (u64)n * b] = acc[t][v];
            }
      }

      // S[a + n b] = sum over chunks (in order); lower triangle mirrored up
      __global__ void gramReduceK(const double *partials, u32 nChunks, u32 n,
                                  double *S)
      {
        const u64 tot = (u64)n * n;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < tot;
             t += (u64)gridDim.x * blockDim.x)
          {
            const u32 a = (u32)(t % n), b = (u32)(t / n);
            const u32 lo = a >= b ? a : b, hi = a >= b ? b : a; // element (lo, hi), lo >= hi
            // blocks with ab >= bb were computed; (lo, hi) lies in such a block
            double s = 0.0;
            for (u32 c = 0; c < nChunks; ++c)
              s += partials[(u64)c * n * n + lo + (u64)n * hi];
            S[t] = s;
          }
      }

      // Qr[a * nbp + b] = Q[a + n b] (or Q[b + n a] if transposed), zero padded
      __global__ void packQK(const double *Q, u32 n, u32 nbp, int transpose, double *Qr)
      {
        const u64 tot = (u64)n * nbp;
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < tot;
             t += (u64)gridDim.x * blockDim.x)
          {
            const u32 a = (u32)(t / nbp), b = (u32)(t % nbp);
            double    v = 0.0;
            if (b < n)
              v = transpose ? Q[b + (u64)n * a] : Q[a + (u64)n * b];
            Qr[t] = v;
          }
      }

      // Xo[i, b] = sum_a X[i, a] Qr[a, b]; a workgroup owns 64 rows and
      // NBT 16-column tiles of the output
      template <class T, int NBT>
      __global__ __launch_bounds__(256) void rotateK(const T *__restrict__ X, u32 ldx,
                                                     u32 N, u32 n,
                                                     const double *__restrict__ Qr,
                                                     u32 nbp, u32 bt0,
                                                     T *__restrict__ Xo, u32 ldo)
      {
        __shared__ double sX[64][KS + 1];
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        const u32 i0 = blockIdx.x * 64;
        d4_t acc[NBT];
#pragma unroll
        for (int t = 0; t < NBT; ++t)
          acc[t] = d4_t{0, 0, 0, 0};
        for (u32 k0 = 0; k0 < n; k0 += KS)
          {
            for (int t = threadIdx.x; t < 64 * KS; t += 256)
              {
                const int r = t / KS, c = t % KS;
                double    v = 0.0;
                if (i0 + r < N && k0 + c < n)
                  v = (double)X[(u64)(i0 + r) * ldx + k0 + c];
                sX[r][c] = v;
              }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < KS; kk += 4)
              {
                const u32    ka = k0 + kk + (l >> 4);
                const double a  = sX[w * 16 + (l & 15)][kk + (l >> 4)];
#pragma unroll
                for (int t = 0; t < NBT; ++t)
                  {
                    const u32 b  = (bt0 + t) * 16 + (l & 15);
                    double    qv = (ka < n) ? Qr[(u64)ka * nbp + b] : 0.0;
                    acc[t]       = mfma64(a, qv, acc[t]);
                  }
              }
            __syncthreads();
          }
#pragma unroll
        for (int t = 0; t < NBT; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            {
              const u32 i = i0 + w * 16 + (l >> 4) + 4 * v;
              const u32 b = (bt0 + t) * 16 + (l & 15);
              if (i < N && b < n)
                Xo[(u64)i * ldo + b] = (T)acc[t][v];
            }
      }

      // S = X^T Y, every lower-triangular 16 x 16 tile of S for one row chunk
      // in one workgroup (8 waves, tile p handled by wave p mod 8), so X and Y
      // are read once; rows run in slabs of GKS staged in LDS.  Per element
      // the rows of a chunk are accumulated in ascending steps of 4.
      constexpr int GKS = 16;
      template <class T, int TPW>
      __global__ __launch_bounds__(512) void gramAllK(const T *__restrict__ X, u32 ldx,
                                                      const T *__restrict__ Y, u32 ldy,
                                                      u32 N, u32 n, u32 nbp, u32 nTiles,
                                                      u32 rowsPerChunk,
                                                      double *__restrict__ partials)
      {
        extern __shared__ __align__(16) unsigned char smemG[];
        double   *sX = reinterpret_cast<double *>(smemG);   // [GKS][nbp]
        double   *sY = sX + GKS * nbp;                      // [GKS][nbp]
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
        // tile p (lower triangle, row-major over ta >= tb) -> (ta, tb)
        int ta[TPW], tb[TPW];
#pragma unroll
        for (int j = 0; j < TPW; ++j)
          {
            u32 p = (u32)(w + 8 * j), a = 0;
            while (p > a)
              {
                p -= a + 1;
                ++a;
              }
            ta[j] = (int)a;
            tb[j] = (int)p;
          }
        const u32 r0 = blockIdx.x * rowsPerChunk;
        const u32 r1 = min(N, r0 + rowsPerChunk);
        d4_t acc[TPW];
#pragma unroll
        for (int j = 0; j < TPW; ++j)
          acc[j] = d4_t{0, 0, 0, 0};
        for (u32 i0 = r0; i0 < r1; i0 += GKS)
          {
            for (u32 t = threadIdx.x; t < GKS * nbp; t += 512)
              {
                const u32 kk = t / nbp, c = t % nbp, i = i0 + kk;
                const bool ok = i < r1 && c < n;
                sX[t]         = ok ? (double)X[(u64)i * ldx + c] : 0.0;
                sY[t]         = ok ? (double)Y[(u64)i * ldy + c] : 0.0;
              }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < GKS; kk += 4)
              {
                const u32 row = (u32)(kk + (l >> 4)) * nbp + (l & 15);
#pragma unroll
                for (int j = 0; j < TPW; ++j)
                  if ((u32)(w + 8 * j) < nTiles)
                    acc[j] = mfma64(sX[row + ta[j] * 16], sY[row + tb[j] * 16], acc[j]);
              }
            __syncthreads();
          }
        double *P = partials + (u64)blockIdx.x * n * n;
#pragma unroll
        for (int j = 0; j < TPW; ++j)
          if ((u32)(w + 8 * j) < nTiles)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              {
                const u32 a = ta[j] * 16 + (l >> 4) + 4 * v;
                const u32 b = tb[j] * 16 + (l & 15);
                if (a < n && b < n)
                  P[a + (u64)n * b] = acc[j][v];
              }
      }

      // Xo[i, b] = sum_a X[i, a] Qr[a, b] for every output column in one pass
      // over X: a workgroup owns 64 rows (one 16-row M tile per wave) and all
      // NBT 16-column tiles (nbt of them live); K runs in slabs of RKS rows of
      // Q staged in LDS beside the matching 64 x RKS tile of X.  The K order
      // of the accumulation is the same as rotateK's (ascending, 4 at a time).
      constexpr int RKS = 16;
      template <class T, int NBT>
      __global__ __launch_bounds__(256, 2) void rotateAllK(const T *__restrict__ X, u32 ldx,
                                                           u32 N, u32 n,
                                                           const double *__restrict__ Qr,
                                                           u32 nbp, u32 nbt,
                                                           T *__restrict__ Xo, u32 ldo)
      {
        __shared__ double sX[64][RKS + 1];
        __shared__ double sQ[RKS][NBT * 16];
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        const u32 i0 = blockIdx.x * 64;
        d4_t acc[NBT];
#pragma unroll
        for (int t = 0; t < NBT; ++t)
          acc[t] = d4_t{0, 0, 0, 0};
        for (u32 k0 = 0; k0 < n; k0 += RKS)
          {
            for (int t = threadIdx.x; t < 64 * RKS; t += 256)
              {
                const int r = t / RKS, c = t % RKS;
                double    v = 0.0;
                if (i0 + r < N && k0 + c < n)
                  v = (double)X[(u64)(i0 + r) * ldx + k0 + c];
                sX[r][c] = v;
              }
            for (int t = threadIdx.x; t < RKS * (int)nbp; t += 256)
              {
                const int r = t / nbp, c = t % nbp;
                sQ[r][c]    = k0 + r < n ? Qr[(u64)(k0 + r) * nbp + c] : 0.0;
              }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < RKS; kk += 4)
              {
                const double a = sX[w * 16 + (l & 15)][kk + (l >> 4)];
#pragma unroll
                for (int t = 0; t < NBT; ++t)
                  if (t < (int)nbt)
                    acc[t] = mfma64(a, sQ[kk + (l >> 4)][t * 16 + (l & 15)], acc[t]);
              }
            __syncthreads();
          }
#pragma unroll
        for (int t = 0; t < NBT; ++t)
          if (t < (int)nbt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              {
                const u32 i = i0 + w * 16 + (l >> 4) + 4 * v;
                const u32 b = t * 16 + (l & 15);
                if (i < N && b < n)
                  Xo[(u64)i * ldo + b] = (T)acc[t][v];
              }
      }

      u32 chunksFor(u32 N) { return std::max(1u, std::min<u32>(MAXCHUNKS, (N + 2047) / 2048)); }
      // row chunks of the all-tile kernel: about two workgroups per CU
      constexpr u32 MAXCHUNKS_ALL = 512;
      u32 chunksForAll(u32 N) { return std::max(1u, std::min<u32>(MAXCHUNKS_ALL, (N + 1023) / 1024)); }
    } // namespace

    size_t gramPartialsNeeded(u32 N, u32 n)
    {
      return (size_t)std::max(chunksFor(N), chunksForAll(N)) * n * n;
    }

    template <class T>
    void gram(hipStream_t s, View<const T> X, View<const T> Y, double *partials,
              size_t partialsCap, double *S)
    {
      const u32 n = X.B, N = X.N;
      DFX_REQUIRE(Y.B == n && Y.N == N, "gram: shape mismatch");
      const u32 nbt = (n + 15) / 16, nbp = nbt * 16, nTiles = nbt * (nbt + 1) / 2;
      if (nTiles <= 8 * 17) // n <= 256: all tiles in one workgroup per row chunk
        {
          const u32 nChunks = chunksForAll(N);
          DFX_REQUIRE(partialsCap >= (size_t)nChunks * n * n, "gram: partials too small");
          const u32    rows = (N + nChunks - 1) / nChunks;
          const size_t lds  = (size_t)2 * GKS * nbp * sizeof(double);
          const int    tpw  = (int)((nTiles + 7) / 8);
          auto go = [&](auto kern) {
            if (lds > 64 * 1024)
              DFX_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds));
            kern<<<nChunks, 512, lds, s>>>(X.p, X.ld, Y.p, Y.ld, N, n, nbp, nTiles, rows,
                                           partials);
          };
          if (tpw <= 4)
            go(gramAllK<T, 4>);
          else if (tpw <= 8)
            go(gramAllK<T, 8>);
          else if (tpw <= 12)
            go(gramAllK<T, 12>);
          else
            go(gramAllK<T, 17>);
          checkLaunch("gramAll");
          const u64 tot = (u64)n * n;
          gramReduceK<<<(unsigned)std::min<u64>(4096, (tot + 255) / 256), 256, 0, s>>>(
            partials, nChunks, n, S);
          checkLaunch("gramReduce");
          return;
        }
      const u32 nChunks = chunksFor(N);
      DFX_REQUIRE(partialsCap >= (size_t)nChunks * n * n, "gram: partials too small");
      const u32 nb     = (n + GB - 1) / GB;
      const u32 nPairs = nb * (nb + 1) / 2;
      const u32 rows   = (N + nChunks - 1) / nChunks;
      gramK<T><<<dim3(nPairs, nChunks), 256, 0, s>>>(X.p, X.ld, Y.p, Y.ld, N, n, rows,
                                                     partials);
      checkLaunch("gram");
      const u64 tot = (u64)n * n;
      gramReduceK<<<(unsigned)std::min<u64>(4096, (tot + 255) / 256), 256, 0, s>>>(
        partials, nChunks, n, S);
      checkLaunch("gramReduce");
    }

    template <class T>
    void rotate(hipStream_t s, View<const T> X, const double *Q, View<T> Xo,
                double *qrScratch, rocblas_handle blas)
    {
      // Q: n x n column-major, device
      const u32 n = X.B, N = X.N;
      DFX_REQUIRE(Xo.B == n && Xo.N == N, "rotate: shape mismatch");
      if constexpr (sizeof(T) == 8)
        if (blas && N <= 0x7fffffffu)
          {
            // column-major view: Xo^T (n x N, ld Xo.ld) = Q^T (n x n) X^T
            // (n x N, ld X.ld).  A plain GEMM: at N = 2.1 M, n = 200 the
            // library takes about 4 ms where rotateAllK below takes 6
            const double   one = 1.0, zero = 0.0;
            CaptureLock    lk(captureMutex());
            rocblas_set_pointer_mode(blas, rocblas_pointer_mode_host);
            if (rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none,
                              (rocblas_int)n, (rocblas_int)N, (rocblas_int)n, &one, Q,
                              (rocblas_int)n, reinterpret_cast<const double *>(X.p),
                              (rocblas_int)X.ld, &zero, reinterpret_cast<double *>(Xo.p),
                              (rocblas_int)Xo.ld) != rocblas_status_success)
              throw Error(DFX_E_HIP, "rocblas_dgemm (rotation) failed");
            return;
          }
      const u32 nbt = (n + 15) / 16, nbp = nbt * 16;
      packQK<<<(unsigned)std::min<u64>(4096, ((u64)n * nbp + 255) / 256), 256, 0, s>>>(
        Q, n, nbp, 0, qrScratch);
      checkLaunch("packQ");
      const unsigned gx = (N + 63) / 64;
      if (nbt <= 16 && (RKS % 4) == 0)
        {
          // one pass over X for all columns (n <= 256)
          if (nbt <= 8)
            rotateAllK<T, 8><<<gx, 256, 0, s>>>(X.p, X.ld, N, n, qrScratch, nbp, nbt, Xo.p,
                                               Xo.ld);
          else if (nbt <= 13)
            rotateAllK<T, 13><<<gx, 256, 0, s>>>(X.p, X.ld, N, n, qrScratch, nbp, nbt, Xo.p,
                                                Xo.ld);
          else
            rotateAllK<T, 16><<<gx, 256, 0, s>>>(X.p, X.ld, N, n, qrScratch, nbp, nbt, Xo.p,
                                                Xo.ld);
          checkLaunch("rotateAll");
          return;
        }
      for (u32 bt = 0; bt < nbt; bt += 4)
        {
          const u32 left = nbt - bt;
          if (left >= 4)
            rotateK<T, 4><<<gx, 256, 0, s>>>(X.p, X.ld, N, n, qrScratch, nbp, bt,
                                            Xo.p, Xo.ld);
          else if (left == 3)
            rotateK<T, 3><<<gx, 256, 0, s>>>(X.p, X.ld, N, n, qrScratch, nbp, bt,
                                            Xo.p, Xo.ld);
          else if (left == 2)
            rotateK<T, 2><<<gx, 256, 0, s>>>(X.p, X.ld, N, n, qrScratch, nbp, bt,
                                            Xo.p, Xo.ld);
          else
            rotateK<T, 1><<<gx, 256, 0, s>>>(X.p, X.ld, N, n, qrScratch, nbp, bt,
                                            Xo.p, Xo.ld);
          checkLaunch("rotate");
        }
    }

    template void gram<double>(hipStream_t, View<const double>, View<const double>,
                               double *, size_t, double *);
    template void rotate<double>(hipStream_t, View<const double>, const double *,
                                 View<double>, double *, rocblas_handle);
  } // namespace k
} // namespace dfx
