// The fused cell apply of libdfx: for every cell e of one colour / size class
//   X_e = Q_e X            (gather through the integer cell->dof map, into LDS)
//   Y_e = A_e^T X_e        (MFMA, FP64 16x16x4 or FP32 16x16x4)
//   Y  += Q_e^T Y_e        (classical rows: read-modify-write, race-free because
//                           cells of one colour share no classical dof;
//                           enriched rows: stored per cell, reduced later in
//                           cell order)
// This is computeAxCellWiseOptimized (src/ksdft/KohnShamOperatorContextFE.t.cpp:
// 958-1199): copyFieldToCellWiseData (src/basis/FECellWiseDataOperations.t.cpp:
// 60-84), gemmStridedVarBatched with m = B, n = k = n_e, yCell = xCell * A_e
// (src/linearAlgebra/BlasLapack.t.cpp:391-436) and addCellWiseDataToFieldData
// (FECellWiseDataOperations.t.cpp:89-117), in one kernel.
//
// Cell matrices are stored on the device in "fragment order": per cell,
// per 16-row M tile, per K block of 4G (G = 16 bytes / sizeof(T)), 64 lanes x
// G values, so each wave loads its MFMA A operand as one fully coalesced
// 1 KiB read.  Entry (lane l, value s) of block (mt, kb) is
// A_e^T[mt*16 + l%16][kb*4G + G*(l/16) + s] = A_e(k, j) (column-major input).
#include "kernels.hpp"

namespace dfx
{
  namespace k
  {
    namespace
    {
      typedef double d4_t __attribute__((ext_vector_type(4)));
      typedef float  f4_t __attribute__((ext_vector_type(4)));
      typedef double d2_t __attribute__((ext_vector_type(2)));

      template <class T>
      struct Mfma;

      template <>
      struct Mfma<double>
      {
        static constexpr int G = 2;
        using acc_t            = d4_t;
        using vec_t            = d2_t;
        __device__ static inline acc_t op(double a, double b, acc_t c)
        {
          return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
        }
        // v_mfma_f64_16x16x4f64: lane l, element v -> row (l/16) + 4v, col l%16
        __device__ static inline int row(int l, int v) { return (l >> 4) + 4 * v; }
      };

      template <>
      struct Mfma<float>
      {
        static constexpr int G = 4;
        using acc_t            = f4_t;
        using vec_t            = f4_t;
        __device__ static inline acc_t op(float a, float b, acc_t c)
        {
          return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
        }
        // v_mfma_f32_16x16x4f32: lane l, element v -> row 4(l/16) + v, col l%16
        __device__ static inline int row(int l, int v) { return 4 * (l >> 4) + v; }
      };

      // first[i / 32] bit i % 32: this cell is the first (in launch order) to
      // touch its classical row i, which is then stored instead of added, so Y
      // needs no zero fill and the first touch of every dof skips the read.
      template <class T, int NT, int MTW>
      __global__ __launch_bounds__(256) void cellApplyK(
        const CellMeta *__restrict__ meta, const u32 *__restrict__ first, u32 begin,
        const u32 *__restrict__ map, const T *__restrict__ frag,
        const T *__restrict__ x, u32 ldx, T *__restrict__ y, u32 ldy,
        T *__restrict__ yE, u32 B)
      {
        using M                = Mfma<T>;
        using acc_t            = typename M::acc_t;
        using vec_t            = typename M::vec_t;
        constexpr int G        = M::G;
        constexpr int KB       = 4 * G;
        constexpr int BC       = 16 * NT;
        constexpr int VC       = BC / G;              // 16-byte vectors per chunk row
        constexpr int KMAX     = 64 * MTW;            // largest padded K of this class
        constexpr int GATHER_IT = (KMAX * VC + 255) / 256;
        extern __shared__ __align__(16) unsigned char smemRaw[];
        T   *sX    = reinterpret_cast<T *>(smemRaw);
        __shared__ u32 sMap[256];
        __shared__ u32 sFirst[8];

        const u32      pos  = begin + blockIdx.x;
        const CellMeta cm   = meta[pos];
        const int      n    = (int)cm.n;
        const int      nc   = (int)cm.nc;
        const int      mt   = (n + 15) >> 4;
        const int      kpad = ((n + KB - 1) / KB) * KB;
        const int      nkb  = kpad / KB;
        const int      w    = threadIdx.x >> 6;
        const int      l    = threadIdx.x & 63;
        const T   *cfrag    = frag + (sizeof(T) == 8 ? cm.fragOff64 : cm.fragOff32);
        if (threadIdx.x < n)
          sMap[threadIdx.x] = map[cm.mapOff + threadIdx.x];
        if (threadIdx.x < 8)
          sFirst[threadIdx.x] = first[(u64)pos * 8 + threadIdx.x];
        __syncthreads();
        const bool vecOk = (ldx % G == 0) && (B % G == 0) &&
                           ((reinterpret_cast<uintptr_t>(x) & 15) == 0);

        for (u32 c0 = 0; c0 < B; c0 += BC)
          {
            const int bw = min(BC, (int)(B - c0));
            // ---- gather X_e (this column chunk) into LDS, K-grouped by G:
            // independent 16-byte loads, all issued before any is consumed
            if (vecOk)
              {
                vec_t v[GATHER_IT];
#pragma unroll
                for (int it = 0; it < GATHER_IT; ++it)
                  {
                    const int t  = threadIdx.x + it * 256;
                    const int kk = t / VC, c = (t - kk * VC) * G;
                    v[it]        = vec_t{};
                    if (kk < n && c < bw)
                      v[it] = *reinterpret_cast<const vec_t *>(
                        x + (u64)sMap[kk] * ldx + c0 + c);
                  }
#pragma unroll
                for (int it = 0; it < GATHER_IT; ++it)
                  {
                    const int t  = threadIdx.x + it * 256;
                    const int kk = t / VC, c = (t - kk * VC) * G;
                    if (kk < kpad)
#pragma unroll
                      for (int g = 0; g < G; ++g)
                        sX[((kk / G) * BC + c + g) * G + (kk % G)] = v[it][g];
                  }
              }
            else
              for (int t = threadIdx.x; t < kpad * BC; t += 256)
                {
                  const int kk = t / BC, c = t - kk * BC;
                  T         v  = (T)0;
                  if (kk < n && c < bw)
                    v = x[(u64)sMap[kk] * ldx + c0 + c];
                  sX[((kk / G) * BC + c) * G + (kk % G)] = v;
                }
            __syncthreads();

            acc_t acc[MTW][NT];
#pragma unroll
            for (int t = 0; t < MTW; ++t)
#pragma unroll
              for (int nt = 0; nt < NT; ++nt)
                acc[t][nt] = acc_t{0, 0, 0, 0};

            vec_t a[MTW], an[MTW];
#pragma unroll
            for (int t = 0; t < MTW; ++t)
              {
                const int mtile = w + 4 * t;
                a[t]            = vec_t{};
                if (mtile < mt)
                  a[t] = *reinterpret_cast<const vec_t *>(
                    cfrag + ((u64)(mtile * nkb) * 64 + l) * G);
              }
            for (int kb = 0; kb < nkb; ++kb)
              {
                // prefetch the next K block of A
#pragma unroll
                for (int t = 0; t < MTW; ++t)
                  {
                    const int mtile = w + 4 * t;
                    an[t]           = vec_t{};
                    if (mtile < mt && kb + 1 < nkb)
                      an[t] = *reinterpret_cast<const vec_t *>(
                        cfrag + ((u64)(mtile * nkb + kb + 1) * 64 + l) * G);
                  }
                vec_t b[NT];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                  b[nt] = *reinterpret_cast<const vec_t *>(
                    sX + (((kb * 4 + (l >> 4)) * BC) + nt * 16 + (l & 15)) * G);
#pragma unroll
                for (int s = 0; s < G; ++s)
#pragma unroll
                  for (int t = 0; t < MTW; ++t)
                    if (w + 4 * t < mt)
#pragma unroll
                      for (int nt = 0; nt < NT; ++nt)
                        acc[t][nt] = M::op(a[t][s], b[nt][s], acc[t][nt]);
#pragma unroll
                for (int t = 0; t < MTW; ++t)
                  a[t] = an[t];
              }

            // ---- scatter: classical rows into y (store on first touch, else
            // read-modify-write; a tile's reads are all issued before its
            // stores), enriched rows into yE
#pragma unroll
            for (int t = 0; t < MTW; ++t)
              {
                if (w + 4 * t >= mt)
                  continue;
                T old[NT][4];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                  for (int v = 0; v < 4; ++v)
                    {
                      const int row = (w + 4 * t) * 16 + M::row(l, v);
                      const int c   = nt * 16 + (l & 15);
                      old[nt][v]    = (T)0;
                      if (c < bw && row < nc && !((sFirst[row >> 5] >> (row & 31)) & 1u))
                        old[nt][v] = y[(u64)sMap[row] * ldy + c0 + c];
                    }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                  for (int v = 0; v < 4; ++v)
                    {
                      const int row = (w + 4 * t) * 16 + M::row(l, v);
                      const int c   = nt * 16 + (l & 15);
                      if (c >= bw || row >= n)
                        continue;
                      if (row < nc)
                        y[(u64)sMap[row] * ldy + c0 + c] = old[nt][v] + acc[t][nt][v];
                      else
                        yE[(u64)(cm.enrOff + (u32)(row - nc)) * B + c0 + c] = acc[t][nt][v];
                    }
              }
            __syncthreads();
          }
      }

      // y[nC + g, j] = sum over the cells of g of yE[slot, j].  One block per
      // (enriched dof g, 16 columns); 16 slot lanes take every 16th slot and
      // keep 8 loads in flight; the 16 lane sums are combined in a fixed
      // order, so the result does not depend on timing.
      template <class T>
      __global__ __launch_bounds__(256) void enrReduceK(const u32 *__restrict__ off,
                                                        const u32 *__restrict__ slots,
                                                        u32 nC, const T *__restrict__ yE,
                                                        u32 B, T *__restrict__ y, u32 ldy)
      {
        constexpr int CB = 16, SL = 16, U = 8;
        __shared__ T part[SL][CB];
        const u32 g  = blockIdx.x;
        const int cj = threadIdx.x % CB, sl = threadIdx.x / CB;
        const u32 j  = blockIdx.y * CB + cj;
        const u32 e0 = off[g], e1 = off[g + 1];
        T         acc = 0;
        if (j < B)
          for (u32 e = e0 + sl; e < e1; e += SL * U)
            {
              T v[U];
#pragma unroll
              for (int u = 0; u < U; ++u)
                {
                  const u32 ee = e + u * SL;
                  v[u]         = ee < e1 ? yE[(u64)slots[ee] * B + j] : (T)0;
                }
#pragma unroll
              for (int u = 0; u < U; ++u)
                acc += v[u];
            }
        part[sl][cj] = acc;
        __syncthreads();
        if (sl == 0 && j < B)
          {
            T s = 0;
#pragma unroll
            for (int q = 0; q < SL; ++q)
              s += part[q][cj];
            y[(u64)(nC + g) * ldy + j] = s;
          }
      }

      template <class T>
      __global__ void repackK(const CellMeta *meta, const u32 *order,
                              const double *raw, const u64 *rawOff, T *frag)
      {
        constexpr int G  = Mfma<T>::G;
        constexpr int KB = 4 * G;
        const CellMeta cm = meta[blockIdx.x];
        const u32      e  = order[blockIdx.x];
        const int      n  = (int)cm.n;
        const int      mt = (n + 15) >> 4;
        const int      nkb  = (n + KB - 1) / KB;
        const u64      tot  = (u64)mt * nkb * 64 * G;
        const double  *A    = raw + rawOff[e];
        T             *F    = frag + (sizeof(T) == 8 ? cm.fragOff64 : cm.fragOff32);
        for (u64 idx = threadIdx.x; idx < tot; idx += blockDim.x)
          {
            const int s    = (int)(idx % G);
            const u64 q    = idx / G;
            const int lane = (int)(q % 64);
            const u64 blk  = q / 64;
            const int kb   = (int)(blk % nkb);
            const int mtile = (int)(blk / nkb);
            const int j    = mtile * 16 + (lane & 15);
            const int kk   = kb * KB + G * (lane >> 4) + s;
            F[idx] = (j < n && kk < n) ? (T)A[(u64)kk + (u64)n * j] : (T)0;
          }
      }

      template <class T, int NT, int MTW>
      void launchCell(hipStream_t s, const dfx_op *op, const Launch &L, const T *frag,
                      View<const T> x, View<T> y, T *yE)
      {
        constexpr int G  = Mfma<T>::G;
        constexpr int BC = 16 * NT;
        (void)G;
        // dynamic tile plus the static map / first-touch words (about 1 KiB)
        const size_t lds = (size_t)L.kpadMax * BC * sizeof(T);
        DFX_REQUIRE(lds + 2048 <= 160 * 1024, "cell tile exceeds the 160 KiB LDS of a CU");
        static size_t attrSet = 60 * 1024;
        if (lds > attrSet)
          {
            DFX_HIP(hipFuncSetAttribute(
              reinterpret_cast<const void *>(&cellApplyK<T, NT, MTW>),
              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            attrSet = lds;
          }
        cellApplyK<T, NT, MTW><<<L.end - L.begin, 256, lds, s>>>(
          op->meta.as<CellMeta>(), op->first.as<u32>(), L.begin, op->map.as<u32>(), frag,
          x.p, x.ld, y.p, y.ld, yE, x.B);
        checkLaunch("cellApply");
      }

      template <class T, int NT>
      void launchCellMtw(hipStream_t s, const dfx_op *op, const Launch &L,
                         const T *frag, View<const T> x, View<T> y, T *yE)
      {
        switch (L.mtw)
          {
            case 1: return launchCell<T, NT, 1>(s, op, L, frag, x, y, yE);
            case 2: return launchCell<T, NT, 2>(s, op, L, frag, x, y, yE);
            case 3: return launchCell<T, NT, 3>(s, op, L, frag, x, y, yE);
            case 4: return launchCell<T, NT, 4>(s, op, L, frag, x, y, yE);
          }
        throw Error(DFX_E_UNSUPPORTED, "cell size above 256 dofs");
      }
    } // namespace

    // cap on the column-chunk width from DFX_CELL_NT_F64 / DFX_CELL_NT_F32
    // (1, 2, 4 or 8; 0 or unset = built-in choice), read once
    static int ntOverride(bool f64)
    {
      static const int v[2] = {
        [] {
          const char *e = getenv("DFX_CELL_NT_F32");
          return e ? atoi(e) : 0;
        }(),
        [] {
          const char *e = getenv("DFX_CELL_NT_F64");
          return e ? atoi(e) : 0;
        }()};
      const int x = v[f64 ? 1 : 0];
      return (x == 1 || x == 2 || x == 4 || x == 8) ? x : 0;
    }

    // column-chunk width (in 16-column tiles) for a launch
    template <class T>
    static int chooseNT(u32 B, u32 kpadMax)
    {
      int nt = (int)((B + 15) / 16);
      int cap;
      if (sizeof(T) == 8)
        cap = kpadMax <= 128 ? 4 : 2;
      else
        cap = kpadMax <= 128 ? 8 : 4;
      if (const int o = ntOverride(sizeof(T) == 8))
        cap = o;
      if (nt > cap)
        nt = cap;
      if (nt == 3)
        nt = 4;
      if (nt > 4 && nt < 8)
        nt = 8;
      return nt;
    }

    template <class T>
    void cellApply(hipStream_t s, const dfx_op *op, const Launch &L, const T *frag,
                   View<const T> x, View<T> y, T *yE)
    {
      if (L.end <= L.begin || !x.B)
        return;
      switch (chooseNT<T>(x.B, L.kpadMax))
        {
          case 1: return launchCellMtw<T, 1>(s, op, L, frag, x, y, yE);
          case 2: return launchCellMtw<T, 2>(s, op, L, frag, x, y, yE);
          case 4: return launchCellMtw<T, 4>(s, op, L, frag, x, y, yE);
          case 8: return launchCellMtw<T, 8>(s, op, L, frag, x, y, yE);
        }
    }

    template <class T>
    void enrichedReduce(hipStream_t s, const dfx_op *op, const T *yE, View<T> y)
    {
      const u32 nE = op->nDofs - op->nClassical;
      if (!nE || !y.B)
        return;
      DFX_REQUIRE((y.B + 15) / 16 <= 65535, "too many columns for the enriched reduction");
      enrReduceK<T><<<dim3(nE, (y.B + 15) / 16), 256, 0, s>>>(
        op->enrSlotOffsets.as<u32>(), op->enrSlots.as<u32>(), op->nClassical, yE, y.B,
        y.p, y.ld);
      checkLaunch("enrichedReduce");
    }

    template <class T>
    void repackCells(hipStream_t s, const dfx_op *op, const double *raw,
                     const u64 *rawOffsets, T *frag)
    {
      if (!op->nCells)
        return;
      // the meta array is in launch order; `order` maps back to caller cells
      DevBuf order;
      upload(order, op->hOrder, s);
      repackK<T><<<op->nCells, 256, 0, s>>>(op->meta.as<CellMeta>(), order.as<u32>(),
                                            raw, rawOffsets, frag);
      checkLaunch("repackCells");
      DFX_HIP(hipStreamSynchronize(s));
    }

    template void cellApply<double>(hipStream_t, const dfx_op *, const Launch &,
                                    const double *, View<const double>, View<double>,
                                    double *);
    template void cellApply<float>(hipStream_t, const dfx_op *, const Launch &,
                                   const float *, View<const float>, View<float>,
                                   float *);
    template void enrichedReduce<double>(hipStream_t, const dfx_op *, const double *,
                                         View<double>);
    template void enrichedReduce<float>(hipStream_t, const dfx_op *, const float *,
                                        View<float>);
    template void repackCells<double>(hipStream_t, const dfx_op *, const double *,
                                      const u64 *, double *);
    template void repackCells<float>(hipStream_t, const dfx_op *, const double *,
                                     const u64 *, float *);
  } // namespace k
} // namespace dfx
