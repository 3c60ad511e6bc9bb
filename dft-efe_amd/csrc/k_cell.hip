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
#include "res_update.hpp"

// Build-time tuning of the cell kernel (defaults measured on MI355X):
// A-operand prefetch depth in K blocks, whether the read-modify-write loads of
// Y are issued before the MFMA loop, and the minimum waves per SIMD the
// register allocation must allow.
#ifndef DFX_CELL_DEPTH
#define DFX_CELL_DEPTH 1
#endif
#ifndef DFX_CELL_EARLY
#define DFX_CELL_EARLY 1
#endif
#ifndef DFX_CELL_MINW
#define DFX_CELL_MINW 2
#endif
#ifndef DFX_CELL_MINW_SMALL
#define DFX_CELL_MINW_SMALL 4
#endif
#ifndef DFX_CELL_EARLY_SMALL
#define DFX_CELL_EARLY_SMALL 0
#endif
#ifndef DFX_CELL_AREG_WIDE
#define DFX_CELL_AREG_WIDE 0
#endif
#ifndef DFX_CELL_MINW_MID
#define DFX_CELL_MINW_MID 3
#endif
#ifndef DFX_CELL_NT5
#define DFX_CELL_NT5 1
#endif
#ifndef DFX_CELL_ORDER
#define DFX_CELL_ORDER 0
#endif
#ifndef DFX_CELL_AREG_MAX
#define DFX_CELL_AREG_MAX 128
#endif
#ifndef DFX_CELL_DEPTH1
#define DFX_CELL_DEPTH1 1
#endif
// whether the thread index is made opaque per column chunk in the kernels
// without the fused residual step too (always in those with it)
#ifndef DFX_CELL_OPAQUE
#define DFX_CELL_OPAQUE 1
#endif
// whether the fused kernels keep the gathered rows of X unscaled in LDS
// (M^-1 applied to the MFMA operand instead) and read x for the residual
// step from there instead of from global memory
#ifndef DFX_RES_LDSX
#define DFX_RES_LDSX 1
#endif
// whether the fused kernels request Y's old values per row group too
#ifndef DFX_RES_OLDG
#define DFX_RES_OLDG 1
#endif
// rows per request group of the fused residual step's operands (1, 2 or 4),
// for cells of up to 128 dofs and for larger ones
#ifndef DFX_RES_ROWS_SMALL
#define DFX_RES_ROWS_SMALL 2
#endif
#ifndef DFX_RES_ROWS
#define DFX_RES_ROWS 2
#endif


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

      // ---- raw buffer access (32-bit byte offsets, hardware range check: a
      // load at an offset past the buffer returns 0 and such a store is
      // dropped, so masked lanes need no branch)
      typedef __amdgpu_buffer_rsrc_t rsrc_t;
      typedef unsigned u2_t __attribute__((ext_vector_type(2)));
      typedef unsigned u4_t __attribute__((ext_vector_type(4)));
      constexpr u32 OFF_NONE = 0x80000000u; // past every buffer (all < 2^31 bytes)

      __device__ inline rsrc_t rsrc(const void *p, u32 bytes)
      {
        // third dword 0x00020000: raw (unswizzled) 32-bit data, gfx9 family
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                                 (int)bytes, 0x00020000);
      }

      template <class T>
      struct Buf;
      template <>
      struct Buf<float>
      {
        __device__ static f4_t load4(rsrc_t r, u32 off)
        {
          return __builtin_bit_cast(f4_t, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        }
        __device__ static float load(rsrc_t r, u32 off)
        {
          return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
        }
        __device__ static void store(float v, rsrc_t r, u32 off)
        {
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
        }
      };
      template <>
      struct Buf<double>
      {
        __device__ static d2_t load4(rsrc_t r, u32 off)
        {
          return __builtin_bit_cast(d2_t, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        }
        __device__ static double load(rsrc_t r, u32 off)
        {
          return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
        }
        __device__ static void store(double v, rsrc_t r, u32 off)
        {
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2_t, v), r, off, 0, 0);
        }
      };

      // The cell kernel on raw buffers: the same data flow as cellApplyK below
      // with no per-lane branches, since every masked load or store goes to
      // OFF_NONE.  Buffers are based at the cell's smallest classical row
      // (rowLo) for x and y and at the first enriched row (nC) for enriched x
      // rows, so offsets stay below 2 GiB when the cells' classical row spans
      // and the enriched block do (checked by the launcher).  The LDS tile is
      // KMAX rows (the class maximum) so the gather writes every row it loads.
      // FP32 launches of at most two M tiles per wave (cells of <= 128 dofs)
      // fit in 128 VGPRs when the read-modify-write loads are issued per tile
      // in the epilogue: four workgroups per CU instead of three
      template <class T, int NT, int MTW>
      struct CellPolicy
      {
        static constexpr bool small = sizeof(T) == 4 && MTW <= 2 && NT > 1 && NT <= 4;
        // FP32 cells of 129..192 dofs: three workgroups fit the LDS of a CU
        static constexpr bool mid   = sizeof(T) == 4 && MTW == 3 && NT > 1 && NT <= 4;
        static constexpr int  minw  = small ? DFX_CELL_MINW_SMALL
                                    : mid   ? DFX_CELL_MINW_MID
                                            : DFX_CELL_MINW;
        static constexpr bool early = small ? DFX_CELL_EARLY_SMALL : DFX_CELL_EARLY;
        // rows per request group of the fused residual step's operands
        static constexpr int resRows = MTW <= 2 ? DFX_RES_ROWS_SMALL : DFX_RES_ROWS;
      };

      // RES (with DSC): the residual Chebyshev step fused into the scatter.
      // last[pos * 8 + i / 32] bit i % 32: this cell is the last (in launch
      // order) to touch its classical row i and the row is unconstrained, so
      // y's row is complete here; res = resUpdate(y, x, res, r0) is stored for
      // it instead of y (x read back raw, res and r0 at x's leading dimension)
      template <class T, int NT, int MTW, bool DSC, bool TAIL, bool RES>
      __global__ __launch_bounds__(256, (CellPolicy<T, NT, MTW>::minw)) void cellApplyBufK(
        const CellMeta *__restrict__ meta, const u32 *__restrict__ first, u32 begin,
        const u32 *__restrict__ map, const T *__restrict__ frag, const T *x, u32 nRowsX,
        u32 ldx, T *y, u32 nRowsY, u32 ldy, T *yE, u32 yEBytes, u32 B, u32 nC,
        const double *__restrict__ dsc, const T *xe, ResFuse<T> rf,
        const u32 *__restrict__ last, u32 nCh, u32 nCellsL)
      {
        using M                 = Mfma<T>;
        using BF                = Buf<T>;
        using acc_t             = typename M::acc_t;
        using vec_t             = typename M::vec_t;
        constexpr int G         = M::G;
        constexpr int KB        = 4 * G;
        constexpr int BC        = 16 * NT;
        constexpr int KMAX      = 64 * MTW;
        constexpr int DEPTH     = NT == 1 ? DFX_CELL_DEPTH1 : DFX_CELL_DEPTH;
        constexpr int D         = (KMAX / KB) < DEPTH ? (KMAX / KB) : DEPTH;
        // LDS column position: within each 16-column segment the column is
        // rotated by the segment index, so that both the gather's stores
        // (columns G apart) and the MFMA's B reads (16 consecutive columns)
        // fall in 16 distinct 16-byte bank groups
        auto swz = [](int col) { return (col & ~15) | ((col + (col >> 4)) & 15); };
        constexpr bool EARLY    = CellPolicy<T, NT, MTW>::early && MTW * NT * 4 * (int)sizeof(T) <= 128;
        constexpr u32 S         = sizeof(T);
        constexpr int  NKBMAX   = KMAX / KB;
        // narrow chunks (NT == 1: few columns, little work to hide latency
        // behind) hold the wave's whole A operand in registers, requested
        // at the start; wider chunks use a D-deep register ring
        constexpr bool AREG     = (NT == 1 && MTW * NKBMAX * 4 <= // 4 VGPRs per vector
                                                  (S == 4 ? DFX_CELL_AREG_MAX : DFX_CELL_AREG_MAX / 2)) ||
                              (NT > 1 && MTW * NKBMAX * 4 <= DFX_CELL_AREG_WIDE);
        constexpr int  NRING    = AREG ? NKBMAX : D;
        // RES with the gathered rows kept unscaled in LDS
        constexpr bool LDSX     = DSC && RES && DFX_RES_LDSX;
        extern __shared__ __align__(16) unsigned char smemRaw[];
        T *sX = reinterpret_cast<T *>(smemRaw);
        __shared__ u32 sMap[256];
        __shared__ u32 sFirst[8];
        __shared__ u32 sLast[RES ? 8 : 1];
        __shared__ __align__(16) T sDsc[DSC ? 256 : 4];

        // nCh > 1: one column chunk per workgroup.  Workgroups are handed to
        // the 8 XCDs in turn (workgroup b to XCD b mod 8), so the nCh chunks
        // of a cell get workgroups b, b + 8, ... of one XCD, dispatched one
        // after the other there: the later chunks find the cell's matrix in
        // that XCD's L2
        u32 cellIdx = blockIdx.x, kOnly = 0;
        if (nCh > 1)
          {
            const u32 grp = blockIdx.x / (8 * nCh), r = blockIdx.x % (8 * nCh);
            cellIdx       = grp * 8 + r % 8;
            kOnly         = r / 8;
            if (cellIdx >= nCellsL) // the last group's missing cells (whole workgroup)
              return;
          }
        const u32      pos  = begin + cellIdx;
        const CellMeta cm   = meta[pos];
        const int      n    = (int)cm.n;
        const int      nc   = (int)cm.nc;
        const int      mt   = (n + 15) >> 4;
        const int      nkb  = (n + KB - 1) / KB;
        const int      w    = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
        const int      l    = threadIdx.x & 63;
        const rsrc_t   rA   = rsrc(frag + (S == 8 ? cm.fragOff64 : cm.fragOff32),
                                   (u32)(mt * nkb * 64 * G) * S);
        const u32      lo   = cm.rowLo;
        auto cap = [](u64 b) { return (u32)(b < 0x7fffffffull ? b : 0x7fffffffull); };
        const rsrc_t rX  = rsrc(x + (u64)lo * ldx, cap((u64)(nRowsX - lo) * ldx * S));
        // enriched rows (>= nC) reachable through rX as well: one load per row
        const bool   oneX = nc == n || (xe == x && (u64)(nRowsX - lo) * ldx * S < 0x7fffffffull);
        const rsrc_t rXE = rsrc(xe + (u64)nC * ldx, cap((u64)(nRowsX - nC) * ldx * S));
        const rsrc_t rY  = rsrc(y + (u64)lo * ldy, cap((u64)(nRowsY - lo) * ldy * S));
        const rsrc_t rE  = rsrc(yE, yEBytes);
        // fused residual step: res and r0 addressed like x
        const rsrc_t rRes = rsrc(RES ? rf.res + (u64)lo * ldx : nullptr,
                                 RES ? cap((u64)(nRowsX - lo) * ldx * S) : 0u);
        const rsrc_t rR0  = rsrc(RES ? rf.r0 + (u64)lo * ldx : nullptr,
                                 RES ? cap((u64)(nRowsX - lo) * ldx * S) : 0u);

        auto loadA = [&](int kb, vec_t *dst) {
#pragma unroll
          for (int t = 0; t < MTW; ++t)
            {
              const int mtile = w + 4 * t;
              const u32 off   = (mtile < mt && kb < nkb)
                                  ? ((u32)(mtile * nkb + kb) * 64 + l) * (G * S)
                                  : OFF_NONE;
              dst[t] = BF::load4(rA, off);
            }
        };
        vec_t ring[NRING][MTW];
        if constexpr (!DFX_CELL_ORDER)
#pragma unroll
          for (int d = 0; d < NRING; ++d)
            loadA(d, ring[d]);

        if (threadIdx.x < n)
          sMap[threadIdx.x] = map[cm.mapOff + threadIdx.x];
        // M^-1 folded in: the gathered rows of X_e are scaled row by row,
        // diag on the classical rows and 1 elsewhere (exact)
        if constexpr (DSC) // dsc holds d[map[slot]] per map slot
          sDsc[threadIdx.x] = threadIdx.x < (unsigned)nc ? (T)dsc[cm.mapOff + threadIdx.x]
                                                         : (T)1;
        if (threadIdx.x < 8)
          {
            sFirst[threadIdx.x] = first[(u64)pos * 8 + threadIdx.x];
            if constexpr (RES)
              sLast[threadIdx.x] = last[(u64)pos * 8 + threadIdx.x];
          }
        __syncthreads();

        const int cl = l & 15;
        auto rowOf = [&](int t, int v) { return (w + 4 * t) * 16 + M::row(l, v); };

        // one column chunk [c0, c0 + bw) with NTL live 16-column tiles (the
        // LDS row stride stays BC).  TAIL kernels (B % BC in [1, 16]) run the
        // last chunk with NTL = 1, so it issues no MFMA, LDS read or store
        // for columns past B; a separate instantiation because the second
        // path costs registers
        auto chunk = [&](auto ntl, const u32 c0, const int bw) {
            // RES (and DFX_CELL_OPAQUE): the thread index is made opaque per chunk, so the row
            // offsets and masks of the gather and the scatter are formed here
            // rather than hoisted out of the chunk loop and held (spilled)
            // throughout
            int lc = l, tx = threadIdx.x;
            if constexpr (RES || DFX_CELL_OPAQUE)
              asm volatile("" : "+v"(lc), "+v"(tx));
            auto rowOf = [&](int t, int v) { return (w + 4 * t) * 16 + M::row(lc, v); };
            constexpr int NTL   = decltype(ntl)::value;
            constexpr int VCL   = 16 * NTL / G;
            constexpr int UNITS = (KMAX / G) * VCL; // G x G gather units of the tile
            constexpr int UPT   = (UNITS + 255) / 256;
            static_assert(UNITS % 64 == 0, "gather units must fill whole waves");
            // ---- gather (columns are whole 16-byte vectors: B % G == 0).  A
            // unit is G rows x G columns: G 16-byte loads, transposed in
            // registers into G 16-byte LDS stores (one per column, holding G
            // consecutive K values), at the swizzled column position swz().
            vec_t v[UPT][G];
            // the rows a wave gathers for unit u form one contiguous block
            // [kLo, kHi] (wave-uniform): wholly classical blocks read rX,
            // wholly enriched blocks rXE, and the one block that straddles
            // nc reads its classical rows here and its enriched rows in a
            // second round of loads, so that every row's load is in flight
            // at once instead of a load pair and a wait per row
            bool strU[UPT];
#pragma unroll
            for (int u = 0; u < UPT; ++u)
              {
                const int  t    = tx + u * 256;
                const int  q    = t / VCL, c = (t - q * VCL) * G; // K group, first column
                const bool live = t < UNITS;                     // wave-uniform
                const int  t0   = u * 256 + 64 * w;
                const int  kLo  = (t0 / VCL) * G, kHi = ((t0 + 63) / VCL) * G + G - 1;
                const bool allE = !oneX && kLo >= nc;
                strU[u]         = !oneX && kLo < nc && kHi >= nc;
                const rsrc_t r1 = allE ? rXE : rX;
                const u32    b1 = allE ? nC : lo;
#pragma unroll
                for (int r = 0; r < G; ++r)
                  {
                    const int kk = q * G + r;
                    const u32 m  = sMap[live ? kk : 0]; // kk < KMAX <= 256
                    v[u][r] = BF::load4(r1, (live & (kk < n) & (c < bw) & !(strU[u] & (kk >= nc)))
                                              ? ((m - b1) * ldx + c0 + c) * S
                                              : OFF_NONE);
                  }
              }
            if (!oneX)
              {
                // at most one unit of this wave straddles nc (blocks are
                // disjoint): its enriched rows, in one round of G loads
                int uS = -1;
#pragma unroll
                for (int u = 0; u < UPT; ++u)
                  uS = strU[u] ? u : uS;
                const int  tS    = tx + (uS < 0 ? 0 : uS) * 256;
                const int  qS    = tS / VCL, cS = (tS - qS * VCL) * G;
                const bool liveS = uS >= 0 && tS < UNITS;
                vec_t      e[G];
#pragma unroll
                for (int r = 0; r < G; ++r)
                  {
                    const int kk = qS * G + r;
                    const u32 m  = sMap[liveS ? kk : 0];
                    e[r] = BF::load4(rXE, (liveS & (kk >= nc) & (kk < n) & (cS < bw))
                                            ? ((m - nC) * ldx + c0 + cS) * S
                                            : OFF_NONE);
                  }
#pragma unroll
                for (int u = 0; u < UPT; ++u)
                  {
                    const int q = (tx + u * 256) / VCL;
#pragma unroll
                    for (int r = 0; r < G; ++r)
                      v[u][r] = (u == uS & q * G + r >= nc) ? e[r] : v[u][r];
                  }
              }
#pragma unroll
            for (int u = 0; u < UPT; ++u)
              {
                const int  t    = tx + u * 256;
                const bool live = t < UNITS;
                const int  q    = t / VCL;
                (void)live;
                if constexpr (DSC && !LDSX)
                  {
                    // M^-1 folded in: rows q G .. q G + G - 1 of the unit
                    // scaled by their diag (1 past nc, exact) before the MFMA
                    // reads them; kk < KMAX <= 256 stays in sDsc
                    const vec_t sc = *reinterpret_cast<const vec_t *>(sDsc + (live ? q * G : 0));
#pragma unroll
                    for (int r = 0; r < G; ++r)
                      v[u][r] = sc[r] * v[u][r];
                  }
              }
            if constexpr (DFX_CELL_ORDER)
              {
                // the A operand is requested after the gather: loads complete
                // in order, so the gather's LDS stores wait for the gather only
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int d = 0; d < NRING; ++d)
                  loadA(d, ring[d]);
              }

#pragma unroll
            for (int u = 0; u < UPT; ++u)
              {
                const int t = tx + u * 256;
                const int q = t / VCL, c = (t - q * VCL) * G;
                if (t < UNITS)
#pragma unroll
                  for (int g = 0; g < G; ++g)
                    {
                      vec_t col;
#pragma unroll
                      for (int r = 0; r < G; ++r)
                        col[r] = v[u][r][g];
                      *reinterpret_cast<vec_t *>(sX + (q * BC + swz(c + g)) * G) = col;
                    }
              }

            // ---- old values of the classical rows touched before
            T    old[EARLY ? MTW : 1][EARLY ? NTL : 1][4];
            auto loadOld = [&](int t, T (*o)[4], int v0 = 0, int nv = 4) {
#pragma unroll
              for (int vv = v0; vv < v0 + nv; ++vv)
                {
                  // row < 64 MTW <= 256: sMap / sFirst reads stay in the arrays
                  const int  row  = rowOf(t, vv);
                  const u32  fb   = sFirst[row >> 5], m = sMap[row];
                  const bool rmw  = (row < nc) & !((fb >> (row & 31)) & 1u);
                  const u32  base = ((m - lo) * ldy + c0 + cl) * S;
#pragma unroll
                  for (int nt = 0; nt < NTL; ++nt)
                    o[nt][vv] = BF::load(rY, (rmw & (nt * 16 + cl < bw)) ? base + nt * 16 * S
                                                                        : OFF_NONE);
                }
            };
            if constexpr (EARLY)
              {
#pragma unroll
                for (int t = 0; t < MTW; ++t)
                  if (w + 4 * t < mt)
                    loadOld(t, old[t]);
              }
            __syncthreads();

            acc_t acc[MTW][NTL];
#pragma unroll
            for (int t = 0; t < MTW; ++t)
#pragma unroll
              for (int nt = 0; nt < NTL; ++nt)
                acc[t][nt] = acc_t{0, 0, 0, 0};

            if constexpr (AREG)
              {
#pragma unroll
                for (int kb = 0; kb < NKBMAX; ++kb)
                  if (kb < nkb) // block-uniform
                    {
                      vec_t b[NTL];
#pragma unroll
                      for (int nt = 0; nt < NTL; ++nt)
                        b[nt] = *reinterpret_cast<const vec_t *>(
                          sX + (((kb * 4 + (l >> 4)) * BC) + swz(nt * 16 + cl)) * G);
                      if constexpr (LDSX)
                        {
                          // M^-1 of the operand's G rows (the product of the
                          // gather's scaling, bit for bit)
                          const vec_t dv =
                            *reinterpret_cast<const vec_t *>(sDsc + (kb * 4 + (l >> 4)) * G);
#pragma unroll
                          for (int nt = 0; nt < NTL; ++nt)
                            b[nt] = dv * b[nt];
                        }
#pragma unroll
                      for (int s = 0; s < G; ++s)
#pragma unroll
                        for (int t = 0; t < MTW; ++t)
                          if (w + 4 * t < mt)
#pragma unroll
                            for (int nt = 0; nt < NTL; ++nt)
                              acc[t][nt] = M::op(ring[kb][t][s], b[nt][s], acc[t][nt]);
                    }
              }
            else
              // ring[0] holds K block kb, ring[d] block kb + d
              for (int kb = 0; kb < nkb; ++kb)
              {
                vec_t b[NTL];
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt)
                  b[nt] = *reinterpret_cast<const vec_t *>(
                    sX + (((kb * 4 + (l >> 4)) * BC) + swz(nt * 16 + cl)) * G);
                if constexpr (LDSX)
                  {
                    const vec_t dv =
                      *reinterpret_cast<const vec_t *>(sDsc + (kb * 4 + (l >> 4)) * G);
#pragma unroll
                    for (int nt = 0; nt < NTL; ++nt)
                      b[nt] = dv * b[nt];
                  }
                vec_t a[MTW];
#pragma unroll
                for (int t = 0; t < MTW; ++t)
                  a[t] = ring[0][t];
#pragma unroll
                for (int d = 0; d + 1 < NRING; ++d)
#pragma unroll
                  for (int t = 0; t < MTW; ++t)
                    ring[d][t] = ring[d + 1][t];
                loadA(kb + NRING, ring[NRING - 1]);
                // one uniform test per M tile
#pragma unroll
                for (int t = 0; t < MTW; ++t)
                  if (w + 4 * t < mt)
#pragma unroll
                    for (int s = 0; s < G; ++s)
#pragma unroll
                      for (int nt = 0; nt < NTL; ++nt)
                        acc[t][nt] = M::op(a[t][s], b[nt][s], acc[t][nt]);
              }
            // the next column chunk starts again at K block 0
            if (!DFX_CELL_ORDER && c0 + BC < B)
#pragma unroll
              for (int d = 0; d < NRING; ++d)
                loadA(d, ring[d]);

            // ---- scatter: classical rows to y (stored on first touch, else
            // added to the old value), enriched rows to yE; with RES the rows
            // this cell finishes get the residual step instead
            T muT[RES ? NTL : 1];
            T a1 = 0, a2 = 0, cc = 0;
            if constexpr (RES)
              {
                a1 = (T)rf.coef[0];
                a2 = (T)rf.coef[1];
                cc = (T)rf.coef[2];
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt)
                  {
                    const int col = nt * 16 + cl;
                    muT[nt]       = (T)rf.mu[c0 + (col < bw ? col : 0)];
                  }
              }
#pragma unroll
            for (int t = 0; t < MTW; ++t)
              {
                if (w + 4 * t >= mt)
                  continue;
                T  oldLate[EARLY ? 1 : NTL][4];
                T(*o)[4] = EARLY ? old[EARLY ? t : 0] : oldLate;
                // with RES the old values are requested per row group
                constexpr bool OLDG = RES && !EARLY && DFX_RES_OLDG;
                if constexpr (!EARLY && !OLDG)
                  loadOld(t, oldLate);
                const bool enr = (w + 4 * t) * 16 + 15 >= nc; // wave-uniform
                // rows in groups of RG: the fused step's operands (0 on the
                // rows it skips) are requested per group, then the group is
                // stored
                constexpr int RG = RES ? CellPolicy<T, NT, MTW>::resRows : 4;
#pragma unroll
                for (int vg = 0; vg < 4; vg += RG)
                  {
                    if constexpr (OLDG)
                      loadOld(t, oldLate, vg, RG);
                    T    fx[RES ? NTL : 1][RG], fr[RES ? NTL : 1][RG], f0[RES ? NTL : 1][RG];
                    bool fin[RG];
                    // FP32 rows 4 (l / 16) + v of a lane are the G values of
                    // one LDS vector: a group's two rows are one 8-byte read
                    constexpr bool FXPAIR = LDSX && G == 4 && RG == 2;
                    if constexpr (FXPAIR)
                      {
                        typedef T pair_t __attribute__((ext_vector_type(2)));
                        const int row0 = rowOf(t, vg); // row0 % G == vg
#pragma unroll
                        for (int nt = 0; nt < NTL; ++nt)
                          {
                            const pair_t p = *reinterpret_cast<const pair_t *>(
                              sX + ((row0 / G) * BC + swz(nt * 16 + cl)) * G + row0 % G);
                            fx[nt][0] = p[0];
                            fx[nt][1] = p[1];
                          }
                      }
#pragma unroll
                    for (int g = 0; g < RG; ++g)
                      {
                        const int row = rowOf(t, vg + g);
                        fin[g] = RES && (row < nc) &&
                                 ((sLast[RES ? row >> 5 : 0] >> (row & 31)) & 1u);
                        if constexpr (RES)
                          {
                            const u32 xb = ((sMap[row] - lo) * ldx + c0 + cl) * S;
#pragma unroll
                            for (int nt = 0; nt < NTL; ++nt)
                              {
                                const u32 off = (fin[g] & (nt * 16 + cl < bw))
                                                  ? xb + nt * 16 * S
                                                  : OFF_NONE;
                                if constexpr (FXPAIR)
                                  ;
                                else if constexpr (LDSX) // row < KMAX: inside the tile
                                  fx[nt][g] = sX[((row / G) * BC + swz(nt * 16 + cl)) * G + row % G];
                                else
                                  fx[nt][g] = BF::load(rX, off);
                                fr[nt][g] = BF::load(rRes, off);
                                f0[nt][g] = BF::load(rR0, off);
                              }
                          }
                      }
#pragma unroll
                    for (int g = 0; g < RG; ++g)
                      {
                        const int vv  = vg + g;
                        const int row = rowOf(t, vv);
                        const u32 yb  = ((sMap[row] - lo) * ldy + c0 + cl) * S;
                        const u32 eb  = ((cm.enrOff + (u32)(row - nc)) * B + c0 + cl) * S;
                        const u32 xb  = ((sMap[row] - lo) * ldx + c0 + cl) * S;
#pragma unroll
                        for (int nt = 0; nt < NTL; ++nt)
                          {
                            const bool colOk = nt * 16 + cl < bw;
                            const T    yv    = o[nt][vv] + acc[t][nt][vv];
                            BF::store(yv, rY,
                                      ((row < nc) & colOk & !fin[g]) ? yb + nt * 16 * S
                                                                     : OFF_NONE);
                            if constexpr (RES)
                              BF::store(resUpdate<T>(yv, fx[nt][g], fr[nt][g], f0[nt][g],
                                                     (double)muT[nt], a1, a2, cc),
                                        rRes, (fin[g] & colOk) ? xb + nt * 16 * S : OFF_NONE);
                            if (enr)
                              BF::store(acc[t][nt][vv], rE,
                                        ((row >= nc) & (row < n) & colOk) ? eb + nt * 16 * S
                                                                          : OFF_NONE);
                          }
                      }
                  }
              }
            __syncthreads();
        };
        const u32 full = TAIL && NT > 1 ? B - B % BC : B;
        const u32 cBeg = nCh > 1 ? kOnly * BC : 0;
        const u32 cEnd = nCh > 1 ? min(full, cBeg + BC) : full;
        for (u32 c0 = cBeg; c0 < cEnd; c0 += BC)
          chunk(std::integral_constant<int, NT>{}, c0, min(BC, (int)(full - c0)));
        if constexpr (TAIL && NT > 1)
          if (full < B && (nCh <= 1 || cBeg >= full))
            chunk(std::integral_constant<int, 1>{}, full, (int)(B - full));
      }

      // One block (4 waves) per cell.  Wave w owns the M tiles w, w + 4, ...
      // (MTW of them) and all NT 16-column tiles of a column chunk.
      //
      // Latency hiding inside the block: the first D K blocks of the cell's A
      // fragments are requested right after the cell's metadata, before the
      // map and the gather; during the MFMA loop A runs D K blocks ahead; the
      // old values of the classical rows of Y (read-modify-write) are
      // requested before the MFMA loop when they fit in registers (EARLY).
      //
      // first[i / 32] bit i % 32: this cell is the first (in launch order) to
      // touch its classical row i, which is then stored instead of added, so Y
      // needs no zero fill and the first touch of every dof skips the read.
      template <class T, int NT, int MTW>
      __global__ __launch_bounds__(256, DFX_CELL_MINW) void cellApplyK(
        const CellMeta *__restrict__ meta, const u32 *__restrict__ first, u32 begin,
        const u32 *__restrict__ map, const T *__restrict__ frag,
        const T *__restrict__ x, u32 ldx, T *__restrict__ y, u32 ldy,
        T *__restrict__ yE, u32 B)
      {
        using M                 = Mfma<T>;
        using acc_t             = typename M::acc_t;
        using vec_t             = typename M::vec_t;
        constexpr int G         = M::G;
        constexpr int KB        = 4 * G;
        constexpr int BC        = 16 * NT;
        constexpr int VC        = BC / G;             // 16-byte vectors per chunk row
        constexpr int KMAX      = 64 * MTW;           // largest padded K of this class
        constexpr int NKB       = KMAX / KB;          // largest K-block count
        constexpr int D = NKB < DFX_CELL_DEPTH ? NKB : DFX_CELL_DEPTH; // A prefetch depth
        constexpr int GATHER_IT = (KMAX * VC + 255) / 256;
        constexpr int GATHER_BATCH = GATHER_IT < 8 ? GATHER_IT : 8; // loads in flight
        constexpr bool EARLY    = DFX_CELL_EARLY && MTW * NT * 4 * (int)sizeof(T) <= 128;
        extern __shared__ __align__(16) unsigned char smemRaw[];
        T *sX = reinterpret_cast<T *>(smemRaw);
        __shared__ u32 sMap[256];
        __shared__ u32 sFirst[8];

        const u32      pos  = begin + blockIdx.x;
        const CellMeta cm   = meta[pos];
        const int      n    = (int)cm.n;
        const int      nc   = (int)cm.nc;
        const int      mt   = (n + 15) >> 4;
        const int      kpad = ((n + KB - 1) / KB) * KB;
        const int      nkb  = kpad / KB;
        const int      w    = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
        const int      l    = threadIdx.x & 63;
        const T       *cfrag = frag + (sizeof(T) == 8 ? cm.fragOff64 : cm.fragOff32);

        auto loadA = [&](int kb, vec_t *dst) {
#pragma unroll
          for (int t = 0; t < MTW; ++t)
            {
              const int mtile = w + 4 * t;
              dst[t]          = vec_t{};
              if (mtile < mt && kb < nkb)
                dst[t] = *reinterpret_cast<const vec_t *>(
                  cfrag + ((u64)(mtile * nkb + kb) * 64 + l) * G);
            }
        };
        vec_t ring[D][MTW];
#pragma unroll
        for (int d = 0; d < D; ++d)
          loadA(d, ring[d]);

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
#pragma unroll
                for (int i0 = 0; i0 < GATHER_IT; i0 += GATHER_BATCH)
                  {
                    vec_t v[GATHER_BATCH];
#pragma unroll
                    for (int i = 0; i < GATHER_BATCH; ++i)
                      {
                        const int t  = threadIdx.x + (i0 + i) * 256;
                        const int kk = t / VC, c = (t - kk * VC) * G;
                        v[i]         = vec_t{};
                        if (i0 + i < GATHER_IT && kk < n && c < bw)
                          v[i] = *reinterpret_cast<const vec_t *>(
                            x + (u64)sMap[kk] * ldx + c0 + c);
                      }
#pragma unroll
                    for (int i = 0; i < GATHER_BATCH; ++i)
                      {
                        const int t  = threadIdx.x + (i0 + i) * 256;
                        const int kk = t / VC, c = (t - kk * VC) * G;
                        if (i0 + i < GATHER_IT && kk < kpad)
#pragma unroll
                          for (int g = 0; g < G; ++g)
                            sX[((kk / G) * BC + c + g) * G + (kk % G)] = v[i][g];
                      }
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

            // ---- old values of the classical rows (read-modify-write).  Row
            // (t, v) of this lane is dof sMap[row]; its NT columns are 16 apart.
            T   old[EARLY ? MTW : 1][EARLY ? NT : 1][4];
            const int cl = (l & 15);
            auto rowOf = [&](int t, int v) { return (w + 4 * t) * 16 + M::row(l, v); };
            auto loadOld = [&](int t, T (*o)[4]) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                {
                  const int row  = rowOf(t, v);
                  const bool rmw = w + 4 * t < mt && row < nc &&
                                   !((sFirst[row >> 5] >> (row & 31)) & 1u);
                  const T *src = y + (u64)sMap[rmw ? row : 0] * ldy + c0 + cl;
#pragma unroll
                  for (int nt = 0; nt < NT; ++nt)
                    o[nt][v] = (rmw && nt * 16 + cl < bw) ? src[nt * 16] : (T)0;
                }
            };
            if constexpr (EARLY)
              {
#pragma unroll
                for (int t = 0; t < MTW; ++t)
                  loadOld(t, old[t]);
              }
            __syncthreads();

            acc_t acc[MTW][NT];
#pragma unroll
            for (int t = 0; t < MTW; ++t)
#pragma unroll
              for (int nt = 0; nt < NT; ++nt)
                acc[t][nt] = acc_t{0, 0, 0, 0};

            // ring[0] holds K block kb; ring[d] block kb + d (a rotating
            // register window, so every index is static)
            for (int kb = 0; kb < nkb; ++kb)
              {
                vec_t b[NT];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                  b[nt] = *reinterpret_cast<const vec_t *>(
                    sX + (((kb * 4 + (l >> 4)) * BC) + nt * 16 + (l & 15)) * G);
                vec_t a[MTW];
#pragma unroll
                for (int t = 0; t < MTW; ++t)
                  a[t] = ring[0][t];
#pragma unroll
                for (int d = 0; d + 1 < D; ++d)
#pragma unroll
                  for (int t = 0; t < MTW; ++t)
                    ring[d][t] = ring[d + 1][t];
                loadA(kb + D, ring[D - 1]);
#pragma unroll
                for (int s = 0; s < G; ++s)
#pragma unroll
                  for (int t = 0; t < MTW; ++t)
                    if (w + 4 * t < mt)
#pragma unroll
                      for (int nt = 0; nt < NT; ++nt)
                        acc[t][nt] = M::op(a[t][s], b[nt][s], acc[t][nt]);
              }
            // the next column chunk starts again at K block 0
            if (c0 + BC < B)
#pragma unroll
              for (int d = 0; d < D; ++d)
                loadA(d, ring[d]);

            // ---- scatter: classical rows into y (store on first touch, else
            // add to the old value), enriched rows into yE
#pragma unroll
            for (int t = 0; t < MTW; ++t)
              {
                if (w + 4 * t >= mt)
                  continue;
                T  oldLate[EARLY ? 1 : NT][4];
                T(*o)[4] = EARLY ? old[EARLY ? t : 0] : oldLate;
                if constexpr (!EARLY)
                  loadOld(t, oldLate);
#pragma unroll
                for (int v = 0; v < 4; ++v)
                  {
                    const int row = rowOf(t, v);
                    if (row >= n)
                      continue;
                    T *dst = row < nc ? y + (u64)sMap[row] * ldy + c0 + cl
                                      : yE + (u64)(cm.enrOff + (u32)(row - nc)) * B + c0 + cl;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                      if (nt * 16 + cl < bw)
                        dst[nt * 16] = (row < nc ? o[nt][v] : (T)0) + acc[t][nt][v];
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
        // slot indices are staged in LDS per chunk of IDX (one coalesced
        // load instead of a dependent index load per value); lane sl adds
        // slots sl, sl + SL, ... in ascending order (IDX is a multiple of SL)
        constexpr int CB = 16, SL = 16, U = 16, IDX = 2048;
        __shared__ T   part[SL][CB];
        __shared__ u32 sIdx[IDX];
        const u32 g  = blockIdx.x;
        const int cj = threadIdx.x % CB, sl = threadIdx.x / CB;
        const u32 j  = blockIdx.y * CB + cj;
        const u32 e0 = off[g], e1 = off[g + 1];
        T         acc = 0;
        for (u32 c0 = e0; c0 < e1; c0 += IDX)
          {
            const u32 cn = min((u32)IDX, e1 - c0);
            __syncthreads(); // the previous chunk's indices are no longer read
            for (u32 t = threadIdx.x; t < cn; t += blockDim.x)
              sIdx[t] = slots[c0 + t];
            __syncthreads();
            if (j < B)
              for (u32 e = sl; e < cn; e += SL * U)
                {
                  T v[U];
#pragma unroll
                  for (int u = 0; u < U; ++u)
                    {
                      const u32 ee = e + u * SL;
                      v[u]         = ee < cn ? yE[(u64)sIdx[ee] * B + j] : (T)0;
                    }
#pragma unroll
                  for (int u = 0; u < U; ++u)
                    acc += v[u];
                }
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

    } // namespace

    template <class T>
    bool cellBufferPath(const dfx_op *op, View<const T> x, View<const T> y)
    {
      constexpr int G  = Mfma<T>::G;
      const u64     S  = sizeof(T);
      const u64 nE     = op->nDofs - op->nClassical;
      const u64 span   = op->maxRowSpan;
      const u64 eBytes = (u64)op->nEnrSlots * x.B * S;
      return span * x.ld * S < (1ull << 31) && span * y.ld * S < (1ull << 31) &&
             nE * x.ld * S < (1ull << 31) && eBytes < (1ull << 31) && x.B % G == 0 &&
             x.ld % G == 0 && (reinterpret_cast<uintptr_t>(x.p) & 15) == 0;
    }
    template bool cellBufferPath<double>(const dfx_op *, View<const double>,
                                         View<const double>);
    template bool cellBufferPath<float>(const dfx_op *, View<const float>, View<const float>);

    namespace
    {
      template <class T, int NT, int MTW>
      void launchCell(hipStream_t s, const dfx_op *op, const Launch &L, const T *frag,
                      View<const T> x, View<T> y, T *yE, const double *dsc, const T *xe,
                      const ResFuse<T> *rf)
      {
        constexpr int BC = 16 * NT;
        const u64     S  = sizeof(T);
        const u64 eBytes = (u64)op->nEnrSlots * x.B * S;
        const bool buf   = cellBufferPath<T>(op, x, View<const T>{y.p, y.N, y.B, y.ld});
        // dynamic tile plus the static map / first-touch words (about 1 KiB)
        const size_t lds = (size_t)(buf ? 64 * MTW : L.kpadMax) * BC * S;
        DFX_REQUIRE(lds + 2048 <= 160 * 1024, "cell tile exceeds the 160 KiB LDS of a CU");
        // a last chunk of at most 16 columns: the kernel with the narrow tail
        // (NT = 5 serves 65..80 columns in one chunk, never with a tail)
        constexpr bool TAILOK = NT != 1 && NT != 5;
        const bool  tail  = TAILOK && buf && x.B % BC != 0 && x.B % BC <= 16;
        DFX_REQUIRE(!rf || (dsc && buf && op->last.p),
                    "the fused residual step needs the folded buffer-addressed cell kernel");
        const int   which = !buf ? 0 : rf ? 5 + (tail ? 1 : 0) : (dsc ? 2 : 1) + (tail ? 2 : 0);
        const void *fns[7] = {
          reinterpret_cast<const void *>(&cellApplyK<T, NT, MTW>),
          reinterpret_cast<const void *>(&cellApplyBufK<T, NT, MTW, false, false, false>),
          reinterpret_cast<const void *>(&cellApplyBufK<T, NT, MTW, true, false, false>),
          reinterpret_cast<const void *>(&cellApplyBufK<T, NT, MTW, false, TAILOK, false>),
          reinterpret_cast<const void *>(&cellApplyBufK<T, NT, MTW, true, TAILOK, false>),
          reinterpret_cast<const void *>(&cellApplyBufK<T, NT, MTW, true, false, true>),
          reinterpret_cast<const void *>(&cellApplyBufK<T, NT, MTW, true, TAILOK, true>)};
        const void *fn = fns[which];
        // raised once per kernel to the largest tile seen (several host
        // threads may launch concurrently: serving slots)
        static std::atomic<size_t> attrSet[7] = {60 * 1024, 60 * 1024, 60 * 1024, 60 * 1024,
                                                 60 * 1024, 60 * 1024, 60 * 1024};
        size_t cur = attrSet[which].load();
        if (lds > cur)
          {
            DFX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds));
            while (lds > cur && !attrSet[which].compare_exchange_weak(cur, lds))
              {
              }
          }
        DFX_REQUIRE(buf || !dsc, "folded diagonal needs the buffer-addressed cell kernel");
        // column chunks of a cell (the narrow tail counts as one); with
        // DFX_CELL_SPLIT (default on) each gets a workgroup of its own
        const char *splitVar = getenv("DFX_CELL_SPLIT"); // read per launch
        const bool  splitEnv = !splitVar || atoi(splitVar) != 0;
        const u32 nCellsL = L.end - L.begin;
        const u32 fullC   = tail ? x.B - x.B % BC : x.B;
        const u32 nChunk  = (fullC + BC - 1) / BC + (tail ? 1u : 0u);
        const u32 nCh     = splitEnv && buf && nChunk > 1 ? nChunk : 1u;
        const u32 grid    = nCh > 1 ? (nCellsL + 7) / 8 * 8 * nCh : nCellsL;
        auto bufLaunch = [&](auto dscTag, auto tailTag, auto resTag) {
          constexpr bool D = decltype(dscTag)::value, TL = decltype(tailTag)::value;
          constexpr bool R = decltype(resTag)::value;
          cellApplyBufK<T, NT, MTW, D, TL, R><<<grid, 256, lds, s>>>(
            op->meta.as<CellMeta>(), op->first.as<u32>(), L.begin, op->map.as<u32>(), frag,
            x.p, x.N, x.ld, y.p, y.N, y.ld, yE, (u32)eBytes, x.B, op->nClassical,
            D ? dsc : nullptr, D && xe ? xe : x.p, R ? *rf : ResFuse<T>{},
            R ? op->last.as<u32>() : nullptr, nCh, nCellsL);
        };
        using F  = std::false_type;
        using Tr = std::true_type;
        using Tt = std::integral_constant<bool, TAILOK>;
        if (which == 1)
          bufLaunch(F{}, F{}, F{});
        else if (which == 2)
          bufLaunch(Tr{}, F{}, F{});
        else if (which == 3)
          bufLaunch(F{}, Tt{}, F{});
        else if (which == 4)
          bufLaunch(Tr{}, Tt{}, F{});
        else if (which == 5)
          bufLaunch(Tr{}, F{}, Tr{});
        else if (which == 6)
          bufLaunch(Tr{}, Tt{}, Tr{});
        else
          cellApplyK<T, NT, MTW><<<L.end - L.begin, 256, lds, s>>>(
            op->meta.as<CellMeta>(), op->first.as<u32>(), L.begin, op->map.as<u32>(), frag,
            x.p, x.ld, y.p, y.ld, yE, x.B);
        checkLaunch("cellApply");
      }

      template <class T, int NT>
      void launchCellMtw(hipStream_t s, const dfx_op *op, const Launch &L,
                         const T *frag, View<const T> x, View<T> y, T *yE,
                         const double *dsc, const T *xe, const ResFuse<T> *rf)
      {
        switch (L.mtw)
          {
            case 1: return launchCell<T, NT, 1>(s, op, L, frag, x, y, yE, dsc, xe, rf);
            case 2: return launchCell<T, NT, 2>(s, op, L, frag, x, y, yE, dsc, xe, rf);
            case 3: return launchCell<T, NT, 3>(s, op, L, frag, x, y, yE, dsc, xe, rf);
            case 4: return launchCell<T, NT, 4>(s, op, L, frag, x, y, yE, dsc, xe, rf);
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
      // FP32, 65..80 columns: one 80-column chunk (the cell matrix read
      // once); the FP64 tile of 80 columns would leave one workgroup per CU
      if (DFX_CELL_NT5 && sizeof(T) == 4 && nt == 5)
        return 5;
      if (nt > cap)
        nt = cap;
      if (nt == 3)
        nt = 4;
      if (nt == 5) // 65..80 columns (FP64): a 64-column chunk and a narrow tail
        nt = 4;
      if (nt > 4 && nt < 8)
        nt = 8;
      return nt;
    }

    template <class T>
    int builtinNT(u32 B, u32 kpadMax)
    {
      return chooseNT<T>(B, kpadMax);
    }

    template <class T>
    void cellApply(hipStream_t s, const dfx_op *op, const Launch &L, const T *frag,
                   View<const T> x, View<T> y, T *yE, const double *dsc, const T *xe, int nt,
                   const ResFuse<T> *rf)
    {
      if (L.end <= L.begin || !x.B)
        return;
      // explicit width, else the autotuner's for (scalar size, B, class) unless
      // the environment caps the width, else the built-in choice
      if (!nt && !ntOverride(sizeof(T) == 8))
        nt = op->tunedNT(sizeof(T), x.B, L.mtw);
      if (!nt)
        nt = chooseNT<T>(x.B, L.kpadMax);
      switch (nt)
        {
          case 1: return launchCellMtw<T, 1>(s, op, L, frag, x, y, yE, dsc, xe, rf);
          case 2: return launchCellMtw<T, 2>(s, op, L, frag, x, y, yE, dsc, xe, rf);
          case 4: return launchCellMtw<T, 4>(s, op, L, frag, x, y, yE, dsc, xe, rf);
          case 8: return launchCellMtw<T, 8>(s, op, L, frag, x, y, yE, dsc, xe, rf);
          case 5: return launchCellMtw<T, 5>(s, op, L, frag, x, y, yE, dsc, xe, rf);
        }
      throw Error(DFX_E_ARG, "column-chunk width must be 1, 2, 4, 5 or 8 tiles of 16");
    }

    namespace
    {
      __global__ void gatherDiagK(const double *d, const u32 *map, u64 n, u32 nC, double *out)
      {
        for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < n;
             t += (u64)gridDim.x * blockDim.x)
          {
            const u32 m = map[t];
            out[t]      = m < nC ? d[m] : 1.0;
          }
      }
    } // namespace

    void gatherDiag(hipStream_t s, const double *d, const u32 *map, u64 n, u32 nC,
                    double *out)
    {
      if (!n)
        return;
      const u64 g = (n + 255) / 256;
      gatherDiagK<<<(unsigned)(g > 65536 ? 65536 : g), 256, 0, s>>>(d, map, n, nC, out);
      checkLaunch("gatherDiag");
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

    namespace
    {
      // one block per cell (launch order): diagonal entry i of the cell's
      // FP64 fragment-ordered matrix.  Entry (lane l, value v) of block
      // (mt, kb) holds A_e^T[mt 16 + l % 16][kb 8 + 2 (l / 16) + v], so A_e(i, i)
      // sits at mt = i / 16, kb = i / 8, l = 16 ((i % 8) / 2) + i % 16, v = i % 2
      __global__ void cellDiagK(const CellMeta *__restrict__ meta,
                                const double *__restrict__ frag, double *__restrict__ out)
      {
        const CellMeta cm  = meta[blockIdx.x];
        const u32      nkb = (cm.n + 7) / 8;
        for (u32 i = threadIdx.x; i < cm.n; i += blockDim.x)
          {
            const u32 mt = i / 16, kb = i / 8, l = 16 * ((i % 8) / 2) + i % 16, v = i % 2;
            out[cm.mapOff + i] = frag[cm.fragOff64 + ((u64)(mt * nkb + kb) * 64 + l) * 2 + v];
          }
      }
    } // namespace

    void cellDiag(hipStream_t s, const dfx_op *op, double *out)
    {
      if (!op->nCells)
        return;
      cellDiagK<<<op->nCells, 256, 0, s>>>(op->meta.as<CellMeta>(), op->frag64.as<double>(),
                                           out);
      checkLaunch("cellDiag");
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
                                    double *, const double *, const double *, int,
                                    const ResFuse<double> *);
    template void cellApply<float>(hipStream_t, const dfx_op *, const Launch &,
                                   const float *, View<const float>, View<float>, float *,
                                   const double *, const float *, int,
                                   const ResFuse<float> *);
    template int builtinNT<double>(u32, u32);
    template int builtinNT<float>(u32, u32);
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
