// CPU timing of the cell-wise operator apply (test infrastructure: used only as
// the CPU baseline that GPU throughput is reported beside; never a product path).
//
// What is timed is the CPU form of KohnShamOperatorContextFE::apply
// (src/ksdft/KohnShamOperatorContextFE.t.cpp:1326-1415) as a threaded
// domain-decomposed CPU run performs it: every thread owns a contiguous slab
// of cells (the stand-in for one MPI rank's locally owned cells), gathers its
// cells' X_e through the cell->dof map (copyFieldToCellWiseData,
// src/basis/FECellWiseDataOperations.t.cpp:60-84), forms Y_e = X_e A_e
// (BlasLapack.t.cpp:391-436, here a vectorizable loop nest instead of a BLAS
// call) and adds Y_e into a thread-local vector over the dofs it touches
// (addCellWiseDataToFieldData, FECellWiseDataOperations.t.cpp:89-117); the
// thread-local vectors are then summed into Y in parallel over dof ranges
// (the stand-in for accumulateAddLocallyOwned).  The O(N B) vector work
// (constraint distribution, zeroing Y) is timed separately on the full
// vectors, since a sample of cells does not scale it.
#include "oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

namespace oracle
{
  namespace
  {
    // Y (B x n, column-major) = X (B x n) A (n x n, column-major); the inner
    // loop runs over the B contiguous columns of the block and vectorizes.
    __attribute__((target_clones("avx512f", "avx2", "default"))) void
    cellGemm(int B, int n, const double *X, const double *A, double *Y)
    {
      for (int j = 0; j < n; ++j)
        {
          double *y = Y + (size_t)B * j;
          for (int i = 0; i < B; ++i)
            y[i] = 0.0;
          for (int l = 0; l < n; ++l)
            {
              const double  a = A[l + (size_t)n * j];
              const double *x = X + (size_t)B * l;
              for (int i = 0; i < B; ++i)
                y[i] += a * x[i];
            }
        }
    }

    struct Slab
    {
      u32                 e0 = 0, e1 = 0;
      std::vector<u32>    dofs;     // sorted global dofs touched by the slab
      std::vector<u32>    localMap; // per cell slot: index into dofs
      std::vector<double> y;        // dofs.size() x B
    };
  } // namespace
} // namespace oracle

extern "C"
{
  // Median seconds per apply over `reps` runs (one untimed run first);
  // tCells / tVec receive the medians of the two parts.
  double orc_time_apply(void *h, int kind, double *X, double *Y, uint32_t B,
                        int reps, int threads, double *tCells, double *tVec)
  {
    using namespace oracle;
    Problem      &P    = *static_cast<Problem *>(h);
    const double *mats = kind == OP_M ? P.M.data() : (kind == OP_K ? P.K.data() : P.H.data());
    const int     T    = std::max(1, threads);
    std::vector<Slab> slab(T);
    for (int t = 0; t < T; ++t)
      {
        Slab &s = slab[t];
        s.e0    = (u32)((u64)P.nCells * t / T);
        s.e1    = (u32)((u64)P.nCells * (t + 1) / T);
        for (u32 e = s.e0; e < s.e1; ++e)
          for (u32 i = 0; i < P.cellNdofs[e]; ++i)
            s.dofs.push_back(P.cellMap[P.cellMapOffset[e] + i]);
        std::sort(s.dofs.begin(), s.dofs.end());
        s.dofs.erase(std::unique(s.dofs.begin(), s.dofs.end()), s.dofs.end());
        for (u32 e = s.e0; e < s.e1; ++e)
          for (u32 i = 0; i < P.cellNdofs[e]; ++i)
            s.localMap.push_back((u32)(std::lower_bound(s.dofs.begin(), s.dofs.end(),
                                                        P.cellMap[P.cellMapOffset[e] + i]) -
                                       s.dofs.begin()));
        s.y.resize(s.dofs.size() * (size_t)B);
      }

    auto parallel = [&](auto &&fn) {
      std::vector<std::thread> th;
      for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] { fn(t); });
      for (auto &x : th)
        x.join();
    };

    std::vector<double> tc, tv;
    for (int r = 0; r < reps + 1; ++r)
      {
        // ---- O(N B) vector work on the full vectors (constraints into X,
        // zeroing Y; the constraint reduction of Y is timed at the end)
        auto t0 = std::chrono::steady_clock::now();
        distributeParentToChild(P.cons, X, B);
        const size_t nY = (size_t)P.nDofs * B;
        parallel([&](int t) {
          const size_t a = nY * t / T, b = nY * (t + 1) / T;
          std::memset(Y + a, 0, (b - a) * sizeof(double));
        });
        auto t1 = std::chrono::steady_clock::now();

        // ---- cell work: gather, per-cell product, local scatter, merge
        parallel([&](int t) {
          Slab &s = slab[t];
          std::fill(s.y.begin(), s.y.end(), 0.0);
          std::vector<double> xc, yc;
          size_t slot = 0;
          for (u32 e = s.e0; e < s.e1; ++e)
            {
              const u32 n = P.cellNdofs[e];
              xc.resize((size_t)n * B);
              yc.resize((size_t)n * B);
              const u32 *m = &P.cellMap[P.cellMapOffset[e]];
              for (u32 i = 0; i < n; ++i)
                std::memcpy(&xc[(size_t)i * B], X + (size_t)m[i] * B, sizeof(double) * B);
              cellGemm((int)B, (int)n, xc.data(), mats + P.cellMatOffset[e], yc.data());
              for (u32 i = 0; i < n; ++i)
                {
                  double       *dst = &s.y[(size_t)s.localMap[slot + i] * B];
                  const double *src = &yc[(size_t)i * B];
                  for (u32 b = 0; b < B; ++b)
                    dst[b] += src[b];
                }
              slot += n;
            }
        });
        parallel([&](int t) {
          const u32 d0 = (u32)((u64)P.nDofs * t / T), d1 = (u32)((u64)P.nDofs * (t + 1) / T);
          for (const Slab &s : slab)
            {
              size_t k = std::lower_bound(s.dofs.begin(), s.dofs.end(), d0) - s.dofs.begin();
              for (; k < s.dofs.size() && s.dofs[k] < d1; ++k)
                {
                  double       *dst = Y + (size_t)s.dofs[k] * B;
                  const double *src = &s.y[k * B];
                  for (u32 b = 0; b < B; ++b)
                    dst[b] += src[b];
                }
            }
        });
        auto t2 = std::chrono::steady_clock::now();
        distributeChildToParent(P.cons, Y, B);
        auto t3 = std::chrono::steady_clock::now();
        if (r > 0)
          {
            tv.push_back(std::chrono::duration<double>((t1 - t0) + (t3 - t2)).count());
            tc.push_back(std::chrono::duration<double>(t2 - t1).count());
          }
      }
    auto med = [](std::vector<double> v) {
      std::sort(v.begin(), v.end());
      return v.empty() ? 0.0 : v[v.size() / 2];
    };
    if (tCells)
      *tCells = med(tc);
    if (tVec)
      *tVec = med(tv);
    return med(tc) + med(tv);
  }
}
