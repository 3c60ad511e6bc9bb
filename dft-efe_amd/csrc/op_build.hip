// Construction of the device-resident cell-wise operator: validation of the
// cell -> dof map, greedy colouring of the cell-adjacency graph (cells that
// share a classical dof get different colours, so one colour's scatter-add is
// race-free), size classes, launch order, fragment offsets, the enriched-dof
// reduction lists and the dof -> slot lists of the stand-alone scatter.
#include <algorithm>
#include <numeric>

#include "api_util.hpp"
#include "kernels.hpp"

namespace dfx
{
  DevBuf &work(dfx_ctx *ctx, const char *name, size_t bytes)
  {
    DevBuf &b = ctx->work[name];
    b.ensure(bytes);
    return b;
  }

  template <class T>
  void colDotsView(dfx_ctx *ctx, View<const T> X, View<const T> Y, double *hostOut)
  {
    const u32 B = X.B;
    ctx->reducePartials.ensure((size_t)1024 * B * sizeof(double) + B * sizeof(double));
    double *part = ctx->reducePartials.as<double>();
    double *out  = part + (size_t)1024 * B;
    k::colDots<T>(ctx->stream, X, Y, part, out);
    DFX_HIP(hipMemcpyAsync(hostOut, out, B * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    DFX_HIP(hipStreamSynchronize(ctx->stream));
  }
  template void colDotsView<double>(dfx_ctx *, View<const double>, View<const double>,
                                    double *);
  template void colDotsView<float>(dfx_ctx *, View<const float>, View<const float>,
                                   double *);

  void colDotsHost(dfx_ctx *ctx, const dfx_mv *X, const dfx_mv *Y, double *out)
  {
    if (X->dtype == DFX_F64)
      colDotsView<double>(ctx, View<const double>{(const double *)X->d, X->N, X->B, X->B},
                          View<const double>{(const double *)Y->d, Y->N, Y->B, Y->B}, out);
    else
      colDotsView<float>(ctx, View<const float>{(const float *)X->d, X->N, X->B, X->B},
                         View<const float>{(const float *)Y->d, Y->N, Y->B, Y->B}, out);
  }

  void buildCellwise(dfx_ctx *ctx, dfx_op *op, u32 nCells, const u32 *cellNdofs,
                     const u32 *map, u32 nDofs, u32 nClassical, dfx_constraints *cons,
                     u32 flags)
  {
    DFX_REQUIRE(nClassical <= nDofs, "nClassical > nDofs");
    DFX_REQUIRE(!cons || cons->nDofs == nDofs, "constraints built for another N");
    op->ctx        = ctx;
    op->kind       = OP_CELLWISE;
    op->nDofs      = nDofs;
    op->nClassical = nClassical;
    op->cons       = cons;
    consRetain(cons);
    op->nCells     = nCells;
    op->hasF32     = (flags & DFX_OP_KEEP_FP32) != 0;
    op->hCellNdofs.assign(cellNdofs, cellNdofs + nCells);
    op->hMapOffset.resize(nCells + 1);
    op->hMatOffset.resize(nCells + 1);
    op->hMapOffset[0] = 0;
    op->hMatOffset[0] = 0;
    u64 sumNe = 0, sumNe2 = 0;
    u32 maxN  = 0;
    for (u32 e = 0; e < nCells; ++e)
      {
        const u32 n = cellNdofs[e];
        DFX_REQUIRE(n > 0 && n <= 256, "cell dof count must be in 1..256");
        sumNe += n;
        sumNe2 += (u64)n * n;
        maxN = std::max(maxN, n);
        DFX_REQUIRE(sumNe < (1ull << 32), "cell->dof map longer than 2^32");
        op->hMapOffset[e + 1] = (u32)sumNe;
        op->hMatOffset[e + 1] = sumNe2;
      }
    op->sumNe  = sumNe;
    op->sumNe2 = sumNe2;
    op->maxN   = maxN;

    // classical entries first in every cell; nc = their count
    std::vector<u32> nc(nCells);
    for (u32 e = 0; e < nCells; ++e)
      {
        u32  cnt = 0;
        bool enr = false;
        for (u32 s = op->hMapOffset[e]; s < op->hMapOffset[e + 1]; ++s)
          {
            DFX_REQUIRE(map[s] < nDofs, "cell->dof map entry out of range");
            if (map[s] < nClassical)
              {
                DFX_REQUIRE(!enr, "classical dofs must precede enriched dofs in a cell");
                ++cnt;
              }
            else
              enr = true;
          }
        nc[e] = cnt;
      }
    {
      // a dof may appear once per cell (two lanes would update one row)
      std::vector<u32> last(nDofs, 0xffffffffu);
      for (u32 e = 0; e < nCells; ++e)
        for (u32 s = op->hMapOffset[e]; s < op->hMapOffset[e + 1]; ++s)
          {
            DFX_REQUIRE(last[map[s]] != e, "a dof appears twice in one cell's map");
            last[map[s]] = e;
          }
    }

    // ---- greedy colouring over shared classical dofs (cell order)
    std::vector<u32> dofCellOff(nClassical + 1, 0), dofCells;
    for (u32 e = 0; e < nCells; ++e)
      for (u32 s = op->hMapOffset[e]; s < op->hMapOffset[e] + nc[e]; ++s)
        ++dofCellOff[map[s] + 1];
    for (u32 d = 0; d < nClassical; ++d)
      dofCellOff[d + 1] += dofCellOff[d];
    dofCells.resize(dofCellOff[nClassical]);
    {
      std::vector<u32> fillp(dofCellOff.begin(), dofCellOff.end() - 1);
      for (u32 e = 0; e < nCells; ++e)
        for (u32 s = op->hMapOffset[e]; s < op->hMapOffset[e] + nc[e]; ++s)
          dofCells[fillp[map[s]]++] = e;
    }
    std::vector<int> color(nCells, -1);
    std::vector<u32> stamp(64, 0xffffffffu);
    int nColors = 0;
    for (u32 e = 0; e < nCells; ++e)
      {
        for (u32 s = op->hMapOffset[e]; s < op->hMapOffset[e] + nc[e]; ++s)
          for (u32 q = dofCellOff[map[s]]; q < dofCellOff[map[s] + 1]; ++q)
            {
              const int c = color[dofCells[q]];
              if (c >= 0)
                {
                  if ((size_t)c >= stamp.size())
                    stamp.resize(c + 1, 0xffffffffu);
                  stamp[c] = e;
                }
            }
        int c = 0;
        while ((size_t)c < stamp.size() && stamp[c] == e)
          ++c;
        color[e] = c;
        nColors  = std::max(nColors, c + 1);
      }
    op->nColors = (u32)nColors;

    // ---- launch order: (colour, M-tile class, cell)
    auto mtwOf = [](u32 n) { return (((n + 15) / 16) + 3) / 4; };
    op->hOrder.resize(nCells);
    std::iota(op->hOrder.begin(), op->hOrder.end(), 0u);
    std::stable_sort(op->hOrder.begin(), op->hOrder.end(), [&](u32 a, u32 b) {
      if (color[a] != color[b])
        return color[a] < color[b];
      return mtwOf(cellNdofs[a]) > mtwOf(cellNdofs[b]);
    });

    // enriched partial rows, in caller cell order
    std::vector<u32> enrOff(nCells + 1, 0);
    for (u32 e = 0; e < nCells; ++e)
      enrOff[e + 1] = enrOff[e] + (cellNdofs[e] - nc[e]);
    op->nEnrSlots = enrOff[nCells];

    op->hMeta.resize(nCells);
    u64 f64 = 0, f32 = 0;
    for (u32 q = 0; q < nCells; ++q)
      {
        const u32 e = op->hOrder[q];
        const u32 n = cellNdofs[e];
        CellMeta  m;
        m.n         = n;
        m.nc        = nc[e];
        m.mapOff    = op->hMapOffset[e];
        m.enrOff    = enrOff[e];
        m.fragOff64 = f64;
        m.fragOff32 = f32;
        u32 lo = 0xffffffffu, hi = 0;
        for (u32 i = 0; i < nc[e]; ++i)
          {
            lo = std::min(lo, map[op->hMapOffset[e] + i]);
            hi = std::max(hi, map[op->hMapOffset[e] + i]);
          }
        m.rowLo         = nc[e] ? lo : 0;
        m.pad_          = 0;
        op->maxRowSpan  = std::max(op->maxRowSpan, nc[e] ? hi - lo + 1 : 0u);
        const u64 mpad = 16 * (u64)((n + 15) / 16);
        f64 += mpad * (8 * (u64)((n + 7) / 8));
        f32 += mpad * (16 * (u64)((n + 15) / 16));
        op->hMeta[q] = m;
      }
    op->fragLen64 = f64;
    op->fragLen32 = f32;

    // one launch per (colour, size class); an operator of few cells (the
    // launches could not fill the GPU anyway, and every launch costs a
    // dispatch) runs each colour as one launch with the largest class's tile.
    // FP32 launches also merge the classes of a colour whenever the largest
    // class fits cells of up to 192 dofs (measured on C4: one launch per
    // colour instead of two side by side takes the FP32 apply from 1.39 to
    // 1.30 ms; the FP64 apply, whose classes use different column-chunk
    // widths, gets slower merged).  DFX_MERGE_CLASSES_BELOW sets the cell
    // count below which both merge (and nothing else does)
    const char *mergeEnv   = getenv("DFX_MERGE_CLASSES_BELOW");
    const u32   mergeBelow = mergeEnv ? (u32)atol(mergeEnv) : 4096u;
    u32         maxMtw     = 0;
    for (u32 e = 0; e < nCells; ++e)
      maxMtw = std::max(maxMtw, mtwOf(cellNdofs[e]));
    const bool merge64 = nCells < mergeBelow;
    const bool merge32 = mergeEnv ? merge64 : (merge64 || maxMtw <= 3);
    auto buildLaunches = [&](std::vector<Launch> &out, bool merge, bool f64) {
      out.clear();
      for (u32 q = 0; q < nCells;)
        {
          const u32 e0 = op->hOrder[q];
          u32       r  = q;
          u32 kp = 0, mtw = 0;
          u64 s2 = 0, s1 = 0, sc = 0;
          while (r < nCells && color[op->hOrder[r]] == color[e0] &&
                 (merge || mtwOf(cellNdofs[op->hOrder[r]]) == mtwOf(cellNdofs[e0])))
            {
              const u32 n = cellNdofs[op->hOrder[r]];
              kp          = std::max(kp, f64 ? 8 * ((n + 7) / 8) : 16 * ((n + 15) / 16));
              mtw         = std::max(mtw, mtwOf(n));
              s2 += (u64)n * n;
              s1 += n;
              sc += nc[op->hOrder[r]];
              ++r;
            }
          out.push_back({q, r, mtw, kp, s2, s1, sc, 0, (u32)color[e0]});
          q = r;
        }
    };
    buildLaunches(op->launches64, merge64, true);
    buildLaunches(op->launches32, merge32, false);

    // first-touch bits in launch order: the first cell to reach a classical
    // dof stores its contribution, later cells add to it
    std::vector<u32> hFirst((size_t)nCells * 8, 0u);
    {
      std::vector<unsigned char> seen(nClassical, 0);
      for (u32 q = 0; q < nCells; ++q)
        {
          const u32 e = op->hOrder[q];
          for (u32 i = 0; i < nc[e]; ++i)
            {
              const u32 d = map[op->hMapOffset[e] + i];
              if (!seen[d])
                {
                  seen[d] = 1;
                  hFirst[(size_t)q * 8 + i / 32] |= 1u << (i % 32);
                }
            }
        }
      op->fillY = std::find(seen.begin(), seen.end(), 0) != seen.end();
      for (auto *L : {&op->launches64, &op->launches32})
        for (Launch &l : *L)
          {
            l.sumFirst = 0;
            for (u64 w = (u64)l.begin * 8; w < (u64)l.end * 8; ++w)
              l.sumFirst += (u64)__builtin_popcount(hFirst[w]);
          }
    }

    // last-touch bits in launch order (the residual step fused into the
    // cell apply): the last cell to reach an unconstrained classical dof
    // finishes it; every other row is listed for the stand-alone step
    std::vector<u32> hLast((size_t)nCells * 8, 0u), hResRows;
    {
      // 0: in no cell, 1: finished by a cell, 2: constrained
      std::vector<unsigned char> st(nClassical, 0);
      if (cons)
        for (u32 r : cons->hRows)
          if (r < nClassical)
            st[r] = 2;
      for (u32 q = nCells; q-- > 0;)
        {
          const u32 e = op->hOrder[q];
          for (u32 i = 0; i < nc[e]; ++i)
            {
              const u32 d = map[op->hMapOffset[e] + i];
              if (!st[d])
                {
                  st[d] = 1;
                  hLast[(size_t)q * 8 + i / 32] |= 1u << (i % 32);
                }
            }
        }
      // homogeneous Dirichlet rows are zero in every residual-filter buffer
      // from the seed on (A and M zero them) and the step keeps them zero,
      // so they are left out; other constrained rows and rows in no cell
      // are listed
      const bool zeroRows = cons && cons->homDirichlet;
      for (u32 d = 0; d < nClassical; ++d)
        if (st[d] == 0 || (st[d] == 2 && !zeroRows))
          hResRows.push_back(d);
      for (u32 d = nClassical; d < nDofs; ++d)
        hResRows.push_back(d);
      op->nResRows = (u32)hResRows.size();
    }

    // enriched dof -> slots (cell order)
    const u32 nE = nDofs - nClassical;
    std::vector<u32> eOff(nE + 1, 0), eSlots(op->nEnrSlots);
    for (u32 e = 0; e < nCells; ++e)
      for (u32 i = nc[e]; i < cellNdofs[e]; ++i)
        ++eOff[map[op->hMapOffset[e] + i] - nClassical + 1];
    for (u32 g = 0; g < nE; ++g)
      eOff[g + 1] += eOff[g];
    {
      std::vector<u32> fp(eOff.begin(), eOff.end() - 1);
      for (u32 e = 0; e < nCells; ++e)
        for (u32 i = nc[e]; i < cellNdofs[e]; ++i)
          eSlots[fp[map[op->hMapOffset[e] + i] - nClassical]++] = enrOff[e] + (i - nc[e]);
    }

    // dof -> map slots (slot order = cell order)
    std::vector<u32> sOff(nDofs + 1, 0), sList(sumNe);
    for (u64 s = 0; s < sumNe; ++s)
      ++sOff[map[s] + 1];
    for (u32 d = 0; d < nDofs; ++d)
      sOff[d + 1] += sOff[d];
    {
      std::vector<u32> fp(sOff.begin(), sOff.end() - 1);
      for (u64 s = 0; s < sumNe; ++s)
        sList[fp[map[s]]++] = (u32)s;
    }

    hipStream_t st = ctx->stream;
    upload(op->map, map, sumNe, st);
    upload(op->meta, op->hMeta, st);
    upload(op->first, hFirst, st);
    upload(op->last, hLast, st);
    if (!hResRows.empty())
      upload(op->resRows, hResRows, st);
    upload(op->enrSlotOffsets, eOff, st);
    upload(op->enrSlots, eSlots, st);
    upload(op->dofSlotOffsets, sOff, st);
    upload(op->dofSlots, sList, st);
  }

  void uploadCellMatrices(dfx_op *op, const double *mats, bool device)
  {
    hipStream_t st = op->ctx->stream;
    DevBuf      raw, rawOff;
    const double *dmats = mats;
    if (!device)
      {
        upload(raw, mats, op->sumNe2, st);
        dmats = raw.as<double>();
      }
    upload(rawOff, op->hMatOffset, st);
    ++op->gen;
    if (!op->frag64.p)
      op->frag64.alloc(op->fragLen64 * sizeof(double));
    k::repackCells<double>(st, op, dmats, rawOff.as<u64>(), op->frag64.as<double>());
    if (op->hasF32)
      {
        if (!op->frag32.p)
          op->frag32.alloc(op->fragLen32 * sizeof(float));
        k::repackCells<float>(st, op, dmats, rawOff.as<u64>(), op->frag32.as<float>());
      }
    DFX_HIP(hipStreamSynchronize(st));
  }
} // namespace dfx
