// Serving: a stream of independent Kohn-Sham eigenproblems on one GPU with
// continuous batching.  Each admitted request runs on a slot (a host thread
// with its own dfx_ctx, hence its own HIP stream), so the kernels of all
// admitted requests overlap on the device; a request is admitted, in arrival
// order, as soon as it has arrived, a slot is free and its estimated device
// footprint fits in the memory budget.  A request keeps its mesh, cell
// matrices and subspace resident for the whole solve and releases everything
// when it finishes.  Every kernel is deterministic and a request touches only
// its own buffers, so its result does not depend on what runs beside it.
//
// There is no serving layer in the reference; a request is the composition
// of its setup (FEBasisManager / computeFEMatrices, here dfx_problem_build)
// and KohnShamEigenSolver::solve (src/ksdft/KohnShamEigenSolver.t.cpp:214-536).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <new>
#include <stdexcept>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "api_util.hpp"
#include "atom.hpp"

namespace dfx
{
  namespace
  {
    using clk = std::chrono::steady_clock;

    // cells whose bounding box meets each atom's enrichment support (the
    // builder's test), hence exact sum n_e and sum n_e^2
    void cellSizes(const dfx_problem_spec &s, u64 &sumN, u64 &sumN2, u32 &maxN, u64 &nCells)
    {
      const u32 nloc = (s.p + 1) * (s.p + 1) * (s.p + 1);
      nCells         = (u64)s.n[0] * s.n[1] * s.n[2];
      std::vector<u32> extra(nCells, 0);
      if (s.enrich && s.nAtoms)
        {
          for (u32 a = 0; a < s.nAtoms; ++a)
            {
              u32    nf;
              double rs;
              atomEnrichment(s, a, nf, rs);
              u32 lo[3], hi[3];
              for (int d = 0; d < 3; ++d)
                {
                  const double h  = (s.hi[d] - s.lo[d]) / s.n[d];
                  const double c  = s.atomXYZ[3 * a + d];
                  const double l  = std::floor((c - rs - s.lo[d]) / h);
                  const double u  = std::floor((c + rs - s.lo[d]) / h);
                  lo[d] = (u32)std::max(0.0, l);
                  hi[d] = (u32)std::min((double)s.n[d] - 1, std::max(0.0, u));
                }
              for (u32 k = lo[2]; k <= hi[2]; ++k)
                for (u32 j = lo[1]; j <= hi[1]; ++j)
                  for (u32 i = lo[0]; i <= hi[0]; ++i)
                    extra[((u64)k * s.n[1] + j) * s.n[0] + i] += nf;
            }
        }
      sumN = sumN2 = 0;
      maxN = 0;
      for (u64 e = 0; e < nCells; ++e)
        {
          const u64 n = nloc + extra[e];
          sumN += n;
          sumN2 += n * n;
          maxN = std::max(maxN, (u32)n);
        }
    }

    // device bytes a request holds at its peak: raw H, M, K (FP64), the three
    // operators' FP64 + FP32 fragment copies, setup tabulation batches (at
    // most 6 GB), and about 16 N x states FP64 vectors of solver work space
    u64 footprint(const dfx_request &r)
    {
      u64 sumN, sumN2, nCells;
      u32 maxN;
      cellSizes(r.spec, sumN, sumN2, maxN, nCells);
      const u32 p  = r.spec.p;
      const u64 nC = (u64)(r.spec.n[0] * p + 1) * (r.spec.n[1] * p + 1) * (r.spec.n[2] * p + 1);
      u64 nE = 0;
      for (u32 a = 0; r.spec.enrich && a < r.spec.nAtoms; ++a)
        {
          u32    nf;
          double rs;
          atomEnrichment(r.spec, a, nf, rs);
          nE += nf;
        }
      const u64 N = nC + nE;
      const u64 nq = r.spec.nqEnr ? r.spec.nqEnr : 2 * (p + 1);
      const u64 nQ = nq * nq * nq;
      const u64 mats  = 3 * 8 * sumN2 + 3 * 12 * (sumN2 + sumN2 / 4);
      const u64 setup = std::min<u64>(6ull << 30, nCells * nQ * maxN * 8 * 9);
      // N x n FP64 arrays: subspaces, filter and CholGS / RR work, and the
      // H X, M X the solver keeps of its eigenvectors
      const u64 vecs  = 18 * N * (u64)r.numStates * 8;
      return (mats + setup + vecs) * 5 / 4 + (64ull << 20);
    }

    struct Shared
    {
      std::mutex              mu;
      std::condition_variable cv;
      u32                     next = 0;   // next request (arrival order) to admit
      u64                     used = 0;   // admitted footprint
      int                     firstError = 0;
    };

    // one request on its slot's context: build, operators, seeded guess, KS
    // solve; everything the request allocated (its objects and the
    // context's named work buffers) is released at the end, the context
    // itself (stream, rocBLAS handle) serves the slot's next request
    int runRequest(dfx_ctx *ctx, int precision, bool useGraph, const dfx_request &r,
                   double *energies, dfx_request_result &out)
    {
      int              rc = 0;
      dfx_problem     *P = nullptr;
      dfx_op          *H = nullptr, *M = nullptr, *K = nullptr, *Mi = nullptr, *Ml = nullptr;
      dfx_constraints *cons = nullptr;
      dfx_mv          *G = nullptr, *L = nullptr, *W = nullptr;
      auto run = [&]() -> int {
        int s = dfx_problem_build(ctx, &r.spec, &P);
        if (s)
          return s;
        s = dfx_problem_create_operators(P, precision == DFX_PREC_MIXED ? DFX_OP_KEEP_FP32 : 0,
                                         &H, &M, &K, &Mi, &Ml, &cons);
        if (s)
          return s;
        uint64_t sz[8];
        dfx_problem_sizes(P, sz);
        const u32 N = (u32)sz[1];
        if ((s = dfx_mv_create(ctx, N, r.numStates, DFX_F64, &G)) ||
            (s = dfx_mv_create(ctx, N, 1, DFX_F64, &L)) ||
            (s = dfx_mv_create(ctx, N, r.numStates, DFX_F64, &W)))
          return s;
        if ((s = dfx_mv_fill_random(G, r.guessSeed, -1.0, 1.0)) ||
            (s = dfx_mv_fill_random(L, r.guessSeed + 1, -1.0, 1.0)) ||
            (s = dfx_constraints_distribute_parent_to_child(cons, G)) ||
            (s = dfx_constraints_distribute_parent_to_child(cons, L)))
          return s;
        dfx_ks_params kp{};
        kp.numWantedEigenvalues        = r.numStates;
        kp.maxChebyshevFilterPass      = r.maxPasses;
        kp.polynomialDegree            = 0;
        kp.eigenSolveResidualTolerance = r.tolerance;
        kp.waveFunctionBatchSize       = 0;
        kp.isResidualChebyshevFilter   = 1;
        kp.precision                   = precision;
        kp.lanczosMaxKrylov            = 20;
        kp.numElectrons                = r.numElectrons;
        kp.smearingTemperature         = 500.0;
        kp.fermiEnergyTolerance        = 1e-10;
        kp.fracOccupancyTolerance      = 1e-3;
        // a small request's degree loop is launch-bound: capture it once
        // and replay it on every pass
        kp.useGraph                    = useGraph ? 1 : 0;
        dfx_ks_result res{};
        s = dfx_ks_eigensolve(ctx, &kp, H, M, Mi, Ml, G, L, energies, W, nullptr, &res);
        out.passes = res.passes;
        return s;
      };
      rc = run();
      static const bool verbose = [] {
        const char *e = getenv("DFX_SERVE_VERBOSE");
        return e && atoi(e) != 0;
      }();
      if (rc < 0 && verbose)
        fprintf(stderr, "dfx_serve: request failed (%d): %s\n", rc, ctx->err.c_str());
      for (dfx_mv *v : {G, L, W})
        if (v)
          dfx_mv_destroy(v);
      for (dfx_op *o : {H, M, K, Mi, Ml})
        if (o)
          dfx_op_destroy(o);
      if (cons)
        dfx_constraints_destroy(cons);
      if (P)
        dfx_problem_destroy(P);
      (void)hipStreamSynchronize(ctx->stream);
      {
        CaptureLock lk(captureMutex());
        ctx->work.clear();
      }
      return rc;
    }

    // status of a request whose run threw (no exception crosses the ABI)
    int statusOf(std::exception_ptr e)
    {
      try
        {
          std::rethrow_exception(e);
        }
      catch (const Error &x)
        {
          return x.code;
        }
      catch (const std::bad_alloc &)
        {
          return DFX_E_OOM;
        }
      catch (const std::invalid_argument &)
        {
          return DFX_E_ARG;
        }
      catch (...)
        {
          return DFX_E_STATE;
        }
    }
  } // namespace
} // namespace dfx

extern "C"
{
  using namespace dfx;

  int dfx_request_footprint(const dfx_request *r, uint64_t *bytes)
  {
    if (!r || !bytes)
      return DFX_E_ARG;
    try
      {
        *bytes = footprint(*r);
        return 0;
      }
    catch (...)
      {
        return DFX_E_ARG;
      }
  }

  int dfx_serve(const dfx_serve_params *p, uint32_t nRequests, const dfx_request *reqs,
                double *energies, dfx_request_result *results)
  {
    if (!p || (nRequests && (!reqs || !energies || !results)) || !p->maxSlots)
      return DFX_E_ARG;
    std::vector<u32> order(nRequests);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) {
      return reqs[a].arrivalSeconds < reqs[b].arrivalSeconds;
    });
    std::vector<u64> energyOff(nRequests + 1, 0), fp(nRequests);
    for (u32 i = 0; i < nRequests; ++i)
      {
        energyOff[i + 1] = energyOff[i] + reqs[i].numStates;
        fp[i]            = footprint(reqs[i]);
        if (fp[i] > p->memoryBudgetBytes)
          return DFX_E_OOM; // could never be admitted
      }
    Shared     sh;
    const auto t0 = clk::now();
    auto since = [&](clk::time_point t) {
      return std::chrono::duration<double>(t - t0).count();
    };
    auto worker = [&](u32 slot) {
      // the slot's context: one stream and rocBLAS handle for all its
      // requests.  Slots already run side by side, so each uses one stream
      // (no second stream for size classes), keeping the process within the
      // device's hardware queues
      dfx_ctx  *ctx   = nullptr;
      const int ctxRc = dfx_ctx_create(p->device, &ctx);
      if (ctxRc == 0)
        ctx->forkClasses = false;
      for (;;)
        {
          u32 idx;
          {
            std::unique_lock<std::mutex> lk(sh.mu);
            for (;;)
              {
                if (sh.next >= nRequests)
                  {
                    lk.unlock();
                    if (ctx)
                      dfx_ctx_destroy(ctx);
                    return;
                  }
                idx = order[sh.next];
                const auto arrive =
                  t0 + std::chrono::duration_cast<clk::duration>(
                         std::chrono::duration<double>(reqs[idx].arrivalSeconds));
                if (clk::now() < arrive)
                  {
                    sh.cv.wait_until(lk, arrive);
                    continue;
                  }
                if (sh.used + fp[idx] <= p->memoryBudgetBytes)
                  break;
                sh.cv.wait(lk); // memory frees when another request finishes
              }
            ++sh.next;
            sh.used += fp[idx];
          }
          dfx_request_result &out = results[idx];
          out                     = dfx_request_result{};
          out.slot                = slot;
          out.footprintBytes      = fp[idx];
          out.admitSeconds        = since(clk::now());
          if (!ctx)
            out.status = ctxRc;
          else
            try
              {
                out.status = runRequest(ctx, p->precision, p->useGraph != 0, reqs[idx],
                                        energies + energyOff[idx], out);
              }
            catch (...)
              {
                out.status = statusOf(std::current_exception());
              }
          out.doneSeconds = since(clk::now());
          {
            std::lock_guard<std::mutex> lk(sh.mu);
            sh.used -= fp[idx];
            if (out.status < 0 && !sh.firstError)
              sh.firstError = out.status;
          }
          sh.cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    for (u32 s = 0; s < p->maxSlots; ++s)
      th.emplace_back(worker, s);
    for (auto &t : th)
      t.join();
    return sh.firstError;
  }
}
