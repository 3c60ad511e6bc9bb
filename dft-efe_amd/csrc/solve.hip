// C ABI of libdfx, part 2: the eigensolver and linear-solver drivers on one
// GPU.  All vector work is on the device; only per-column scalars (dots,
// eigenvalues, recurrence coefficients) cross to the host, as in the
// reference's host-side control flow.
#include <hipblaslt/hipblaslt.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>

#include "api_util.hpp"
#include "kernels.hpp"

using namespace dfx;

struct dfx_chfsi
{
  dfx_ctx         *ctx = nullptr;
  dfx_chfsi_params p{};
  dfx_mv          *guess = nullptr;
  double           ms[3] = {0, 0, 0};
  // storeIntermediateSubspaces: filtered and orthonormalized-filtered
  // subspaces of the last solve (N x n FP64)
  dfx::DevBuf      filtered, filteredOrtho;
  dfx::u32         storedN = 0, storedB = 0;
  // H X and M X of the last solve's eigenvectors X (FP64 N x n), formed by
  // rotating the products Cholesky-Gram-Schmidt (M X_f) and Rayleigh-Ritz
  // (H X_o) compute anyway; valid for the operators (uid, generation) and
  // the content hash of X they were made with
  dfx::DevBuf      kpHX, kpMX;
  bool             kpValid = false;
  uint64_t         kpHash = 0, kpA = 0, kpAGen = 0, kpB = 0, kpBGen = 0;
  dfx::u32         kpN = 0, kpn = 0;
};

namespace dfx
{
  namespace
  {
    template <class T>
    View<T> sub(View<T> v, u32 s, u32 nb)
    {
      return View<T>{v.p + s, v.N, nb, v.ld};
    }
    template <class T>
    View<const T> cv(View<T> v)
    {
      return View<const T>{v.p, v.N, v.B, v.ld};
    }

    template <class T>
    View<T> workView(dfx_ctx *ctx, const char *name, u32 N, u32 B)
    {
      DevBuf &b = work(ctx, name, (size_t)N * B * sizeof(T) + 256);
      return View<T>{b.as<T>(), N, B, B};
    }

    double *workDoubles(dfx_ctx *ctx, const char *name, size_t n)
    {
      return work(ctx, name, n * sizeof(double) + 64).as<double>();
    }

    // small host -> device upload; completes before returning, so the host
    // buffer (often a temporary) may be released right after
    // pinned host doubles (asynchronous device-to-host copies)
    struct PinnedDoubles
    {
      double *p = nullptr;
      explicit PinnedDoubles(size_t n)
      {
        DFX_HIP(hipHostMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(n, 1) * sizeof(double),
                              hipHostMallocDefault));
      }
      ~PinnedDoubles()
      {
        if (p)
          (void)hipHostFree(p);
      }
      PinnedDoubles(const PinnedDoubles &)            = delete;
      PinnedDoubles &operator=(const PinnedDoubles &) = delete;
    };

    // two events without timing, for waits on one iteration's results
    struct EventPair
    {
      hipEvent_t e[2] = {nullptr, nullptr};
      EventPair()
      {
        for (auto &x : e)
          DFX_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
      }
      ~EventPair()
      {
        for (auto &x : e)
          if (x)
            (void)hipEventDestroy(x);
      }
      EventPair(const EventPair &)            = delete;
      EventPair &operator=(const EventPair &) = delete;
    };

    void toDevice(dfx_ctx *ctx, double *d, const double *h, size_t n)
    {
      if (!n)
        return;
      DFX_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      DFX_HIP(hipStreamSynchronize(ctx->stream));
    }

    struct Timer
    {
      hipEvent_t a, b;
      hipStream_t s;
      explicit Timer(hipStream_t st) : s(st)
      {
        DFX_HIP(hipEventCreate(&a));
        DFX_HIP(hipEventCreate(&b));
        DFX_HIP(hipEventRecord(a, s));
      }
      double stop()
      {
        DFX_HIP(hipEventRecord(b, s));
        DFX_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        DFX_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
      }
      ~Timer()
      {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    };

    // M^-1 of a diagonal (+ enriched block) operator applied in the fused
    // plain-filter step: only the enriched rows need the block GEMM; the
    // diagonal rows are scaled inside chebStep.
    template <class T>
    void enrichedPart(dfx_op *BInv, View<const T> x, View<T> tE)
    {
      if (BInv->nEnriched)
        k::enrichedBlock<T>(BInv->ctx->stream, tE, BInv->enrBlock.as<double>(), x,
                            BInv->nDofs - BInv->nEnriched, BInv->nEnriched);
    }

    // ------------------------------------------------------------ filters
    // ChebyshevFilter (src/linearAlgebra/ChebyshevFilter.t.cpp:39-135) on one
    // block, bulk in T.  Xin (FP64) receives the filtered block as well (the
    // reference assigns eigenSubspaceGuess = filteredSubspace at the end).
    template <class T>
    void plainFilter(dfx_ctx *ctx, dfx_op *A, dfx_op *BInv, View<double> Xin, u32 m,
                     double a, double b, double c, View<double> Xout)
    {
      hipStream_t s  = ctx->stream;
      const u32   N  = Xin.N, nb = Xin.B;
      DFX_REQUIRE(BInv->kind == OP_DIAGONAL,
                  "the fused filter needs a diagonal (+ enriched block) BInv");
      const double e  = 0.5 * (c - b);
      const double c0 = 0.5 * (c + b);
      double       sigma  = e / (a - c0);
      const double sigma1 = sigma;
      const double gamma  = 2.0 / sigma1;
      View<T> xk = workView<T>(ctx, "pf.xk", N, nb);
      View<T> yk = workView<T>(ctx, "pf.yk", N, nb);
      View<T> s1 = workView<T>(ctx, "pf.s1", N, nb);
      View<T> tE = workView<T>(ctx, "pf.tE", N, nb);
      k::convert<T, double>(s, xk, cv(Xin));
      const double *dinv = BInv->diag.as<double>();
      const u32     nC   = BInv->nDofs - BInv->nEnriched;
      // Y1 = (sigma1/e)(BInv A X - c0 X)
      applyOp<T>(A, xk, s1);
      enrichedPart<T>(BInv, cv(s1), tE);
      // yk = alpha * Minv s1 + beta * xk + 0 * yk  (yk zeroed first)
      k::fill<T>(s, yk, 0.0);
      k::chebStep<T>(s, yk, cv(s1), dinv, cv(tE), nC, cv(xk), sigma1 / e,
                     -sigma1 / e * c0, 0.0);
      for (u32 deg = 2; deg <= m; ++deg)
        {
          const double sigma2 = 1.0 / (gamma - sigma);
          applyOp<T>(A, yk, s1);
          enrichedPart<T>(BInv, cv(s1), tE);
          // xk <- (2 sigma2/e)(Minv s1 - c0 yk) - sigma sigma2 xk
          k::chebStep<T>(s, xk, cv(s1), dinv, cv(tE), nC, cv(yk), 2.0 * sigma2 / e,
                         -2.0 * sigma2 / e * c0, -sigma * sigma2);
          std::swap(xk, yk);
          sigma = sigma2;
        }
      k::convert<double, T>(s, Xout, cv(yk));
      k::convert<double, T>(s, Xin, cv(yk));
    }

    // ResidualChebyshevFilterGEP (src/linearAlgebra/ChebyshevFilter.t.cpp:
    // 242-444) on one block.  The residual seed R0 = A X - lambda B X and the
    // final X_m = BInv R_m + mu_m X are FP64; the degree loop runs in T.
    // Runs body(deg) for deg = first..last, either directly or as one
    // replay of a hipGraph captured from the same calls (cached in the
    // context under `name` until `key` changes).  body must launch kernels
    // only (no allocation, host copy or synchronization once first - 1 has
    // run eagerly) and must leave the same host state after the loop as the
    // captured run did: `replayHost` re-creates that state when the cached
    // graph is replayed without running body.
    template <class Body, class Replay>
    void degreeLoop(dfx_ctx *ctx, bool useGraph, const std::string &name,
                    const std::vector<uint64_t> &key, const std::vector<uint64_t> &uids,
                    u32 first, u32 last, Body body, Replay replayHost)
    {
      if (!useGraph || first > last)
        {
          for (u32 d = first; d <= last; ++d)
            body(d);
          return;
        }
      hipStream_t         s = ctx->stream;
      dfx_ctx::Graph     &g = ctx->graphs[name];
      if (g.exec && g.key == key)
        {
          replayHost();
          DFX_HIP(hipGraphLaunch(g.exec, s));
          ++ctx->graphReplays;
          return;
        }
      if (g.exec)
        {
          DFX_HIP(hipGraphExecDestroy(g.exec));
          g.exec = nullptr;
        }
      hipGraph_t  graph = nullptr;
      CaptureLock lk(captureMutex());
      DFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      try
        {
          for (u32 d = first; d <= last; ++d)
            body(d);
        }
      catch (...)
        {
          (void)hipStreamEndCapture(s, &graph);
          if (graph)
            (void)hipGraphDestroy(graph);
          throw;
        }
      DFX_HIP(hipStreamEndCapture(s, &graph));
      hipGraphExec_t exec = nullptr;
      const hipError_t e  = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      DFX_HIP(e);
      g.exec = exec;
      g.key  = key;
      g.uids = uids;
      ++ctx->graphCaptures;
      DFX_HIP(hipGraphLaunch(g.exec, s));
    }

    // DFX_FUSE_RES=0: the residual degree step as a launch of its own after
    // each apply (read on every filter call)
    bool fuseResidualStep()
    {
      const char *e = getenv("DFX_FUSE_RES");
      return !e || atoi(e) != 0;
    }

    // AXin / BXin (optional, leading dimension ld): A Xin and B Xin already
    // formed (the products kept from the solve that made Xin); the seed then
    // takes them instead of two applies
    template <class T>
    void residualFilter(dfx_ctx *ctx, dfx_op *A, dfx_op *Bop, dfx_op *BInv,
                        const double *lam, View<double> Xin, u32 m, double a,
                        double b, double c, View<double> Y, bool useGraph = false,
                        const double *AXin = nullptr, const double *BXin = nullptr,
                        u32 ld = 0)
    {
      hipStream_t s  = ctx->stream;
      const u32   N  = Xin.N, nb = Xin.B;
      const double e  = 0.5 * (c - b);
      const double c0 = 0.5 * (c + b);
      double       sigma  = e / (a - c0);
      const double sigma1 = sigma;
      const double gamma  = 2.0 / sigma1;

      // ---- scalar recurrence (eigenValuesFiltered1/2), host, reference order
      std::vector<double> lamF(lam, lam + nb), l1(nb, 1.0), l2(nb), ones(nb, 1.0);
      double alpha1 = sigma1 / e, alpha2 = -c0;
      for (u32 j = 0; j < nb; ++j)
        l2[j] = alpha1 * alpha2;
      for (u32 j = 0; j < nb; ++j)
        l2[j] = 1.0 * ones[j] * l2[j] + alpha1 * lamF[j] * l1[j];
      const double rnScale = alpha1;
      // per-column mu rows 0..m+1, ones, lambda, then {alpha1, alpha2, c0}
      // per degree: the degree loop reads every bound-dependent scalar from
      // this device table, so a captured loop is replayed for new bounds
      const size_t coefOff = (size_t)(m + 2) * nb + 2 * nb;
      std::vector<double> table(coefOff + 3 * (size_t)(m + 1), 0.0);
      {
        double sg = sigma;
        for (u32 deg = 2; deg <= m; ++deg)
          {
            const double sigma2 = 1.0 / (gamma - sg);
            const double al1    = 2.0 * sigma2 / e, al2 = -(sg * sigma2);
            table[coefOff + 3 * deg]     = al1;
            table[coefOff + 3 * deg + 1] = al2;
            table[coefOff + 3 * deg + 2] = c0;
            std::copy(l2.begin(), l2.end(), table.begin() + (size_t)deg * nb);
            for (u32 j = 0; j < nb; ++j)
              l1[j] = (-c0 * al1) * l2[j] + al2 * l1[j];
            for (u32 j = 0; j < nb; ++j)
              l1[j] = 1.0 * ones[j] * l1[j] + al1 * lamF[j] * l2[j];
            std::swap(l1, l2);
            sg = sigma2;
          }
      }
      // final mu (l2) in row m+1; ones and lambda after it
      std::copy(l2.begin(), l2.end(), table.begin() + (size_t)(m + 1) * nb);
      std::copy(ones.begin(), ones.end(), table.begin() + (size_t)(m + 2) * nb);
      std::copy(lamF.begin(), lamF.end(), table.begin() + (size_t)(m + 2) * nb + nb);
      double *dtab = workDoubles(ctx, "rf.table", table.size());
      toDevice(ctx, dtab, table.data(), table.size());
      const double *dOnes = dtab + (size_t)(m + 2) * nb;
      const double *dLam  = dOnes + nb;

      View<T> r0 = workView<T>(ctx, "rf.r0", N, nb);
      View<T> rn = workView<T>(ctx, "rf.rn", N, nb);
      View<T> rs = workView<T>(ctx, "rf.res", N, nb);
      View<T> s1 = workView<T>(ctx, "rf.s1", N, nb);
      View<T> s2 = workView<T>(ctx, "rf.s2", N, nb);
      // ---- FP64 seed: Y = B X, S3 = A X, Y = S3 - lambda o Y; then
      // ResidualNew = alpha1 * Y, Residual = 0 (one fused launch).  From
      // kept products the seed is formed in registers (the same bits)
      if (AXin && BXin)
        k::resSeedFrom<T>(s, r0, rn, rs, View<const double>{AXin, N, nb, ld},
                          View<const double>{BXin, N, nb, ld}, dOnes, dLam, rnScale);
      else
        {
          View<double> s3 = workView<double>(ctx, "rf.s3", N, nb);
          applyOp<double>(Bop, Xin, Y);
          applyOp<double>(A, Xin, s3);
          k::axpbyBlocked<double>(s, Y, 1.0, dOnes, cv(s3), -1.0, dLam, cv(Y));
          k::resSeed<T>(s, r0, rn, rs, cv(Y), rnScale);
        }
      // the degree step fused into the folded cell apply where it qualifies
      // (the same bits as the separate step).  The fused step leaves the
      // constrained rows out, which needs them zero in every buffer from the
      // seed on: A and B zero them when B shares A's constraints
      const bool zeroRows = !A->cons || Bop->cons == A->cons;
      const bool fuse = fuseResidualStep() && zeroRows && rs.ld == rn.ld && r0.ld == rn.ld;
      auto step = [&](u32 deg) {
        const k::ResFuse<T> rf{rs.p, r0.p, dtab + (size_t)deg * nb,
                               dtab + coefOff + 3 * (size_t)deg};
        if (!(fuse && applyInvThenCellwise<T>(BInv, A, rn, s1, s2, &rf)))
          {
            if (!applyInvThenCellwise<T>(BInv, A, rn, s1, s2))
              {
                applyOp<T>(BInv, rn, s1);
                applyOp<T>(A, s1, s2);
              }
            k::resStep<T>(s, rs, cv(s2), cv(rn), cv(r0), dtab + (size_t)deg * nb,
                          dtab + coefOff + 3 * (size_t)deg);
          }
        std::swap(rn, rs);
      };
      // degree 2 runs directly (first-use allocations and kernel attributes);
      // degrees 3..m go through one captured graph when asked for and when no
      // operator is being profiled (profiling records host-side per launch)
      if (m >= 2)
        step(2);
      const bool graph = useGraph && !A->profiling && !BInv->profiling;
      // the captured launches depend on operators, buffers and shapes only
      // (coefficients and mu come from dtab), so one graph per block width
      // serves every pass of a solve; an operator whose buffers were
      // reallocated since (its generation) is captured again
      const std::vector<uint64_t> key = {A->uid, A->gen, BInv->uid, BInv->gen,
                                         (uint64_t)(uintptr_t)rn.p, (uint64_t)(uintptr_t)rs.p,
                                         (uint64_t)(uintptr_t)s1.p, (uint64_t)(uintptr_t)s2.p,
                                         (uint64_t)(uintptr_t)r0.p, (uint64_t)(uintptr_t)dtab,
                                         N, nb, m, fuse ? 1u : 0u};
      const std::string gname = std::string(sizeof(T) == 8 ? "rf.f64." : "rf.f32.") +
                                std::to_string(nb);
      degreeLoop(ctx, graph, gname, key, {A->uid, BInv->uid}, 3, m, step, [&] {
        for (u32 d = 3; d <= m; ++d)
          std::swap(rn, rs);
      });
      // X_m = M^-1 R_m + Lambda X: with the step fused and a diagonal M^-1
      // under the same (homogeneous Dirichlet, classical) constraints, R_m is
      // zero on the constrained rows and the diagonal product is formed in
      // the result's launch (the same bits as the apply followed by resFinal)
      const dfx_constraints *ci = BInv->cons;
      const bool finalFused =
        fuse && BInv->kind == OP_DIAGONAL && ci == A->cons &&
        (!ci || (ci->homDirichlet &&
                 (ci->hRows.empty() || ci->hRows.back() < BInv->nDofs - BInv->nEnriched)));
      if (finalFused)
        {
          const u32 nC = BInv->nDofs - BInv->nEnriched;
          if (BInv->nEnriched)
            k::enrichedBlock<T>(s, rs, BInv->enrBlock.as<double>(), cv(rn), nC,
                                BInv->nEnriched);
          k::resFinalDiag<T>(s, Y, 1.0, dOnes, cv(rn), BInv->diag.as<double>(), nC, cv(rs), 1.0,
                             dtab + (size_t)(m + 1) * nb, cv(Xin));
        }
      else
        {
          applyOp<T>(BInv, rn, rs);
          k::resFinal<T>(s, Y, 1.0, dOnes, cv(rs), 1.0, dtab + (size_t)(m + 1) * nb, cv(Xin));
        }
    }

    // ------------------------------------------------ subspace rotations
    // Xo = X Q (N x n row-major, Q n x n column-major FP64) as one hipBLASLt
    // matmul in column-major terms, Xo^T = Q^T X^T, with the first algorithm
    // of the library's heuristic for the shape (a fixed choice, so results do
    // not depend on timing); C4 (N = 2.1 M, n = 200): 3.2 ms against 3.7 for
    // the rocBLAS DGEMM, which remains the fallback when hipBLASLt offers no
    // algorithm
    struct LtState
    {
      hipblasLtHandle_t h = nullptr;
      DevBuf            ws;
      struct Plan
      {
        u32                     n, N, ldx, ldo;
        bool                    ok = false;
        hipblasLtMatmulDesc_t   desc = nullptr;
        hipblasLtMatrixLayout_t a = nullptr, b = nullptr, c = nullptr;
        hipblasLtMatmulAlgo_t   algo{};
      };
      std::vector<Plan> plans;
      ~LtState()
      {
        for (Plan &p : plans)
          {
            if (p.a)
              (void)hipblasLtMatrixLayoutDestroy(p.a);
            if (p.b)
              (void)hipblasLtMatrixLayoutDestroy(p.b);
            if (p.c)
              (void)hipblasLtMatrixLayoutDestroy(p.c);
            if (p.desc)
              (void)hipblasLtMatmulDescDestroy(p.desc);
          }
        if (h)
          (void)hipblasLtDestroy(h);
      }
    };
    constexpr size_t LT_WORKSPACE = (size_t)64 << 20;

    void ltFree(void *p)
    {
      delete static_cast<LtState *>(p);
    }

    void rotate64(dfx_ctx *ctx, View<const double> X, const double *Q, View<double> Xo,
                  double *qr)
    {
      const u32 n = X.B, N = X.N;
      DFX_REQUIRE(Xo.B == n && Xo.N == N, "rotate: shape mismatch");
      // DFX_ROTATE_LT=0: the rocBLAS DGEMM (read on every call)
      const char *env = getenv("DFX_ROTATE_LT");
      if (env && atoi(env) == 0)
        {
          k::rotate<double>(ctx->stream, X, Q, Xo, qr, ctx->blas);
          return;
        }
      CaptureLock lk(captureMutex());
      if (!ctx->lt)
        {
          auto st = new LtState();
          if (hipblasLtCreate(&st->h) != HIPBLAS_STATUS_SUCCESS)
            st->h = nullptr;
          ctx->lt     = st;
          ctx->ltFree = ltFree;
        }
      LtState *st = static_cast<LtState *>(ctx->lt);
      LtState::Plan *plan = nullptr;
      for (auto &p : st->plans)
        if (p.n == n && p.N == N && p.ldx == X.ld && p.ldo == Xo.ld)
          plan = &p;
      if (!plan && st->h)
        {
          st->plans.push_back({n, N, X.ld, Xo.ld});
          plan = &st->plans.back();
          bool ok = hipblasLtMatmulDescCreate(&plan->desc, HIPBLAS_COMPUTE_64F, HIP_R_64F) ==
                    HIPBLAS_STATUS_SUCCESS;
          const hipblasOperation_t tA = HIPBLAS_OP_T, tB = HIPBLAS_OP_N;
          ok = ok && hipblasLtMatmulDescSetAttribute(plan->desc, HIPBLASLT_MATMUL_DESC_TRANSA,
                                                     &tA, sizeof(tA)) == HIPBLAS_STATUS_SUCCESS;
          ok = ok && hipblasLtMatmulDescSetAttribute(plan->desc, HIPBLASLT_MATMUL_DESC_TRANSB,
                                                     &tB, sizeof(tB)) == HIPBLAS_STATUS_SUCCESS;
          ok = ok && hipblasLtMatrixLayoutCreate(&plan->a, HIP_R_64F, n, n, n) ==
                       HIPBLAS_STATUS_SUCCESS;
          ok = ok && hipblasLtMatrixLayoutCreate(&plan->b, HIP_R_64F, n, N, X.ld) ==
                       HIPBLAS_STATUS_SUCCESS;
          ok = ok && hipblasLtMatrixLayoutCreate(&plan->c, HIP_R_64F, n, N, Xo.ld) ==
                       HIPBLAS_STATUS_SUCCESS;
          if (ok)
            {
              hipblasLtMatmulPreference_t pref = nullptr;
              size_t                      wsb  = LT_WORKSPACE;
              int                         got  = 0;
              hipblasLtMatmulHeuristicResult_t res{};
              ok = hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS &&
                   hipblasLtMatmulPreferenceSetAttribute(
                     pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsb, sizeof(wsb)) ==
                     HIPBLAS_STATUS_SUCCESS &&
                   hipblasLtMatmulAlgoGetHeuristic(st->h, plan->desc, plan->a, plan->b,
                                                   plan->c, plan->c, pref, 1, &res, &got) ==
                     HIPBLAS_STATUS_SUCCESS &&
                   got > 0;
              if (pref)
                (void)hipblasLtMatmulPreferenceDestroy(pref);
              if (ok)
                plan->algo = res.algo;
            }
          plan->ok = ok;
          if (ok && st->ws.bytes < LT_WORKSPACE)
            st->ws.alloc(LT_WORKSPACE);
        }
      if (plan && plan->ok)
        {
          const double one = 1.0, zero = 0.0;
          if (hipblasLtMatmul(st->h, plan->desc, &one, Q, plan->a, X.p, plan->b, &zero, Xo.p,
                              plan->c, Xo.p, plan->c, &plan->algo, st->ws.p, LT_WORKSPACE,
                              ctx->stream) == HIPBLAS_STATUS_SUCCESS)
            return;
          throw Error(DFX_E_HIP, "hipblasLtMatmul (rotation) failed");
        }
      k::rotate<double>(ctx->stream, X, Q, Xo, qr, ctx->blas);
    }

    // --------------------------------------------------- CholGS and RR
    int cholGS(dfx_ctx *ctx, View<double> X, View<double> Xo, dfx_op *Bop)
    {
      hipStream_t s = ctx->stream;
      const u32   N = X.N, n = X.B;
      if (N < n)
        return DFX_ORTHO_NON_ORTHONORMALIZABLE_MULTIVECTOR;
      View<double> MX = workView<double>(ctx, "cg.mx", N, n);
      applyOp<double>(Bop, X, MX);
      double *S    = workDoubles(ctx, "cg.S", (size_t)n * n);
      double *part = workDoubles(ctx, "gram.part", k::gramPartialsNeeded(N, n));
      k::gram<double>(s, cv(X), cv(MX), part, k::gramPartialsNeeded(N, n), S);
      int *info = work(ctx, "info", 64).as<int>();
      rocblas_status st;
      {
        CaptureLock lk(captureMutex());
        st = rocsolver_dpotrf(ctx->blas, rocblas_fill_lower, (rocblas_int)n, S, (rocblas_int)n,
                           info);
      }
      if (st != rocblas_status_success)
        throw Error(DFX_E_HIP, "rocsolver_dpotrf failed");
      int hinfo = 0;
      DFX_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
      std::vector<double> hS((size_t)n * n);
      DFX_HIP(hipMemcpyAsync(hS.data(), S, hS.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      DFX_HIP(hipStreamSynchronize(s));
      if (hinfo != 0)
        return DFX_ORTHO_LAPACK_ERROR;
      // |L_ii| < 1e-14 is reported as failure (OrthonormalizationFunctions.t.cpp:274-306)
      for (u32 i = 0; i < n; ++i)
        if (!(std::fabs(hS[i + (size_t)n * i]) >= 1e-14))
          return DFX_ORTHO_LAPACK_ERROR;
      for (u32 j = 0; j < n; ++j)
        for (u32 i = 0; i < j; ++i)
          hS[i + (size_t)n * j] = 0.0;
      toDevice(ctx, S, hS.data(), hS.size());
      rocblas_status st2;
      {
        CaptureLock lk(captureMutex());
        st2 = rocsolver_dtrtri(ctx->blas, rocblas_fill_lower, rocblas_diagonal_non_unit,
                           (rocblas_int)n, S, (rocblas_int)n, info);
      }
      if (st2 != rocblas_status_success)
        throw Error(DFX_E_HIP, "rocsolver_dtrtri failed");
      DFX_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
      // Xo = X L^-T: Q = (L^-1)^T
      double *Q = workDoubles(ctx, "cg.Q", (size_t)n * n);
      DFX_HIP(hipMemcpyAsync(hS.data(), S, hS.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      DFX_HIP(hipStreamSynchronize(s));
      if (hinfo != 0)
        return DFX_ORTHO_LAPACK_ERROR;
      std::vector<double> hQ((size_t)n * n, 0.0);
      for (u32 j = 0; j < n; ++j)
        for (u32 i = j; i < n; ++i)
          hQ[j + (size_t)n * i] = hS[i + (size_t)n * j];
      toDevice(ctx, Q, hQ.data(), hQ.size());
      double *qr = workDoubles(ctx, "rot.qr", (size_t)n * 16 * ((n + 15) / 16));
      rotate64(ctx, cv(X), Q, Xo, qr);
      return 0;
    }

    // hxName: the work buffer that receives H X (kept past the call when it
    // is not CholGS's "cg.mx")
    int rayleighRitz(dfx_ctx *ctx, dfx_op *A, View<double> X, double *eig,
                     View<double> Xo, const char *hxName = "cg.mx")
    {
      hipStream_t s = ctx->stream;
      const u32   N = X.N, n = X.B;
      View<double> HX = workView<double>(ctx, hxName, N, n);
      applyOp<double>(A, X, HX);
      double *S    = workDoubles(ctx, "rr.S", (size_t)n * n);
      double *part = workDoubles(ctx, "gram.part", k::gramPartialsNeeded(N, n));
      k::gram<double>(s, cv(X), cv(HX), part, k::gramPartialsNeeded(N, n), S);
      double *D    = workDoubles(ctx, "rr.D", 2 * (size_t)n);
      int    *info = work(ctx, "info", 64).as<int>();
      rocblas_status st;
      {
        CaptureLock lk(captureMutex());
        st = rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_lower,
                           (rocblas_int)n, S, (rocblas_int)n, D, D + n, info);
      }
      if (st != rocblas_status_success)
        throw Error(DFX_E_HIP, "rocsolver_dsyevd failed");
      int hinfo = 0;
      DFX_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
      DFX_HIP(hipMemcpyAsync(eig, D, n * sizeof(double), hipMemcpyDeviceToHost, s));
      DFX_HIP(hipStreamSynchronize(s));
      if (hinfo != 0)
        return DFX_EIG_LAPACK_ERROR;
      double *qr = workDoubles(ctx, "rot.qr", (size_t)n * 16 * ((n + 15) / 16));
      rotate64(ctx, cv(X), S, Xo, qr);
      return 0;
    }

    View<double> mvView(dfx_mv *m)
    {
      DFX_REQUIRE(m->dtype == DFX_F64, "FP64 multivector expected");
      return View<double>{static_cast<double *>(m->d), m->N, m->B, m->B};
    }

    // content hash of an N x n FP64 view (contiguous)
    u64 hashOf(dfx_ctx *ctx, View<const double> X)
    {
      DFX_REQUIRE(X.ld == X.B, "hash of a contiguous multivector");
      u64 *part = work(ctx, "hash.part", k::HASH_BLOCKS * sizeof(u64)).as<u64>();
      return k::hashWords(ctx->stream, X.p, (u64)X.N * X.B, part);
    }

    // A X and B X of a solve's eigenvectors can be kept when both operators
    // are linear in X: constraints (if any) without inhomogeneities
    bool productsUsable(const dfx_op *A, const dfx_op *Bop)
    {
      return (!A->cons || A->cons->homogeneous) && (!Bop->cons || Bop->cons->homogeneous);
    }

    bool productsFor(const dfx_chfsi *cs, const dfx_op *A, const dfx_op *Bop, u32 N, u32 n)
    {
      return cs->kpValid && cs->kpA == A->uid && cs->kpAGen == A->gen && cs->kpB == Bop->uid &&
             cs->kpBGen == Bop->gen && cs->kpN == N && cs->kpn == n;
    }

    // C = A B, n x n column-major, device
    void gemmSquare(dfx_ctx *ctx, const double *A, const double *B, double *C, u32 n)
    {
      const double one = 1.0, zero = 0.0;
      CaptureLock  lk(captureMutex());
      rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
      if (rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none,
                        (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, &one, A, (rocblas_int)n,
                        B, (rocblas_int)n, &zero, C, (rocblas_int)n) != rocblas_status_success)
        throw Error(DFX_E_HIP, "rocblas_dgemm failed");
    }

    // res_j = || -A x_j + lambda_j B x_j || (host) from AX and BX (N x n,
    // contiguous), without forming the residual vectors: the bits of the
    // blocked axpby followed by the column norms
    void residualNormsOf(dfx_ctx *ctx, View<const double> AX, View<const double> BX,
                         const double *eig, double *res)
    {
      const u32 n = AX.B;
      std::vector<double> h(2 * (size_t)n);
      for (u32 j = 0; j < n; ++j)
        {
          h[j]     = -1.0;
          h[n + j] = eig[j];
        }
      double *d = workDoubles(ctx, "rn.coef", 2 * (size_t)n);
      toDevice(ctx, d, h.data(), h.size());
      ctx->reducePartials.ensure((size_t)1024 * n * sizeof(double) + n * sizeof(double));
      double *part = ctx->reducePartials.as<double>();
      double *out  = part + (size_t)1024 * n;
      k::combinationNorms(ctx->stream, 1.0, d, AX, 1.0, d + n, BX, part, out);
      DFX_HIP(hipMemcpyAsync(res, out, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      DFX_HIP(hipStreamSynchronize(ctx->stream));
      for (u32 j = 0; j < n; ++j)
        res[j] = std::sqrt(res[j]);
    }

    // res_j = || A x_j - lambda_j B x_j || from the kept products
    void productResidualNorms(dfx_chfsi *cs, const double *eig, double *res)
    {
      const u32 N = cs->kpN, n = cs->kpn;
      residualNormsOf(cs->ctx, View<const double>{cs->kpHX.as<double>(), N, n, n},
                      View<const double>{cs->kpMX.as<double>(), N, n, n}, eig, res);
    }

    // Column counts of the filter's launches for n eigenvectors in batches
    // of bs: batches of bs, a remainder of at most bs / 4 joining the last
    // full batch, then consecutive batches grouped while a group holds at
    // most DFX_FILTER_GROUP_COLS columns (default 128, eight 16-column tiles;
    // 0 keeps the batches).  Per-column results do not depend on the
    // grouping (the same bits); the per-cell work of the cell kernel (its
    // metadata, map and masks) is shared by more columns.  C4 (200 states,
    // batches of 64): launches of 128 + 72 columns, filter 143.4 -> 134.6 ms
    // per mixed pass against 64 + 64 + 72 (one launch of 200: 138.2 ms, its
    // 8-column tail chunk streams every cell matrix once more)
    std::vector<u32> filterLaunchWidths(u32 n, u32 bs)
    {
      const char *e         = getenv("DFX_FILTER_GROUP_COLS"); // read per solve
      const u32   groupCols = e ? (u32)atol(e) : 128u;
      std::vector<u32> batches, out;
      for (u32 s0 = 0, nb = 0; s0 < n; s0 += nb)
        {
          nb = std::min(bs, n - s0);
          if (n - s0 > nb && (n - s0 - nb) * 4 <= bs)
            nb = n - s0;
          batches.push_back(nb);
        }
      u32 acc = 0;
      for (u32 w : batches)
        {
          if (acc && acc + w > groupCols)
            {
              out.push_back(acc);
              acc = 0;
            }
          acc += w;
        }
      if (acc)
        out.push_back(acc);
      return out;
    }

    // ChebyshevFilteredEigenSolver::solve (ChebyshevFilteredEigenSolver.t.cpp:194-437).
    // Unless p.noProductReuse, the solve keeps A X and B X of its eigenvectors
    // (two rotations of products it forms anyway); a later solve whose guess
    // has the same content (hash) takes them for the residual filter's seed
    // A X - B X Lambda instead of two applies, and the residual norms of X
    // take them instead of two more.
    int chfsiSolve(dfx_chfsi *cs, dfx_op *A, double *eig, dfx_mv *X, dfx_op *Bop,
                   dfx_op *BInv)
    {
      dfx_ctx *ctx = cs->ctx;
      const auto &p = cs->p;
      View<double> G  = mvView(cs->guess);
      View<double> XV = mvView(X);
      DFX_REQUIRE(G.N == XV.N && G.B == XV.B, "guess and eigenvectors differ in shape");
      DFX_REQUIRE(A->nDofs == XV.N, "operator size does not match the eigenvectors");
      const u32 n  = XV.B;
      const bool keep  = !p.noProductReuse && productsUsable(A, Bop);
      const bool reuse = keep && p.isResidualChebyshevFilter && productsFor(cs, A, Bop, G.N, n) &&
                         hashOf(ctx, cv(G)) == cs->kpHash;
      cs->kpValid = false;
      const u32 bs = (p.eigenVectorBatchSize == 0 || p.eigenVectorBatchSize > n)
                       ? n
                       : p.eigenVectorBatchSize;
      const bool mixed = p.precision == DFX_PREC_MIXED;
      Timer tf(ctx->stream);
      // every column runs the same recurrence, so each column's result is the
      // same whatever batch it is filtered in: the batches are launched in
      // groups (filterLaunchWidths), and a narrow batch would stream all cell
      // matrices for a few columns
      const std::vector<u32> widths = filterLaunchWidths(n, bs);
      for (u32 s0 = 0, bi = 0, nb = 0; s0 < n; s0 += nb, ++bi)
        {
          nb = widths[bi];
          View<double> Gin = sub(G, s0, nb), Xo = sub(XV, s0, nb);
          if (p.isResidualChebyshevFilter)
            {
              const double *ax = reuse ? cs->kpHX.as<double>() + s0 : nullptr;
              const double *bx = reuse ? cs->kpMX.as<double>() + s0 : nullptr;
              if (mixed)
                residualFilter<float>(ctx, A, Bop, BInv, eig + s0, Gin, p.polynomialDegree,
                                      p.wantedSpectrumLowerBound, p.wantedSpectrumUpperBound,
                                      p.unWantedSpectrumUpperBound, Xo, p.useGraph != 0, ax,
                                      bx, n);
              else
                residualFilter<double>(ctx, A, Bop, BInv, eig + s0, Gin, p.polynomialDegree,
                                       p.wantedSpectrumLowerBound, p.wantedSpectrumUpperBound,
                                       p.unWantedSpectrumUpperBound, Xo, p.useGraph != 0, ax,
                                       bx, n);
            }
          else
            {
              if (mixed)
                plainFilter<float>(ctx, A, BInv, Gin, p.polynomialDegree,
                                   p.wantedSpectrumLowerBound, p.wantedSpectrumUpperBound,
                                   p.unWantedSpectrumUpperBound, Xo);
              else
                plainFilter<double>(ctx, A, BInv, Gin, p.polynomialDegree,
                                    p.wantedSpectrumLowerBound, p.wantedSpectrumUpperBound,
                                    p.unWantedSpectrumUpperBound, Xo);
            }
        }
      cs->ms[0] = tf.stop();
      const size_t subBytes = (size_t)XV.N * n * sizeof(double);
      if (p.storeIntermediateSubspaces)
        {
          cs->filtered.ensure(subBytes);
          cs->filteredOrtho.ensure(subBytes);
          cs->storedN = XV.N;
          cs->storedB = n;
          DFX_HIP(hipMemcpyAsync(cs->filtered.p, XV.p, subBytes, hipMemcpyDeviceToDevice,
                                 ctx->stream));
        }
      Timer to(ctx->stream);
      int oc = cholGS(ctx, XV, G, Bop);
      cs->ms[1] = to.stop();
      if (p.storeIntermediateSubspaces && oc == 0)
        DFX_HIP(hipMemcpyAsync(cs->filteredOrtho.p, G.p, subBytes, hipMemcpyDeviceToDevice,
                               ctx->stream));
      if (oc != 0)
        return DFX_EIG_CHFSI_ORTHONORMALIZATION_ERROR;
      Timer tr(ctx->stream);
      int rc = rayleighRitz(ctx, A, G, eig, XV, keep ? "rr.hx" : "cg.mx");
      if (rc == 0 && keep)
        {
          // X = X_f Qc S (Qc: CholGS's L^-T, S: the Ritz vectors), so
          // A X = (A X_o) S and B X = (B X_f)(Qc S).  A diagonal B is
          // applied to X instead (one read of X and one write against the
          // 2 N n^2 flops of the rotation; the same bits as the apply of a
          // solve without kept products, and X's constrained rows get the
          // parent->child values that apply would give them)
          const u32    N    = XV.N;
          const size_t by   = (size_t)N * n * sizeof(double);
          double      *S    = workDoubles(ctx, "rr.S", (size_t)n * n);
          double      *qr   = workDoubles(ctx, "rot.qr", (size_t)n * 16 * ((n + 15) / 16));
          cs->kpHX.ensure(by);
          cs->kpMX.ensure(by);
          rotate64(ctx, cv(workView<double>(ctx, "rr.hx", N, n)), S,
                   View<double>{cs->kpHX.as<double>(), N, n, n}, qr);
          View<double> MXk{cs->kpMX.as<double>(), N, n, n};
          if (Bop->kind == OP_DIAGONAL)
            applyOp<double>(Bop, XV, MXk);
          else
            {
              double *Qc = workDoubles(ctx, "cg.Q", (size_t)n * n);
              double *W  = workDoubles(ctx, "kp.W", (size_t)n * n);
              gemmSquare(ctx, Qc, S, W, n);
              rotate64(ctx, cv(workView<double>(ctx, "cg.mx", N, n)), W, MXk, qr);
            }
          cs->kpHash  = hashOf(ctx, cv(XV));
          cs->kpA     = A->uid;
          cs->kpAGen  = A->gen;
          cs->kpB     = Bop->uid;
          cs->kpBGen  = Bop->gen;
          cs->kpN     = N;
          cs->kpn     = n;
          cs->kpValid = true;
        }
      cs->ms[2] = tr.stop();
      if (rc != 0)
        return DFX_EIG_CHFSI_RAYLEIGH_RITZ_ERROR;
      return 0;
    }

    // LanczosExtremeEigenSolver::solve, non-adaptive (LanczosExtremeEigenSolver.t.cpp:224-540)
    void lanczos(dfx_ctx *ctx, dfx_op *A, dfx_op *Bop, dfx_op *BInv, View<double> guess,
                 u32 maxK, double *lo, double *hi, double *betaLast)
    {
      hipStream_t s = ctx->stream;
      const u32   N = guess.N;
      DFX_REQUIRE(guess.B == 1, "Lanczos guess must be a single vector");
      View<double> g     = workView<double>(ctx, "lz.g", N, 1);
      View<double> temp  = workView<double>(ctx, "lz.t", N, 1);
      View<double> v     = workView<double>(ctx, "lz.v", N, 1);
      View<double> q     = workView<double>(ctx, "lz.q", N, 1);
      View<double> qPrev = workView<double>(ctx, "lz.qp", N, 1);
      k::convert<double, double>(s, g, cv(guess));
      k::fill<double>(s, qPrev, 0.0);
      double dot;
      applyOp<double>(Bop, g, temp);
      colDotsView<double>(ctx, cv(g), cv(temp), &dot);
      double alpha = std::sqrt(dot);
      k::axpby<double>(s, q, 1.0 / alpha, cv(g), 0.0, cv(g));
      double              beta = 0.0;
      std::vector<double> av, bv;
      for (u32 it = 1; it <= maxK; ++it)
        {
          applyOp<double>(A, q, temp);
          applyOp<double>(BInv, temp, v);
          colDotsView<double>(ctx, cv(q), cv(temp), &dot);
          alpha = dot;
          av.push_back(alpha);
          k::axpby<double>(s, v, -alpha, cv(q), 1.0, cv(v));
          k::axpby<double>(s, v, -beta, cv(qPrev), 1.0, cv(v));
          applyOp<double>(Bop, v, temp);
          colDotsView<double>(ctx, cv(v), cv(temp), &dot);
          beta = std::sqrt(dot);
          bv.push_back(beta);
          k::convert<double, double>(s, qPrev, cv(q));
          k::axpby<double>(s, q, 1.0 / beta, cv(v), 0.0, cv(v));
        }
      // eigenvalues of the tridiagonal (alpha, beta[0..k-2]) on the device
      const u32 kk = (u32)av.size();
      double   *d  = workDoubles(ctx, "lz.d", 2 * (size_t)kk + 2);
      toDevice(ctx, d, av.data(), kk);
      if (kk > 1)
        toDevice(ctx, d + kk, bv.data(), kk - 1);
      int *info = work(ctx, "info", 64).as<int>();
      rocblas_status st;
      {
        CaptureLock lk(captureMutex());
        st = rocsolver_dsterf(ctx->blas, (rocblas_int)kk, d, d + kk, info);
      }
      if (st != rocblas_status_success)
        throw Error(DFX_E_HIP, "rocsolver_dsterf failed");
      std::vector<double> w(kk);
      DFX_HIP(hipMemcpyAsync(w.data(), d, kk * sizeof(double), hipMemcpyDeviceToHost, s));
      DFX_HIP(hipStreamSynchronize(s));
      std::sort(w.begin(), w.end());
      *lo       = w.front();
      *hi       = w.back();
      *betaLast = bv.back();
    }

    // CHEBY_ORDER_LOOKUP (src/ksdft/Defaults.cpp:51-58), lower_bound semantics
    // of KohnShamEigenSolver.t.cpp:36-46
    u32 chebyDegree(double upper)
    {
      static const std::map<double, u32> lookup = {
        {10, 6},      {50, 9},      {100, 12},    {200, 16},    {300, 19},
        {500, 24},    {750, 30},    {1000, 39},   {1500, 50},   {2000, 53},
        {3000, 57},   {4000, 62},   {5000, 69},   {9000, 77},   {14000, 104},
        {20000, 119}, {30000, 162}, {50000, 300}, {80000, 450}, {100000, 550},
        {200000, 700}, {500000, 1000}};
      // the reference looks the bound up as a size_type key
      auto it = lookup.lower_bound((double)(u32)upper);
      return it != lookup.end() ? it->second : 1250;
    }

    // fermiDirac / fermiDiracDer (src/ksdft/FractionalOccupancyFunction.cpp:9-33)
    double fermiDirac(double e, double mu, double kb, double T)
    {
      const double f = (e - mu) / (kb * T);
      return f >= 0 ? std::exp(-f) / (1.0 + std::exp(-f)) : 1.0 / (1.0 + std::exp(f));
    }
    double fermiDiracDer(double e, double mu, double kb, double T)
    {
      const double f = (e - mu) / (kb * T), beta = 1.0 / (kb * T);
      return f >= 0 ? beta * std::exp(-f) / (1.0 + std::exp(-f)) / (1.0 + std::exp(-f))
                    : beta * std::exp(f) / (1.0 + std::exp(f)) / (1.0 + std::exp(f));
    }
    constexpr double kBoltzmannHartree = 3.166811429e-06; // ksdft/Defaults.cpp:69

    // Fermi level: NewtonRaphsonSolver on sum_i 2 f_i - Ne
    // (src/linearAlgebra/NewtonRaphsonSolver.t.cpp, FractionalOccupancyFunction.cpp:49-71),
    // started at eps[ceil(Ne/2) - 1] (KohnShamEigenSolver.t.cpp:396-425)
    double fermiLevel(const std::vector<double> &eps, double Ne, double T, double tol)
    {
      size_t i0 = (size_t)std::ceil(Ne / 2.0);
      i0        = i0 ? std::min(i0 - 1, eps.size() - 1) : 0;
      double x  = eps[i0];
      for (int it = 0; it < 100000; ++it)
        {
          double g = -Ne, gp = 0.0;
          for (double e : eps)
            {
              g += 2 * fermiDirac(e, x, kBoltzmannHartree, T);
              gp += 2 * fermiDiracDer(e, x, kBoltzmannHartree, T);
            }
          if (gp == 0.0)
            break;
          const double x1 = x - g / gp;
          if (std::fabs(x1 - x) < tol)
            return x1;
          x = x1;
        }
      return x;
    }

    void eigenResidualNorms(dfx_ctx *ctx, dfx_op *H, dfx_op *M, View<double> X,
                       const double *eig, double *res)
    {
      hipStream_t s = ctx->stream;
      const u32   N = X.N, n = X.B;
      View<double> x  = workView<double>(ctx, "rn.x", N, n);
      View<double> hx = workView<double>(ctx, "rn.hx", N, n);
      View<double> mx = workView<double>(ctx, "rn.mx", N, n);
      k::convert<double, double>(s, x, cv(X));
      applyOp<double>(H, x, hx);
      applyOp<double>(M, x, mx);
      residualNormsOf(ctx, cv(hx), cv(mx), eig, res);
    }
  } // namespace
} // namespace dfx

namespace dfx
{
  namespace
  {
    // CG step scalars per column, as the host formed them: coef[j] = 1,
    // coef[B + j] = zr / pw (alpha), coef[2B + j] = -zr / pw
    __global__ void cgAlphaK(const double *zr, const double *pw, u32 B, double *coef)
    {
      const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
      if (j >= B)
        return;
      coef[j]         = 1.0;
      coef[B + j]     = zr[j] / pw[j];
      coef[2 * B + j] = -zr[j] / pw[j];
    }
    // beta[j] = zrNew / zr
    __global__ void cgBetaK(const double *zrNew, const double *zr, u32 B, double *beta)
    {
      const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
      if (j < B)
        beta[j] = zrNew[j] / zr[j];
    }
  } // namespace
} // namespace dfx

extern "C"
{
  int dfx_chfsi_create(dfx_ctx *ctx, const dfx_chfsi_params *p, dfx_mv *guess,
                       dfx_chfsi **out)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(p && guess && out, "null argument");
      DFX_REQUIRE(guess->dtype == DFX_F64, "the subspace guess must be FP64");
      if (p->orthogonalizationType == DFX_ORTHO_MULTIPASS_LOWDIN)
        throw Error(DFX_E_UNSUPPORTED, "Orthogonalization type not present");
      if (p->orthogonalizationType != DFX_ORTHO_CHOLESKY_GRAMSCHMIDT)
        throw Error(DFX_E_UNSUPPORTED,
                    "only CHOLESKY_GRAMSCHMIDT orthogonalization is implemented");
      auto cs   = new dfx_chfsi();
      cs->ctx   = ctx;
      cs->p     = *p;
      cs->guess = guess;
      ctxRetain(ctx);
      *out      = cs;
      return 0;
    });
  }

  int dfx_chfsi_reinit(dfx_chfsi *cs, double a, double b, double c, uint32_t degree,
                       double tol, dfx_mv *guess)
  {
    if (!cs)
      return DFX_E_ARG;
    cs->p.wantedSpectrumLowerBound   = a;
    cs->p.wantedSpectrumUpperBound   = b;
    cs->p.unWantedSpectrumUpperBound = c;
    cs->p.polynomialDegree           = degree;
    cs->p.illConditionTolerance      = tol;
    if (guess)
      cs->guess = guess;
    return 0;
  }

  int dfx_chfsi_solve(dfx_chfsi *cs, dfx_op *A, double *eig, dfx_mv *X, int computeEV,
                      dfx_op *Bop, dfx_op *BInv)
  {
    if (!cs || !A || !eig || !X || !Bop || !BInv)
      return DFX_E_ARG;
    (void)computeEV;
    return guard(cs->ctx, [&] { return chfsiSolve(cs, A, eig, X, Bop, BInv); });
  }

  int dfx_chfsi_get_filtered_subspace(const dfx_chfsi *cs, int ortho, dfx_mv *out)
  {
    if (!cs || !out)
      return DFX_E_ARG;
    return guard(cs->ctx, [&] {
      if (!cs->p.storeIntermediateSubspaces)
        throw Error(DFX_E_STATE, "storeIntermediateSubspaces is false in CHFSI class. Cannot "
                                 "return the filtered Subspace.");
      DFX_REQUIRE(cs->storedN, "no solve has run yet");
      DFX_REQUIRE(out->dtype == DFX_F64 && out->N == cs->storedN && out->B == cs->storedB,
                  "output must be FP64 with the eigenvectors' shape");
      const DevBuf &b = ortho ? cs->filteredOrtho : cs->filtered;
      DFX_HIP(hipMemcpyAsync(out->d, b.p, (size_t)cs->storedN * cs->storedB * sizeof(double),
                             hipMemcpyDeviceToDevice, cs->ctx->stream));
      DFX_HIP(hipStreamSynchronize(cs->ctx->stream));
      return 0;
    });
  }

  int dfx_chfsi_timings(const dfx_chfsi *cs, double *ms3)
  {
    if (!cs || !ms3)
      return DFX_E_ARG;
    std::copy(cs->ms, cs->ms + 3, ms3);
    return 0;
  }

  int dfx_chfsi_destroy(dfx_chfsi *cs)
  {
    if (!cs)
      return DFX_E_ARG;
    dfx_ctx *ctx = cs->ctx;
    delete cs;
    ctxRelease(ctx);
    return 0;
  }

  int dfx_cholesky_gram_schmidt(dfx_ctx *ctx, dfx_mv *X, dfx_mv *Xo, dfx_op *Bop)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(X && Xo && Bop, "null argument");
      return cholGS(ctx, mvView(X), mvView(Xo), Bop);
    });
  }

  int dfx_rayleigh_ritz(dfx_ctx *ctx, dfx_op *A, dfx_mv *X, double *eig, dfx_mv *Xo)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(A && X && Xo && eig, "null argument");
      return rayleighRitz(ctx, A, mvView(X), eig, mvView(Xo));
    });
  }

  int dfx_lanczos_extreme(dfx_ctx *ctx, dfx_op *A, dfx_op *Bop, dfx_op *BInv,
                          dfx_mv *guess, uint32_t maxK, double *lo, double *hi,
                          double *betaLast)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(A && Bop && BInv && guess && lo && hi && betaLast, "null argument");
      DFX_REQUIRE(maxK >= 2, "maxKrylov must be >= 2");
      lanczos(ctx, A, Bop, BInv, mvView(guess), maxK, lo, hi, betaLast);
      return 0;
    });
  }

  int dfx_ks_eigensolve(dfx_ctx *ctx, const dfx_ks_params *p, dfx_op *H, dfx_op *M,
                        dfx_op *MInv, dfx_op *MLanczos, dfx_mv *guess,
                        dfx_mv *lanczosGuess, double *energies, dfx_mv *waveFns,
                        double *residualNorms, dfx_ks_result *result)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(p && H && M && MInv && MLanczos && guess && lanczosGuess && energies &&
                    waveFns,
                  "null argument");
      const u32 n = waveFns->B, N = waveFns->N;
      DFX_REQUIRE(p->numWantedEigenvalues == n, "numWantedEigenvalues must equal the width");
      dfx_ks_result R{};
      // ---- spectrum bounds (KohnShamEigenSolver.t.cpp:240-297)
      auto t0 = std::chrono::steady_clock::now();
      double lo, hi, beta;
      lanczos(ctx, H, MLanczos, MInv, mvView(lanczosGuess),
              p->lanczosMaxKrylov ? p->lanczosMaxKrylov : 20, &lo, &hi, &beta);
      R.lanczosMs = std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t0).count();
      const double residual = beta / 10.0;
      double a = lo;
      double b = (hi + residual - lo) * ((double)(n * 200.0) / N) + lo;
      if (b >= hi + residual)
        b = (hi + residual + lo) * 0.5;
      const double c  = hi + residual;
      const u32    m  = p->polynomialDegree ? p->polynomialDegree : chebyDegree(c);
      R.lanczosLow  = lo;
      R.lanczosHigh = hi;
      R.lanczosBeta = beta;
      R.degree      = m;

      dfx_chfsi cs;
      cs.ctx   = ctx;
      cs.guess = guess;
      cs.p.illConditionTolerance     = 1e-14;
      cs.p.eigenVectorBatchSize      = p->waveFunctionBatchSize;
      cs.p.precision                 = p->precision;
      cs.p.useGraph                  = p->useGraph;
      // residual-filter runs start from one degree-1 plain-filter pass
      // (src/ksdft/KohnShamDFT.t.cpp:462-487)
      if (p->isResidualChebyshevFilter)
        {
          cs.p.isResidualChebyshevFilter  = 0;
          cs.p.wantedSpectrumLowerBound   = a;
          cs.p.wantedSpectrumUpperBound   = b;
          cs.p.unWantedSpectrumUpperBound = c;
          cs.p.polynomialDegree           = 1;
          int rc = chfsiSolve(&cs, H, energies, waveFns, M, MInv);
          if (rc != 0)
            return DFX_EIG_KS_CHFSI_ERROR;
          k::convert<double, double>(ctx->stream, mvView(guess), cv(mvView(waveFns)));
        }
      cs.p.isResidualChebyshevFilter  = p->isResidualChebyshevFilter;
      cs.p.wantedSpectrumLowerBound   = a;
      cs.p.wantedSpectrumUpperBound   = b;
      cs.p.unWantedSpectrumUpperBound = c;
      cs.p.polynomialDegree           = m;
      std::vector<double> res(n, 0.0);
      u32 pass = 0;
      int status = DFX_EIG_KS_MAX_PASS_ERROR;
      for (; pass < p->maxChebyshevFilterPass; ++pass)
        {
          int rc = chfsiSolve(&cs, H, energies, waveFns, M, MInv);
          R.filterMs += cs.ms[0];
          R.orthoMs += cs.ms[1];
          R.rrMs += cs.ms[2];
          if (rc != 0)
            {
              status = DFX_EIG_KS_CHFSI_ERROR;
              break;
            }
          auto tr = std::chrono::steady_clock::now();
          // waveFns is the solve's output, so the kept products are its own
          if (productsFor(&cs, H, M, N, n))
            productResidualNorms(&cs, energies, res.data());
          else
            eigenResidualNorms(ctx, H, M, mvView(waveFns), energies, res.data());
          R.residualMs += std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - tr).count();
          k::convert<double, double>(ctx->stream, mvView(guess), cv(mvView(waveFns)));
          u32 nOcc = n, nOccConv = 0;
          if (p->numElectrons > 0)
            {
              std::vector<double> eps(energies, energies + n);
              R.fermiEnergy = fermiLevel(eps, p->numElectrons, p->smearingTemperature,
                                         p->fermiEnergyTolerance);
              nOcc = 0;
              for (u32 j = 0; j < n; ++j)
                if (fermiDirac(eps[j], R.fermiEnergy, kBoltzmannHartree,
                               p->smearingTemperature) > p->fracOccupancyTolerance)
                  {
                    ++nOcc;
                    if (res[j] <= p->eigenSolveResidualTolerance)
                      ++nOccConv;
                  }
            }
          else
            for (u32 j = 0; j < n; ++j)
              if (res[j] <= p->eigenSolveResidualTolerance)
                ++nOccConv;
          R.numOccupied = nOcc;
          if (nOccConv == nOcc)
            {
              status = 0;
              ++pass;
              break;
            }
          cs.p.wantedSpectrumLowerBound = energies[0];
          cs.p.wantedSpectrumUpperBound = energies[n - 1];
        }
      R.passes = pass;
      R.a      = cs.p.wantedSpectrumLowerBound;
      R.b      = cs.p.wantedSpectrumUpperBound;
      R.c      = c;
      if (residualNorms)
        std::copy(res.begin(), res.end(), residualNorms);
      if (result)
        *result = R;
      return status;
    });
  }

  // CGLinearSolver::solve (src/linearAlgebra/CGLinearSolver.t.cpp:68-313).
  // The per-column scalars alpha = z.r / p.w and beta = z.r_new / z.r are
  // formed on the device from device-resident dots (the same IEEE divisions
  // as on the host), so one iteration synchronizes once, for the residual
  // norms of the convergence test.
  int dfx_eigen_residual_norms(dfx_ctx *ctx, dfx_op *H, dfx_op *M, dfx_mv *X,
                               const double *eigenValues, double *residualNorms)
  {
    if (!ctx || !H || !M || !X || !eigenValues || !residualNorms)
      return DFX_E_ARG;
    return guard(ctx, [&] {
      View<double> x = mvView(X);
      DFX_REQUIRE(H->nDofs == x.N && M->nDofs == x.N, "operator size does not match X");
      eigenResidualNorms(ctx, H, M, x, eigenValues, residualNorms);
      return 0;
    });
  }

  int dfx_chfsi_residual_norms(dfx_chfsi *cs, dfx_op *H, dfx_op *M, dfx_mv *X,
                               const double *eigenValues, double *residualNorms)
  {
    if (!cs || !H || !M || !X || !eigenValues || !residualNorms)
      return DFX_E_ARG;
    return guard(cs->ctx, [&] {
      View<double> x = mvView(X);
      DFX_REQUIRE(H->nDofs == x.N && M->nDofs == x.N, "operator size does not match X");
      if (productsFor(cs, H, M, x.N, x.B) && hashOf(cs->ctx, cv(x)) == cs->kpHash)
        productResidualNorms(cs, eigenValues, residualNorms);
      else
        eigenResidualNorms(cs->ctx, H, M, x, eigenValues, residualNorms);
      return 0;
    });
  }

  int dfx_cg_solve(dfx_ctx *ctx, dfx_op *A, dfx_op *pc, dfx_mv *bmv, dfx_mv *xmv,
                   double absTol, double relTol, double divTol, uint32_t maxIter,
                   uint32_t *iterations, double *residualNorms)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(A && pc && bmv && xmv, "null argument");
      hipStream_t  s = ctx->stream;
      View<double> b = mvView(bmv), xo = mvView(xmv);
      DFX_REQUIRE(b.N == xo.N && b.B == xo.B, "b and x differ in shape");
      const u32 N = b.N, B = b.B;
      View<double> x  = workView<double>(ctx, "cgs.x", N, B);
      View<double> r  = workView<double>(ctx, "cgs.r", N, B);
      View<double> w  = workView<double>(ctx, "cgs.w", N, B);
      View<double> z  = workView<double>(ctx, "cgs.z", N, B);
      View<double> pv = workView<double>(ctx, "cgs.p", N, B);
      double *coef = workDoubles(ctx, "cgs.coef", 4 * (size_t)B);
      // device dots: z.r (two, alternating), p.w, r.r; partial sums
      double *dots = workDoubles(ctx, "cgs.dots", 4 * (size_t)B);
      double *part = workDoubles(ctx, "cgs.part", 2048 * (size_t)B);
      double *dZr = dots, *dZrNew = dots + B, *dPw = dots + 2 * B, *dRr = dots + 3 * B;
      // a Jacobi preconditioner (diagonal, no constraints, no enriched block)
      // lets one launch do the updates of x, r, z and the dots z.r, r.r and
      // one more form beta and the search direction (the same bits;
      // DFX_FUSE_CG=0 keeps the separate launches); z.r and r.r of an
      // iteration land side by side in one of two 2B slots
      const char *fuseEnv = getenv("DFX_FUSE_CG");
      const bool  fusedCG = (!fuseEnv || atoi(fuseEnv) != 0) && pc->kind == OP_DIAGONAL &&
                           !pc->cons && pc->nEnriched == 0 && pc->nDofs == N;
      double *slots = fusedCG ? workDoubles(ctx, "cgs.zrrr", 4 * (size_t)B) : nullptr;
      if (fusedCG)
        dZr = slots;
      k::convert<double, double>(s, x, cv(xo));
      std::vector<double> bNorm(B), rNorm(B);
      colDotsView<double>(ctx, cv(b), cv(b), bNorm.data());
      for (auto &v : bNorm)
        v = std::sqrt(v);
      std::vector<char> conv(B, 0);
      bool diverged = false, allConv = false;
      int  err  = DFX_LS_OTHER_ERROR;
      // the convergence test of an iteration (CGLinearSolver.t.cpp's per
      // column test): converged columns of x are copied to the output
      // (from xk), divergence is flagged; true when the loop stops there.
      // Pipelined (DFX_CG_PIPELINE, default on), iteration k's test runs
      // while iteration k + 1 is on the device, from the r.r and the copy of
      // x that iteration k left: the same results and iteration count, and
      // one iteration of work past the last is computed and dropped
      const char *pipeEnv = getenv("DFX_CG_PIPELINE");
      const bool  pipe    = !pipeEnv || atoi(pipeEnv) != 0;
      View<double> xs[2];
      if (pipe)
        {
          xs[0] = workView<double>(ctx, "cgs.xs0", N, B);
          xs[1] = workView<double>(ctx, "cgs.xs1", N, B);
        }
      PinnedDoubles hr(2 * (size_t)B);
      EventPair     ev;
      auto check = [&](u32 k) {
        const double *h = hr.p + (size_t)(k % 2) * B;
        for (u32 j = 0; j < B; ++j)
          rNorm[j] = std::sqrt(h[j]);
        std::vector<double> mask(2 * (size_t)B, 0.0);
        bool any = false;
        for (u32 j = 0; j < B; ++j)
          {
            mask[B + j] = 1.0;
            if (rNorm[j] < std::max(absTol, bNorm[j] * relTol) && !conv[j])
              {
                err         = 0;
                conv[j]     = 1;
                mask[j]     = 1.0;
                mask[B + j] = 0.0;
                any         = true;
              }
            if (rNorm[j] > divTol && !diverged)
              {
                err      = DFX_LS_RESIDUAL_DIVERGENCE;
                diverged = true;
              }
          }
        if (any)
          {
            double *dm = workDoubles(ctx, "cgs.mask", 2 * (size_t)B);
            toDevice(ctx, dm, mask.data(), mask.size());
            k::axpbyBlocked<double>(s, xo, 1.0, dm, pipe ? cv(xs[k % 2]) : cv(x), 1.0, dm + B,
                                    cv(xo));
          }
        allConv = std::all_of(conv.begin(), conv.end(), [](char c) { return c != 0; });
        return diverged || allConv;
      };
      u32      iter = 0;
      uint64_t stop = (uint64_t)maxIter + 1; // the iteration whose test ended the loop
      for (; iter <= maxIter; ++iter)
        {
          if (iter == 0)
            {
              applyOp<double>(A, x, w);
              k::axpby<double>(s, r, 1.0, cv(b), -1.0, cv(w));
              applyOp<double>(pc, r, z);
              k::convert<double, double>(s, pv, cv(z));
              k::colDots<double>(s, cv(z), cv(r), part, dZr);
            }
          else
            {
              // z.r of this iteration is the previous iteration's z.r_new
              applyOp<double>(A, pv, w);
              k::colDots<double>(s, cv(pv), cv(w), part, dPw);
            }
          if (iter > 0 && fusedCG)
            {
              double *next = dZr == slots ? slots + 2 * B : slots;
              k::cgUpdate(s, x, r, z, cv(pv), cv(w), pc->diag.as<double>(), dZr, dPw, part,
                          next, pipe ? xs[iter % 2].p : nullptr);
              k::cgDirection(s, pv, cv(z), next, dZr);
              dZr = next;
              dRr = next + B;
            }
          else if (iter > 0)
            {
              cgAlphaK<<<(B + 63) / 64, 64, 0, s>>>(dZr, dPw, B, coef);
              checkLaunch("cgAlpha");
              k::axpbyBlocked<double>(s, x, 1.0, coef, cv(x), 1.0, coef + B, cv(pv));
              k::axpbyBlocked<double>(s, r, 1.0, coef, cv(r), 1.0, coef + 2 * B, cv(w));
              applyOp<double>(pc, r, z);
              k::colDots<double>(s, cv(z), cv(r), part, dZrNew);
              cgBetaK<<<(B + 63) / 64, 64, 0, s>>>(dZrNew, dZr, B, coef + 3 * B);
              checkLaunch("cgBeta");
              k::axpbyBlocked<double>(s, pv, 1.0, coef, cv(z), 1.0, coef + 3 * B, cv(pv));
              std::swap(dZr, dZrNew);
            }
          if (!(fusedCG && iter > 0))
            k::colDots<double>(s, cv(r), cv(r), part, dRr);
          // r.r of the iteration to pinned memory; pipelined, x as it stands
          // now is kept for the copy to the output
          if (pipe && !(fusedCG && iter > 0)) // the fused update wrote it
            k::convert<double, double>(s, xs[iter % 2], cv(x));
          DFX_HIP(hipMemcpyAsync(hr.p + (size_t)(iter % 2) * B, dRr, B * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
          DFX_HIP(hipEventRecord(ev.e[iter % 2], s));
          if (!pipe || iter > 0)
            {
              const u32 k = pipe ? iter - 1 : iter;
              DFX_HIP(hipEventSynchronize(ev.e[k % 2]));
              if (check(k))
                {
                  stop = k;
                  break;
                }
            }
        }
      // pipelined, the last iteration's test is still to be made
      if (pipe && stop > maxIter)
        {
          DFX_HIP(hipEventSynchronize(ev.e[maxIter % 2]));
          if (check(maxIter))
            stop = maxIter;
        }
      DFX_HIP(hipStreamSynchronize(s));
      iter = stop <= maxIter ? (u32)stop : maxIter + 1;
      if (iter > maxIter)
        err = DFX_LS_FAILED_TO_CONVERGE;
      if (iterations)
        *iterations = iter;
      if (residualNorms)
        std::copy(rNorm.begin(), rNorm.end(), residualNorms);
      return err;
    });
  }
}
