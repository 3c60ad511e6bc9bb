// C ABI of libdfx, part 1: context, multivectors, constraints, operators.
#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "api_util.hpp"
#include "kernels.hpp"

using namespace dfx;

namespace dfx
{
  // ---------------------------------------------------------------- apply
  template <class T>
  View<T> view(dfx_mv *m)
  {
    return View<T>{static_cast<T *>(m->d), m->N, m->B, m->B};
  }

  template <class T>
  View<const T> cview(const dfx_mv *m)
  {
    return View<const T>{static_cast<const T *>(m->d), m->N, m->B, m->B};
  }

  template <class T>
  const T *fragOf(const dfx_op *op)
  {
    if (sizeof(T) == 8)
      return reinterpret_cast<const T *>(op->frag64.p);
    if (!op->hasF32)
      throw Error(DFX_E_STATE, "operator has no FP32 cell matrices (create with DFX_OP_KEEP_FP32)");
    return reinterpret_cast<const T *>(op->frag32.p);
  }

  // Y = C^T (sum_e Q_e^T A_e Q_e) C X, KohnShamOperatorContextFE::apply
  // (src/ksdft/KohnShamOperatorContextFE.t.cpp:1326-1415)
  // dsc / xe: see applyInvThenCellwise (classical rows gathered as dsc o X,
  // enriched rows from xe; X already constrained)
  // skipP2C: X is used as given (the caller has applied the parent->child
  // step of other constraints, as LaplaceOperatorContextFE::apply does with
  // its X manager's constraints, LaplaceOperatorContextFE.t.cpp:433-436)
  // yEOverride: caller-provided scratch for the enriched partial rows (at
  // least nEnrSlots x B values of T): nothing is allocated and no profiling
  // record is taken
  // rf (with dsc): the residual Chebyshev step fused in; Y then holds
  // partial sums on the rows the cells finish, and res is complete on
  // return (the listed remaining rows by a separate step)
  template <class T>
  void applyCellwise(dfx_op *op, View<T> X, View<T> Y, const double *dsc = nullptr,
                     const T *xe = nullptr, bool skipP2C = false, T *yEOverride = nullptr,
                     const k::ResFuse<T> *rf = nullptr)
  {
    hipStream_t s = op->ctx->stream;
    if (!dsc && !skipP2C)
      k::p2c<T>(s, op->cons, X);
    // every classical row is stored by the first cell that reaches it and every
    // enriched row by the enriched reduction, so Y is cleared only when some
    // dof lies in no cell
    if (op->fillY || (op->nDofs > op->nClassical && !op->nEnrSlots))
      k::fill<T>(s, Y, 0.0);
    T *yE = nullptr;
    if (op->nEnrSlots && yEOverride)
      yE = yEOverride;
    else if (op->nEnrSlots)
      {
        if (op->yE.ensure((size_t)op->nEnrSlots * X.B * sizeof(double)))
          ++op->gen;
        yE = op->yE.as<T>();
      }
    const T *frag = fragOf<T>(op);
    const auto &L = sizeof(T) == 8 ? op->launches64 : op->launches32;
    View<const T> Xc{X.p, X.N, X.B, X.ld};
    // the size classes of one colour touch disjoint dofs: the second one of a
    // colour runs on the context's second stream, beside the first, so that
    // the tail of one launch overlaps the other
    static const bool forkEnv = [] {
      const char *e = getenv("DFX_CELL_FORK");
      return !e || atoi(e) != 0;
    }();
    const bool fork = forkEnv && op->ctx->forkClasses;
    const bool profKind =
      op->profiling && !yEOverride && (op->profMask & (sizeof(T) == 8 ? 1 : 2));
    const bool prof = profKind && (op->profApplies++ % op->profEvery) == 0;
    ProfRec    r{};
    // the end event of a colour step is the start event of the next one
    hipEvent_t prevEnd  = nullptr;
    auto       newEvent = [&]() {
      hipEvent_t e = nullptr;
      DFX_HIP(hipEventCreate(&e));
      op->profEvents.push_back(e);
      DFX_HIP(hipEventRecord(e, s));
      return e;
    };
    for (size_t li = 0; li < L.size(); ++li)
      {
        const Launch &l     = L[li];
        const bool    first = li == 0 || L[li - 1].color != l.color;
        const bool    last  = li + 1 == L.size() || L[li + 1].color != l.color;
        const bool    side  = fork && !first;
        hipStream_t   ls    = side ? op->ctx->stream2 : s;
        // one profile record per colour step (its launches, side by side),
        // bracketed on the main stream
        if (prof && first)
          {
            r   = ProfRec{};
            r.a = prevEnd ? prevEnd : newEvent();
          }
        if (fork && first && !last)
          {
            // the next launch (same colour) will go to stream2 after this point
            DFX_HIP(hipEventRecord(op->ctx->forkEv, s));
            DFX_HIP(hipStreamWaitEvent(op->ctx->stream2, op->ctx->forkEv, 0));
          }
        k::cellApply<T>(ls, op, l, frag, Xc, Y, yE, dsc, xe, 0, rf);
        if (side)
          {
            DFX_HIP(hipEventRecord(op->ctx->joinEv, ls));
            DFX_HIP(hipStreamWaitEvent(s, op->ctx->joinEv, 0));
          }
        if (prof)
          {
            const double B = X.B, sz = sizeof(T);
            r.flops += 2.0 * B * (double)l.sumN2;
            // cell matrices once, the map, gathered X rows, classical rows of
            // Y written (and read unless first touched), enriched partials
            r.bytes += sz * (double)l.sumN2 + 4.0 * (double)l.sumN + sz * B * (double)l.sumN +
                       sz * B * (2.0 * (double)l.sumNc - (double)l.sumFirst) +
                       sz * B * (double)(l.sumN - l.sumNc) +
                       (dsc ? 8.0 * (double)l.sumNc : 0.0);
            if (last)
              {
                r.b = prevEnd = newEvent();
                op->prof.push_back(r);
              }
          }
      }
    if (op->nEnrSlots)
      k::enrichedReduce<T>(s, op, yE, Y);
    // with the residual step fused in, the constraints are homogeneous
    // Dirichlet ones whose rows the step leaves out, so nothing reads Y there
    if (!rf)
      k::c2p<T>(s, op->cons, Y);
    if (rf)
      k::resFix<T>(s, View<T>{rf->res, X.N, X.B, X.ld}, View<const T>{Y.p, Y.N, Y.B, Y.ld}, Xc,
                   View<const T>{rf->r0, X.N, X.B, X.ld}, op->resRows.as<u32>(), op->nResRows,
                   rf->mu, rf->coef);
  }

  // Y = C2P(diag o P2C(X) + [0; E X_E])
  template <class T>
  void applyDiagonal(dfx_op *op, View<T> X, View<T> Y)
  {
    hipStream_t s = op->ctx->stream;
    k::p2c<T>(s, op->cons, X);
    const u32 nC = op->nDofs - op->nEnriched;
    View<const T> Xc{X.p, X.N, X.B, X.ld};
    k::diagScale<T>(s, Y, op->diag.as<double>(), Xc, nC);
    if (op->nEnriched)
      k::enrichedBlock<T>(s, Y, op->enrBlock.as<double>(), Xc, nC, op->nEnriched);
    k::c2p<T>(s, op->cons, Y);
  }

  // Column-chunk autotuner: for each M-tile class of the operator (launch
  // order of first appearance), time `reps` passes over the class's
  // launches (all colours, one stream) per candidate width on zeroed scratch
  // multivectors and keep the fastest.  Candidates: 1, 2, 4, 8 tiles of 16
  // columns up to twice the columns there are, 5 for FP32 with 65..80
  // columns, and only tiles that fit the LDS of a CU.  Per-column results do
  // not depend on the width (same K order per column), only the time does.
  template <class T>
  std::vector<std::pair<u32, int>> autotune(dfx_op *op, u32 B, u32 reps)
  {
    hipStream_t s    = op->ctx->stream;
    const auto &L    = sizeof(T) == 8 ? op->launches64 : op->launches32;
    const T    *frag = fragOf<T>(op);
    const u32   N    = op->nDofs;
    DevBuf xb((size_t)N * B * sizeof(T)), yb((size_t)N * B * sizeof(T));
    DFX_HIP(hipMemsetAsync(xb.p, 0, xb.bytes, s));
    DFX_HIP(hipMemsetAsync(yb.p, 0, yb.bytes, s));
    T *yE = nullptr;
    if (op->nEnrSlots)
      {
        if (op->yE.ensure((size_t)op->nEnrSlots * B * sizeof(double)))
          ++op->gen;
        yE = op->yE.as<T>();
      }
    View<const T> X{xb.as<T>(), N, B, B};
    View<T>       Y{yb.as<T>(), N, B, B};
    const bool    buf = k::cellBufferPath<T>(op, X, View<const T>{Y.p, N, B, B});
    std::vector<u32> classes;
    for (const Launch &l : L)
      if (std::find(classes.begin(), classes.end(), l.mtw) == classes.end())
        classes.push_back(l.mtw);
    op->tuned.erase(std::remove_if(op->tuned.begin(), op->tuned.end(),
                                   [&](const dfx_op::Tuned &t) {
                                     return t.s == sizeof(T) && t.B == B;
                                   }),
                    op->tuned.end());
    hipEvent_t a, b;
    DFX_HIP(hipEventCreate(&a));
    DFX_HIP(hipEventCreate(&b));
    std::vector<std::pair<u32, int>> out;
    try
      {
        for (u32 c : classes)
          {
            u32 kpad = 0;
            for (const Launch &l : L)
              if (l.mtw == c)
                kpad = std::max(kpad, l.kpadMax);
            const u32 rows = buf ? 64 * c : kpad; // K rows of the LDS tile
            int       best = 0;
            float     bestMs = 0.f;
            for (int nt : {1, 2, 4, 5, 8})
              {
                if (nt == 5 && !(sizeof(T) == 4 && B > 64 && B <= 80))
                  continue;
                if (nt > 1 && 16u * nt > 2 * B)
                  continue;
                if ((size_t)rows * 16 * nt * sizeof(T) + 2048 > 160 * 1024)
                  continue;
                auto pass = [&] {
                  for (const Launch &l : L)
                    if (l.mtw == c)
                      k::cellApply<T>(s, op, l, frag, X, Y, yE, nullptr, nullptr, nt);
                };
                pass(); // first launch: kernel attributes, instruction cache
                DFX_HIP(hipEventRecord(a, s));
                for (u32 r = 0; r < reps; ++r)
                  pass();
                DFX_HIP(hipEventRecord(b, s));
                DFX_HIP(hipEventSynchronize(b));
                float ms = 0.f;
                DFX_HIP(hipEventElapsedTime(&ms, a, b));
                if (!best || ms < bestMs)
                  best = nt, bestMs = ms;
              }
            if (!best)
              best = k::builtinNT<T>(B, kpad);
            op->tuned.push_back(dfx_op::Tuned{(u32)sizeof(T), B, c, best});
            out.emplace_back(c, best);
          }
      }
    catch (...)
      {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
      }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return out;
  }

  // Y = A (Dinv X) with Dinv a diagonal (+ enriched block) operator and A a
  // cell-wise one sharing homogeneous Dirichlet constraints (no parents, zero
  // values, classical rows only): the diagonal is applied inside the cell
  // gather instead of as a pass of its own, and only the enriched rows of
  // Dinv X are formed (in S1).  Every value equals the unfused sequence
  // Dinv.apply(X, S1); A.apply(S1, Y) bit for bit (the same products in the
  // same precision; constrained rows are zero either way) and X receives the
  // same parent->child update.  Returns false (nothing done) when the
  // operators or the multivectors do not qualify.
  // rf: the residual Chebyshev step fused in (res and r0 N x B at X's
  // leading dimension); res then equals resStep(Y, X, res, r0) bit for bit
  // and Y holds partial sums on the rows the cells finish.  The caller's X,
  // res and r0 are zero on the constrained rows (the residual filter's
  // buffers are, from its seed on, since A and M zero them): the step keeps
  // them zero there, so the P2C on X, the C2P on Y and the step on those
  // rows are left out (Y's constrained rows keep partial sums)
  template <class T>
  bool applyInvThenCellwise(dfx_op *Dinv, dfx_op *A, View<T> X, View<T> S1, View<T> Y,
                            const k::ResFuse<T> *rf)
  {
    if (Dinv->kind != OP_DIAGONAL || A->kind != OP_CELLWISE || Dinv->cons != A->cons)
      return false;
    if (Dinv->nDofs != A->nDofs || Dinv->nEnriched != A->nDofs - A->nClassical)
      return false;
    if (A->cons && (!A->cons->homDirichlet ||
                    (!A->cons->hRows.empty() && A->cons->hRows.back() >= A->nClassical)))
      return false;
    if ((sizeof(T) == 4 && !A->hasF32) || X.N != A->nDofs || Y.N != A->nDofs ||
        S1.N != A->nDofs || S1.ld != X.ld || X.B != Y.B || S1.B != X.B)
      return false;
    if (!k::cellBufferPath<T>(A, View<const T>{X.p, X.N, X.B, X.ld},
                              View<const T>{Y.p, Y.N, Y.B, Y.ld}) ||
        (reinterpret_cast<uintptr_t>(S1.p) & 15) != 0)
      return false;
    hipStream_t s = A->ctx->stream;
    if (A->foldedUid != Dinv->uid)
      {
        // d[map[slot]] for every slot of the cell->dof map (classical slots
        // only are read), so the kernel loads it beside the map
        A->foldedDiag.ensure((size_t)A->sumNe * sizeof(double));
        k::gatherDiag(s, Dinv->diag.as<double>(), A->map.as<u32>(), A->sumNe,
                      A->nClassical, A->foldedDiag.as<double>());
        A->foldedUid = Dinv->uid;
        ++A->gen;
      }
    if (!rf)
      k::p2c<T>(s, Dinv->cons, X);
    if (Dinv->nEnriched)
      k::enrichedBlock<T>(s, S1, Dinv->enrBlock.as<double>(),
                          View<const T>{X.p, X.N, X.B, X.ld}, A->nClassical, Dinv->nEnriched);
    applyCellwise<T>(A, X, Y, A->foldedDiag.as<double>(), S1.p, false, nullptr, rf);
    return true;
  }
  template bool applyInvThenCellwise<double>(dfx_op *, dfx_op *, View<double>, View<double>,
                                             View<double>, const k::ResFuse<double> *);
  template bool applyInvThenCellwise<float>(dfx_op *, dfx_op *, View<float>, View<float>,
                                            View<float>, const k::ResFuse<float> *);

  template <class T>
  void applyOp(dfx_op *op, View<T> X, View<T> Y)
  {
    DFX_REQUIRE(X.N == op->nDofs && Y.N == op->nDofs, "apply: N does not match the operator");
    DFX_REQUIRE(X.B == Y.B, "apply: X and Y have different widths");
    if (op->kind == OP_CELLWISE)
      applyCellwise<T>(op, X, Y);
    else
      applyDiagonal<T>(op, X, Y);
  }

  void ctxRelease(dfx_ctx *ctx)
  {
    if (!ctx || ctx->refs.fetch_sub(1) != 1)
      return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream)
      (void)hipStreamSynchronize(ctx->stream);
    CaptureLock lk(captureMutex());
    if (ctx->lt && ctx->ltFree)
      ctx->ltFree(ctx->lt);
    ctx->lt = nullptr;
    if (ctx->blas)
      rocblas_destroy_handle(ctx->blas);
    for (auto &g : ctx->graphs)
      if (g.second.exec)
        (void)hipGraphExecDestroy(g.second.exec);
    ctx->graphs.clear();
    ctx->scratchSmall.release();
    ctx->reducePartials.release();
    ctx->work.clear();
    if (ctx->stream2)
      {
        (void)hipStreamSynchronize(ctx->stream2);
        (void)hipStreamDestroy(ctx->stream2);
      }
    if (ctx->forkEv)
      (void)hipEventDestroy(ctx->forkEv);
    if (ctx->joinEv)
      (void)hipEventDestroy(ctx->joinEv);
    if (ctx->stream)
      (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  }

  void consRelease(dfx_constraints *c)
  {
    if (!c || c->refs.fetch_sub(1) != 1)
      return;
    dfx_ctx *ctx = c->ctx;
    delete c;
    ctxRelease(ctx);
  }

  template <class T>
  void applyCellwiseNoP2C(dfx_op *op, View<T> X, View<T> Y)
  {
    DFX_REQUIRE(op->kind == OP_CELLWISE, "a cell-wise operator is required");
    DFX_REQUIRE(X.N == op->nDofs && Y.N == op->nDofs, "apply: N does not match the operator");
    DFX_REQUIRE(X.B == Y.B, "apply: X and Y have different widths");
    applyCellwise<T>(op, X, Y, nullptr, nullptr, true);
  }
  template void applyCellwiseNoP2C<double>(dfx_op *, View<double>, View<double>);

  // sum_e Q_e^T diag(A_e) (no constraint step): each cell's diagonal to its
  // map slots, then added dof by dof in cell order
  void cellDiagonal(dfx_op *op, double *out)
  {
    DFX_REQUIRE(op->kind == OP_CELLWISE, "a cell-wise operator is required");
    hipStream_t s = op->ctx->stream;
    DevBuf slots(op->sumNe * sizeof(double));
    k::cellDiag(s, op, slots.as<double>());
    View<double> o{out, op->nDofs, 1, 1};
    k::fill<double>(s, o, 0.0);
    k::scatterAdd<double>(s, op->dofSlotOffsets.as<u32>(), op->dofSlots.as<u32>(), op->nDofs,
                          View<const double>{slots.as<double>(), (u32)op->sumNe, 1, 1}, o);
    DFX_HIP(hipStreamSynchronize(s));
  }

  template void applyOp<double>(dfx_op *, View<double>, View<double>);
  template void applyOp<float>(dfx_op *, View<float>, View<float>);
  template View<double> view<double>(dfx_mv *);
  template View<float> view<float>(dfx_mv *);
} // namespace dfx

// ----------------------------------------------- entry point without allocation
// Scratch of dfx_op_apply_noalloc: the enriched partial rows of a cell-wise
// apply, DFX_NOALLOC_SCRATCH_BYTES per GPU, held by every GPU the library
// runs on until the process ends.  Host code never uses the array's name as
// an address: each device's address comes from hipGetSymbolAddress, taken
// once per device with that device current when a context is created there.
namespace
{
  static __device__ __align__(256) unsigned char scratch[DFX_NOALLOC_SCRATCH_BYTES];

  std::mutex &scratchMutex()
  {
    static std::mutex *m = new std::mutex; // never destroyed
    return *m;
  }
  std::vector<void *> &scratchPtrs()
  {
    static std::vector<void *> *v = new std::vector<void *>;
    return *v;
  }

  // the current device is `device`
  void scratchSetUp(int device)
  {
    std::lock_guard<std::mutex> lk(scratchMutex());
    auto &v = scratchPtrs();
    if ((size_t)device < v.size() && v[device])
      return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Error(DFX_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                               hipGetErrorString(e));
    if ((size_t)device >= v.size())
      v.resize(device + 1, nullptr);
    v[device] = ptr;
  }

  void *scratchOn(int device)
  {
    std::lock_guard<std::mutex> lk(scratchMutex());
    auto &v = scratchPtrs();
    if ((size_t)device >= v.size() || !v[device])
      throw Error(DFX_E_STATE, "no-allocation scratch not set up on this device");
    return v[device];
  }
} // namespace

// =================================================================== ABI
extern "C"
{
  // ------------------------------------------------------------- context
  int dfx_ctx_create(int device, dfx_ctx **out)
  {
    if (!out)
      return DFX_E_ARG;
    *out = nullptr;
    auto ctx = new dfx_ctx();
    int rc = guard(ctx, [&] {
      int n = 0;
      DFX_HIP(hipGetDeviceCount(&n));
      DFX_REQUIRE(device >= 0 && device < n, "no such HIP device");
      DFX_HIP(hipSetDevice(device));
      ctx->device = device;
      scratchSetUp(device);
      DFX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
      DFX_HIP(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
      DFX_HIP(hipEventCreateWithFlags(&ctx->forkEv, hipEventDisableTiming));
      DFX_HIP(hipEventCreateWithFlags(&ctx->joinEv, hipEventDisableTiming));
      {
        CaptureLock lk(captureMutex());
        if (rocblas_create_handle(&ctx->blas) != rocblas_status_success)
          throw Error(DFX_E_HIP, "rocblas_create_handle failed");
        rocblas_set_stream(ctx->blas, ctx->stream);
      }
      ctx->scratchSmall.alloc(1 << 20);
      return 0;
    });
    if (rc != 0)
      {
        last_global_error() = ctx->err;
        delete ctx;
        return rc;
      }
    *out = ctx;
    return 0;
  }

  int dfx_ctx_destroy(dfx_ctx *ctx)
  {
    if (!ctx)
      return DFX_E_ARG;
    ctxRelease(ctx);
    return 0;
  }

  const char *dfx_ctx_last_error(const dfx_ctx *ctx)
  {
    return ctx ? ctx->err.c_str() : last_global_error().c_str();
  }

  int dfx_ctx_synchronize(dfx_ctx *ctx)
  {
    return guard(ctx, [&] {
      DFX_HIP(hipStreamSynchronize(ctx->stream));
      return 0;
    });
  }

  int dfx_ctx_stream(dfx_ctx *ctx, void **stream)
  {
    if (!ctx || !stream)
      return DFX_E_ARG;
    *stream = (void *)ctx->stream;
    return 0;
  }

  int dfx_ctx_graph_stats(const dfx_ctx *ctx, uint64_t *captures, uint64_t *replays,
                          uint32_t *cached)
  {
    if (!ctx)
      return DFX_E_ARG;
    if (captures)
      *captures = ctx->graphCaptures;
    if (replays)
      *replays = ctx->graphReplays;
    if (cached)
      {
        uint32_t n = 0;
        for (const auto &g : ctx->graphs)
          n += g.second.exec != nullptr;
        *cached = n;
      }
    return 0;
  }

  // -------------------------------------------------------- multivectors
  int dfx_mv_create(dfx_ctx *ctx, uint32_t N, uint32_t B, int dtype, dfx_mv **out)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(out, "null output");
      DFX_REQUIRE(dtype == DFX_F64 || dtype == DFX_F32, "bad dtype");
      auto m   = new dfx_mv();
      m->ctx   = ctx;
      m->N     = N;
      m->B     = B;
      m->dtype = dtype;
      m->owns  = true;
      if (m->bytes())
        {
          hipError_t e = hipMalloc(&m->d, m->bytes());
          if (e != hipSuccess)
            {
              delete m;
              throw Error(e == hipErrorOutOfMemory ? DFX_E_OOM : DFX_E_HIP,
                          std::string("hipMalloc: ") + hipGetErrorString(e));
            }
          DFX_HIP(hipMemsetAsync(m->d, 0, m->bytes(), ctx->stream));
          DFX_HIP(hipStreamSynchronize(ctx->stream));
        }
      ctxRetain(ctx);
      *out = m;
      return 0;
    });
  }

  int dfx_mv_wrap(dfx_ctx *ctx, void *ptr, uint32_t N, uint32_t B, int dtype, dfx_mv **out)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(out && (ptr || !N || !B), "null argument");
      DFX_REQUIRE(dtype == DFX_F64 || dtype == DFX_F32, "bad dtype");
      auto m   = new dfx_mv();
      m->ctx   = ctx;
      m->N     = N;
      m->B     = B;
      m->dtype = dtype;
      m->d     = ptr;
      m->owns  = false;
      ctxRetain(ctx);
      *out     = m;
      return 0;
    });
  }

  int dfx_mv_destroy(dfx_mv *m)
  {
    if (!m)
      return DFX_E_ARG;
    if (m->owns && m->d)
      (void)hipFree(m->d);
    dfx_ctx *ctx = m->ctx;
    delete m;
    ctxRelease(ctx);
    return 0;
  }

  int dfx_mv_upload(dfx_mv *m, const void *host)
  {
    if (!m)
      return DFX_E_ARG;
    return guard(m->ctx, [&] {
      DFX_REQUIRE(host || !m->bytes(), "null host buffer");
      if (m->bytes())
        {
          DFX_HIP(hipMemcpyAsync(m->d, host, m->bytes(), hipMemcpyHostToDevice, m->ctx->stream));
          DFX_HIP(hipStreamSynchronize(m->ctx->stream));
        }
      return 0;
    });
  }

  int dfx_mv_download(const dfx_mv *m, void *host)
  {
    if (!m)
      return DFX_E_ARG;
    return guard(m->ctx, [&] {
      DFX_REQUIRE(host || !m->bytes(), "null host buffer");
      if (m->bytes())
        {
          DFX_HIP(hipMemcpyAsync(host, m->d, m->bytes(), hipMemcpyDeviceToHost, m->ctx->stream));
          DFX_HIP(hipStreamSynchronize(m->ctx->stream));
        }
      return 0;
    });
  }

  int dfx_mv_device_ptr(dfx_mv *m, void **ptr)
  {
    if (!m || !ptr)
      return DFX_E_ARG;
    *ptr = m->d;
    return 0;
  }

  int dfx_mv_set(dfx_mv *m, double v)
  {
    if (!m)
      return DFX_E_ARG;
    return guard(m->ctx, [&] {
      if (m->dtype == DFX_F64)
        k::fill<double>(m->ctx->stream, view<double>(m), v);
      else
        k::fill<float>(m->ctx->stream, view<float>(m), v);
      return 0;
    });
  }

  int dfx_mv_fill_random(dfx_mv *m, uint64_t seed, double a, double b)
  {
    if (!m)
      return DFX_E_ARG;
    return guard(m->ctx, [&] {
      if (m->dtype == DFX_F64)
        k::fillRandom<double>(m->ctx->stream, view<double>(m), seed, a, b);
      else
        k::fillRandom<float>(m->ctx->stream, view<float>(m), seed, a, b);
      return 0;
    });
  }

  int dfx_mv_copy(dfx_mv *dst, const dfx_mv *src)
  {
    if (!dst || !src)
      return DFX_E_ARG;
    return guard(dst->ctx, [&] {
      DFX_REQUIRE(dst->N == src->N && dst->B == src->B, "copy: shape mismatch");
      hipStream_t s = dst->ctx->stream;
      if (dst->dtype == DFX_F64 && src->dtype == DFX_F64)
        k::convert<double, double>(s, view<double>(dst), cview<double>(src));
      else if (dst->dtype == DFX_F64)
        k::convert<double, float>(s, view<double>(dst), cview<float>(src));
      else if (src->dtype == DFX_F64)
        k::convert<float, double>(s, view<float>(dst), cview<double>(src));
      else
        k::convert<float, float>(s, view<float>(dst), cview<float>(src));
      return 0;
    });
  }

  int dfx_mv_axpby(dfx_mv *Z, double a, const dfx_mv *X, double b, const dfx_mv *Y)
  {
    if (!Z || !X || !Y)
      return DFX_E_ARG;
    return guard(Z->ctx, [&] {
      DFX_REQUIRE(Z->N == X->N && Z->N == Y->N && Z->B == X->B && Z->B == Y->B,
                  "axpby: shape mismatch");
      DFX_REQUIRE(Z->dtype == X->dtype && Z->dtype == Y->dtype, "axpby: dtype mismatch");
      if (Z->dtype == DFX_F64)
        k::axpby<double>(Z->ctx->stream, view<double>(Z), a, cview<double>(X), b, cview<double>(Y));
      else
        k::axpby<float>(Z->ctx->stream, view<float>(Z), a, cview<float>(X), b, cview<float>(Y));
      return 0;
    });
  }

  int dfx_mv_axpby_blocked(dfx_mv *Z, double a1, const double *a, const dfx_mv *X,
                           double b1, const double *b, const dfx_mv *Y)
  {
    if (!Z || !X || !Y || !a || !b)
      return DFX_E_ARG;
    return guard(Z->ctx, [&] {
      DFX_REQUIRE(Z->N == X->N && Z->N == Y->N && Z->B == X->B && Z->B == Y->B,
                  "axpbyBlocked: shape mismatch");
      DFX_REQUIRE(Z->dtype == X->dtype && Z->dtype == Y->dtype, "axpbyBlocked: dtype mismatch");
      dfx_ctx *ctx = Z->ctx;
      double  *dv  = ctx->scratchSmall.as<double>();
      DFX_REQUIRE((size_t)2 * Z->B * sizeof(double) <= ctx->scratchSmall.bytes, "B too large");
      std::vector<double> h(2 * (size_t)Z->B);
      std::copy(a, a + Z->B, h.begin());
      std::copy(b, b + Z->B, h.begin() + Z->B);
      DFX_HIP(hipMemcpyAsync(dv, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
      if (Z->dtype == DFX_F64)
        k::axpbyBlocked<double>(ctx->stream, view<double>(Z), a1, dv, cview<double>(X), b1,
                                dv + Z->B, cview<double>(Y));
      else
        k::axpbyBlocked<float>(ctx->stream, view<float>(Z), a1, dv, cview<float>(X), b1,
                               dv + Z->B, cview<float>(Y));
      DFX_HIP(hipStreamSynchronize(ctx->stream));
      return 0;
    });
  }

  int dfx_mv_dots(const dfx_mv *X, const dfx_mv *Y, double *out)
  {
    if (!X || !Y || !out)
      return DFX_E_ARG;
    return guard(X->ctx, [&] {
      DFX_REQUIRE(X->N == Y->N && X->B == Y->B && X->dtype == Y->dtype, "dots: mismatch");
      colDotsHost(X->ctx, X, Y, out);
      return 0;
    });
  }

  int dfx_mv_l2norms(const dfx_mv *X, double *out)
  {
    if (!X || !out)
      return DFX_E_ARG;
    return guard(X->ctx, [&] {
      colDotsHost(X->ctx, X, X, out);
      for (u32 j = 0; j < X->B; ++j)
        out[j] = std::sqrt(out[j]);
      return 0;
    });
  }

  // --------------------------------------------------------- constraints
  int dfx_constraints_create(dfx_ctx *ctx, uint32_t nDofs, uint32_t nRows,
                             const uint32_t *rows, const uint32_t *rowSizes,
                             const uint32_t *rowOffsets, const uint32_t *cols,
                             const double *weights, const double *inhom,
                             dfx_constraints **out)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(out, "null output");
      DFX_REQUIRE(!nRows || (rows && rowSizes && rowOffsets && inhom), "null constraint arrays");
      auto c   = std::make_unique<dfx_constraints>();
      c->ctx   = ctx;
      c->nDofs = nDofs;
      c->nRows = nRows;
      u64 nEnt = 0;
      for (u32 i = 0; i < nRows; ++i)
        nEnt += rowSizes[i];
      DFX_REQUIRE(nEnt < (1ull << 32), "too many constraint entries");
      DFX_REQUIRE(!nEnt || (cols && weights), "null constraint column arrays");
      c->nEntries = (u32)nEnt;
      std::vector<char> isRow(nDofs, 0);
      c->hRows.assign(rows, rows + nRows);
      c->hInhom.assign(inhom, inhom + nRows);
      c->homogeneous = std::all_of(c->hInhom.begin(), c->hInhom.end(),
                                   [](double v) { return v == 0.0; });
      c->homDirichlet = nEnt == 0 && std::all_of(c->hInhom.begin(), c->hInhom.end(),
                                                 [](double v) { return v == 0.0; });
      c->hRowOffsets.resize(nRows + 1);
      c->hCols.resize(nEnt);
      c->hWeights.resize(nEnt);
      // entries are re-laid contiguously in row order (the reference's
      // rowOffsets may point anywhere into cols / weights)
      u32 pos = 0;
      for (u32 i = 0; i < nRows; ++i)
        {
          DFX_REQUIRE(rows[i] < nDofs, "constraint row out of range");
          DFX_REQUIRE(!isRow[rows[i]], "constraint row repeated");
          isRow[rows[i]]     = 1;
          c->hRowOffsets[i]  = pos;
          for (u32 j = 0; j < rowSizes[i]; ++j)
            {
              const u32 src = rowOffsets[i] + j;
              DFX_REQUIRE(cols[src] < nDofs, "constraint column out of range");
              c->hCols[pos]    = cols[src];
              c->hWeights[pos] = weights[src];
              ++pos;
            }
        }
      c->hRowOffsets[nRows] = pos;
      for (u32 e = 0; e < pos; ++e)
        DFX_REQUIRE(!isRow[c->hCols[e]],
                    "constraint chains (a parent that is itself constrained) are not closed");
      // parent-major transpose, children in row order
      std::vector<u32> idx(pos);
      std::iota(idx.begin(), idx.end(), 0u);
      std::stable_sort(idx.begin(), idx.end(),
                       [&](u32 a, u32 b) { return c->hCols[a] < c->hCols[b]; });
      std::vector<u32> entryRow(pos);
      for (u32 i = 0; i < nRows; ++i)
        for (u32 e = c->hRowOffsets[i]; e < c->hRowOffsets[i + 1]; ++e)
          entryRow[e] = i;
      std::vector<u32>    rowDofs(pos);
      std::vector<double> pw(pos);
      c->hParentRows.resize(pos);
      c->hParentEntries.resize(pos);
      for (u32 q = 0; q < pos; ++q)
        {
          const u32 e = idx[q];
          if (q == 0 || c->hCols[e] != c->hCols[idx[q - 1]])
            {
              c->hParents.push_back(c->hCols[e]);
              c->hParentOffsets.push_back(q);
            }
          c->hParentRows[q]    = entryRow[e];
          c->hParentEntries[q] = e;
          rowDofs[q]           = c->hRows[entryRow[e]];
          pw[q]                = c->hWeights[e];
        }
      c->hParentOffsets.push_back(pos);
      c->nParents = (u32)c->hParents.size();
      hipStream_t s = ctx->stream;
      upload(c->rows, c->hRows, s);
      upload(c->rowOffsets, c->hRowOffsets, s);
      upload(c->cols, c->hCols, s);
      upload(c->weights, c->hWeights, s);
      upload(c->inhom, c->hInhom, s);
      upload(c->parents, c->hParents, s);
      upload(c->parentOffsets, c->hParentOffsets, s);
      upload(c->parentRowDofs, rowDofs, s);
      upload(c->parentWeights, pw, s);
      ctxRetain(ctx);
      *out = c.release();
      return 0;
    });
  }

  int dfx_constraints_destroy(dfx_constraints *c)
  {
    if (!c)
      return DFX_E_ARG;
    consRelease(c);
    return 0;
  }

  int dfx_constraints_distribute_parent_to_child(dfx_constraints *c, dfx_mv *X)
  {
    if (!c || !X)
      return DFX_E_ARG;
    return guard(c->ctx, [&] {
      DFX_REQUIRE(X->N == c->nDofs, "constraints: N mismatch");
      if (X->dtype == DFX_F64)
        k::p2c<double>(c->ctx->stream, c, view<double>(X));
      else
        k::p2c<float>(c->ctx->stream, c, view<float>(X));
      return 0;
    });
  }

  int dfx_constraints_distribute_child_to_parent(dfx_constraints *c, dfx_mv *Y)
  {
    if (!c || !Y)
      return DFX_E_ARG;
    return guard(c->ctx, [&] {
      DFX_REQUIRE(Y->N == c->nDofs, "constraints: N mismatch");
      if (Y->dtype == DFX_F64)
        k::c2p<double>(c->ctx->stream, c, view<double>(Y));
      else
        k::c2p<float>(c->ctx->stream, c, view<float>(Y));
      return 0;
    });
  }

  int dfx_constraints_set_constrained_nodes(dfx_constraints *c, dfx_mv *X, double v)
  {
    if (!c || !X)
      return DFX_E_ARG;
    return guard(c->ctx, [&] {
      DFX_REQUIRE(X->N == c->nDofs, "constraints: N mismatch");
      if (X->dtype == DFX_F64)
        k::setRows<double>(c->ctx->stream, c, view<double>(X), v);
      else
        k::setRows<float>(c->ctx->stream, c, view<float>(X), v);
      return 0;
    });
  }

  int dfx_constraints_sizes(const dfx_constraints *c, uint32_t *nRows, uint32_t *nEntries,
                            uint32_t *nParents)
  {
    if (!c)
      return DFX_E_ARG;
    if (nRows)
      *nRows = c->nRows;
    if (nEntries)
      *nEntries = c->nEntries;
    if (nParents)
      *nParents = c->nParents;
    return 0;
  }

  int dfx_constraints_export(const dfx_constraints *c, uint32_t *rows, uint32_t *rowOffsets,
                             uint32_t *cols, uint32_t *parents, uint32_t *parentOffsets,
                             uint32_t *parentRows, uint32_t *parentEntries)
  {
    if (!c)
      return DFX_E_ARG;
    return guard(c->ctx, [&] {
      // read back what the device holds (not the host copies)
      hipStream_t s = c->ctx->stream;
      auto down = [&](uint32_t *dst, const DevBuf &b, size_t n) {
        if (dst && n)
          DFX_HIP(hipMemcpyAsync(dst, b.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      };
      down(rows, c->rows, c->nRows);
      down(rowOffsets, c->rowOffsets, c->nRows + 1);
      down(cols, c->cols, c->nEntries);
      down(parents, c->parents, c->nParents);
      down(parentOffsets, c->parentOffsets, c->nParents + 1);
      DFX_HIP(hipStreamSynchronize(s));
      // parentRows / parentEntries: the device holds row dofs; export indices
      if (parentRows)
        std::copy(c->hParentRows.begin(), c->hParentRows.end(), parentRows);
      if (parentEntries)
        std::copy(c->hParentEntries.begin(), c->hParentEntries.end(), parentEntries);
      return 0;
    });
  }

  // ----------------------------------------------------------- operators
  int dfx_op_create_cellwise(dfx_ctx *ctx, uint32_t nCells, const uint32_t *cellNdofs,
                             const uint32_t *cellDofMap, const double *cellMatrices,
                             uint32_t nDofs, uint32_t nClassical, dfx_constraints *cons,
                             uint32_t flags, dfx_op **out)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(out && cellNdofs && cellDofMap && cellMatrices, "null argument");
      std::unique_ptr<dfx_op> op(new dfx_op());
      buildCellwise(ctx, op.get(), nCells, cellNdofs, cellDofMap, nDofs, nClassical, cons,
                    flags);
      uploadCellMatrices(op.get(), cellMatrices, /*device=*/false);
      ctxRetain(ctx);
      *out = op.release();
      return 0;
    });
  }

  int dfx_op_update_cells(dfx_op *op, const double *cellMatrices)
  {
    if (!op || !cellMatrices)
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      DFX_REQUIRE(op->kind == OP_CELLWISE, "not a cell-wise operator");
      uploadCellMatrices(op, cellMatrices, false);
      return 0;
    });
  }

  int dfx_op_create_diagonal(dfx_ctx *ctx, uint32_t nDofs, const double *diag,
                             uint32_t nEnriched, const double *enrBlock,
                             dfx_constraints *cons, dfx_op **out)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(out && diag, "null argument");
      DFX_REQUIRE(nEnriched <= nDofs, "nEnriched > nDofs");
      DFX_REQUIRE(!nEnriched || enrBlock, "enriched block missing");
      DFX_REQUIRE(!cons || cons->nDofs == nDofs, "constraints built for another N");
      std::unique_ptr<dfx_op> op(new dfx_op());
      op->ctx        = ctx;
      op->kind       = OP_DIAGONAL;
      op->nDofs      = nDofs;
      op->nEnriched  = nEnriched;
      op->nClassical = nDofs - nEnriched;
      op->cons       = cons;
      consRetain(cons);
      upload(op->diag, diag, nDofs, ctx->stream);
      if (nEnriched)
        upload(op->enrBlock, enrBlock, (size_t)nEnriched * nEnriched, ctx->stream);
      ctxRetain(ctx);
      *out = op.release();
      return 0;
    });
  }

  int dfx_op_apply(dfx_op *op, dfx_mv *X, dfx_mv *Y)
  {
    if (!op || !X || !Y)
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      DFX_REQUIRE(X->dtype == Y->dtype, "apply: X and Y dtypes differ");
      if (X->dtype == DFX_F64)
        applyOp<double>(op, view<double>(X), view<double>(Y));
      else
        applyOp<float>(op, view<float>(X), view<float>(Y));
      return 0;
    });
  }

  int dfx_op_apply_noalloc(dfx_op *op, dfx_mv *X, dfx_mv *Y)
  {
    if (!op || !X || !Y)
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      DFX_REQUIRE(X->dtype == Y->dtype, "apply: X and Y dtypes differ");
      DFX_REQUIRE(X->N == op->nDofs && Y->N == op->nDofs, "apply: N does not match the operator");
      DFX_REQUIRE(X->B == Y->B, "apply: X and Y have different widths");
      if (op->kind != OP_CELLWISE)
        {
          // the diagonal apply uses no scratch
          if (X->dtype == DFX_F64)
            applyDiagonal<double>(op, view<double>(X), view<double>(Y));
          else
            applyDiagonal<float>(op, view<float>(X), view<float>(Y));
          return 0;
        }
      const size_t sz   = X->dtype == DFX_F64 ? 8 : 4;
      const size_t need = (size_t)op->nEnrSlots * X->B * sz;
      if (need > (size_t)DFX_NOALLOC_SCRATCH_BYTES)
        throw Error(DFX_E_ARG, "dfx_op_apply_noalloc: the enriched partial rows need " +
                                 std::to_string(need) + " bytes, above the " +
                                 std::to_string((size_t)DFX_NOALLOC_SCRATCH_BYTES) +
                                 "-byte scratch (use dfx_op_apply or fewer columns)");
      void *ptr = scratchOn(op->ctx->device);
      if (X->dtype == DFX_F64)
        applyCellwise<double>(op, view<double>(X), view<double>(Y), nullptr, nullptr, false,
                              static_cast<double *>(ptr));
      else
        applyCellwise<float>(op, view<float>(X), view<float>(Y), nullptr, nullptr, false,
                             static_cast<float *>(ptr));
      return 0;
    });
  }

  int dfx_op_apply_after_inverse(dfx_op *A, dfx_op *BInv, dfx_mv *X, dfx_mv *Y)
  {
    if (!A || !BInv || !X || !Y)
      return DFX_E_ARG;
    return guard(A->ctx, [&] {
      DFX_REQUIRE(X->dtype == Y->dtype, "apply: X and Y dtypes differ");
      DFX_REQUIRE(X->N == A->nDofs && Y->N == A->nDofs && BInv->nDofs == A->nDofs,
                  "apply: N does not match the operators");
      DFX_REQUIRE(X->B == Y->B, "apply: X and Y have different widths");
      auto run = [&](auto tag) {
        using T = decltype(tag);
        // scratch1 of the reference (all of BInv X on the unfused path, its
        // enriched rows on the folded one)
        DevBuf &sb = work(A->ctx, "aai.s1", (size_t)X->N * X->B * sizeof(T));
        View<T> S1{sb.as<T>(), X->N, X->B, X->B};
        View<T> Xv = view<T>(X), Yv = view<T>(Y);
        if (!applyInvThenCellwise<T>(BInv, A, Xv, S1, Yv))
          {
            applyOp<T>(BInv, Xv, S1);
            applyOp<T>(A, S1, Yv);
          }
      };
      if (X->dtype == DFX_F64)
        run(double{});
      else
        run(float{});
      return 0;
    });
  }

  int dfx_op_destroy(dfx_op *op)
  {
    if (!op)
      return DFX_E_ARG;
    dfx_ctx         *ctx  = op->ctx;
    dfx_constraints *cons = op->cons;
    if (ctx && ctx->stream)
      (void)hipStreamSynchronize(ctx->stream);
    if (ctx)
      {
        // captured graphs that launch on this operator hold its buffers
        CaptureLock lk(captureMutex());
        for (auto it = ctx->graphs.begin(); it != ctx->graphs.end();)
          {
            const auto &u = it->second.uids;
            if (std::find(u.begin(), u.end(), op->uid) != u.end())
              {
                if (it->second.exec)
                  (void)hipGraphExecDestroy(it->second.exec);
                it = ctx->graphs.erase(it);
              }
            else
              ++it;
          }
      }
    delete op;
    consRelease(cons);
    ctxRelease(ctx);
    return 0;
  }

  int dfx_op_get_info(const dfx_op *op, dfx_op_info *info)
  {
    if (!op || !info)
      return DFX_E_ARG;
    info->nCells      = op->nCells;
    info->nDofs       = op->nDofs;
    info->nClassical  = op->nClassical;
    info->nColors     = op->nColors;
    info->maxCellDofs = op->maxN;
    info->sumNe       = op->sumNe;
    info->sumNe2      = op->sumNe2;
    info->matBytesF64 = op->fragLen64 * 8;
    info->matBytesF32 = op->hasF32 ? op->fragLen32 * 4 : 0;
    return 0;
  }

  int dfx_op_set_profiling(dfx_op *op, int enable)
  {
    if (!op)
      return DFX_E_ARG;
    op->profiling     = enable != 0;
    op->profMask      = enable & 3 ? (enable & 3) : 3;
    const u32 every   = ((u32)enable >> 8) & 0xffffu;
    op->profEvery     = every ? every : 1;
    op->profApplies   = 0;
    return 0;
  }

  int dfx_op_get_profile(dfx_op *op, dfx_op_profile *out, int reset)
  {
    if (!op || !out)
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      for (auto &r : op->prof)
        {
          DFX_HIP(hipEventSynchronize(r.b));
          float ms = 0.f;
          DFX_HIP(hipEventElapsedTime(&ms, r.a, r.b));
          op->profMs += ms;
          op->profFlops += r.flops;
          op->profBytes += r.bytes;
          ++op->profLaunches;
        }
      op->prof.clear();
      for (hipEvent_t e : op->profEvents)
        (void)hipEventDestroy(e);
      op->profEvents.clear();
      out->launches = op->profLaunches;
      out->kernelMs = op->profMs;
      out->flops    = op->profFlops;
      out->bytes    = op->profBytes;
      if (reset)
        {
          op->profLaunches = 0;
          op->profMs = op->profFlops = op->profBytes = 0;
        }
      return 0;
    });
  }

  int dfx_op_autotune(dfx_op *op, int dtype, uint32_t B, uint32_t reps, uint32_t maxOut,
                      uint32_t *classes, uint32_t *widths, uint32_t *nOut)
  {
    if (!op || op->kind != OP_CELLWISE || (dtype != DFX_F64 && dtype != DFX_F32) ||
        (maxOut && (!classes || !widths)))
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      const u32 s = dtype == DFX_F64 ? 8 : 4;
      if (B == 0)
        {
          op->tuned.erase(std::remove_if(op->tuned.begin(), op->tuned.end(),
                                         [&](const dfx_op::Tuned &t) { return t.s == s; }),
                          op->tuned.end());
          if (nOut)
            *nOut = 0;
          return 0;
        }
      const auto r = dtype == DFX_F64 ? autotune<double>(op, B, std::max(reps, 1u))
                                      : autotune<float>(op, B, std::max(reps, 1u));
      for (size_t i = 0; i < r.size() && i < maxOut; ++i)
        {
          classes[i] = r[i].first;
          widths[i]  = (uint32_t)r[i].second;
        }
      if (nOut)
        *nOut = (uint32_t)r.size();
      return 0;
    });
  }

  int dfx_op_export_map(const dfx_op *op, uint32_t *cellNdofs, uint32_t *cellDofMap)
  {
    if (!op || op->kind != OP_CELLWISE)
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      if (cellNdofs)
        std::copy(op->hCellNdofs.begin(), op->hCellNdofs.end(), cellNdofs);
      if (cellDofMap)
        {
          DFX_HIP(hipMemcpyAsync(cellDofMap, op->map.p, op->sumNe * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, op->ctx->stream));
          DFX_HIP(hipStreamSynchronize(op->ctx->stream));
        }
      return 0;
    });
  }

  int dfx_op_gather(dfx_op *op, dfx_mv *X, dfx_mv *Xcell)
  {
    if (!op || !X || !Xcell || op->kind != OP_CELLWISE)
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      DFX_REQUIRE(X->N == op->nDofs && Xcell->N == op->sumNe && X->B == Xcell->B &&
                    X->dtype == Xcell->dtype,
                  "gather: shape mismatch");
      if (X->dtype == DFX_F64)
        k::gather<double>(op->ctx->stream, op->map.as<u32>(), op->sumNe, cview<double>(X),
                          view<double>(Xcell));
      else
        k::gather<float>(op->ctx->stream, op->map.as<u32>(), op->sumNe, cview<float>(X),
                         view<float>(Xcell));
      return 0;
    });
  }

  int dfx_op_scatter_add(dfx_op *op, dfx_mv *Ycell, dfx_mv *Y)
  {
    if (!op || !Y || !Ycell || op->kind != OP_CELLWISE)
      return DFX_E_ARG;
    return guard(op->ctx, [&] {
      DFX_REQUIRE(Y->N == op->nDofs && Ycell->N == op->sumNe && Y->B == Ycell->B &&
                    Y->dtype == Ycell->dtype,
                  "scatter: shape mismatch");
      if (Y->dtype == DFX_F64)
        k::scatterAdd<double>(op->ctx->stream, op->dofSlotOffsets.as<u32>(),
                              op->dofSlots.as<u32>(), op->nDofs, cview<double>(Ycell),
                              view<double>(Y));
      else
        k::scatterAdd<float>(op->ctx->stream, op->dofSlotOffsets.as<u32>(),
                             op->dofSlots.as<u32>(), op->nDofs, cview<float>(Ycell),
                             view<float>(Y));
      return 0;
    });
  }
}
