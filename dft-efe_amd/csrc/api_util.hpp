// Helpers shared by the C ABI translation units of libdfx.
#pragma once
#include <cmath>
#include <string>

#include "common.hpp"

namespace dfx
{
  namespace k
  {
    template <class T>
    struct ResFuse;
  }
} // namespace dfx

namespace dfx
{
  inline std::string &last_global_error()
  {
    static thread_local std::string e;
    return e;
  }

  // Run f, translating exceptions into status codes and the context's
  // last-error string.  No exception crosses the C ABI.
  template <class F>
  int guard(dfx_ctx *ctx, F &&f)
  {
    try
      {
        if (ctx)
          {
            ctx->err.clear();
            DFX_HIP(hipSetDevice(ctx->device));
          }
        return f();
      }
    catch (const Error &e)
      {
        (ctx ? ctx->err : last_global_error()) = e.what();
        return e.code;
      }
    catch (const std::bad_alloc &)
      {
        (ctx ? ctx->err : last_global_error()) = "host allocation failed";
        return DFX_E_OOM;
      }
    catch (const std::exception &e)
      {
        (ctx ? ctx->err : last_global_error()) = e.what();
        return DFX_E_ARG;
      }
  }

  // a named scratch buffer of the context, grown on demand
  DevBuf &work(dfx_ctx *ctx, const char *name, size_t bytes);

  // context lifetime: every object created on a context holds a reference
  inline void ctxRetain(dfx_ctx *c)
  {
    if (c)
      c->refs.fetch_add(1);
  }
  void ctxRelease(dfx_ctx *c);
  inline void consRetain(dfx_constraints *c)
  {
    if (c)
      c->refs.fetch_add(1);
  }
  void consRelease(dfx_constraints *c);

  template <class T>
  View<T> view(dfx_mv *m);
  template <class T>
  void applyOp(dfx_op *op, View<T> X, View<T> Y);
  template <class T>
  bool applyInvThenCellwise(dfx_op *Dinv, dfx_op *A, View<T> X, View<T> S1, View<T> Y,
                            const k::ResFuse<T> *rf = nullptr);
  // Y = C2P(sum_e Q_e^T A_e Q_e X) with X used as given (no parent->child step)
  template <class T>
  void applyCellwiseNoP2C(dfx_op *op, View<T> X, View<T> Y);
  // out (device, nDofs) = sum_e Q_e^T diag(A_e), no constraint step
  void cellDiagonal(dfx_op *op, double *out);
  // FEBasisOperations::integrateWithBasisValues on the problem's basis
  // (out = C2P(sum_e Q_e^T N_e^T (JxW o f_e)); cons NULL: no C2P)
  void problemIntegrate(dfx_problem *P, const dfx_constraints *cons, View<const double> f,
                        View<double> out);

  // column dots of two device multivectors -> host (FP64), deterministic
  void colDotsHost(dfx_ctx *ctx, const dfx_mv *X, const dfx_mv *Y, double *out);
  template <class T>
  void colDotsView(dfx_ctx *ctx, View<const T> X, View<const T> Y, double *hostOut);

  void buildCellwise(dfx_ctx *ctx, dfx_op *op, u32 nCells, const u32 *cellNdofs,
                     const u32 *cellDofMap, u32 nDofs, u32 nClassical,
                     dfx_constraints *cons, u32 flags);
  // cellMatrices: raw concatenated column-major blocks, on the host or (if
  // device) already on the GPU
  void uploadCellMatrices(dfx_op *op, const double *cellMatrices, bool device);
} // namespace dfx
