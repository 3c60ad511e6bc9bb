// C ABI of libdfx, part 4: the Poisson linear-solver function.
//
// electrostatics::PoissonLinearSolverFunctionFE (src/electrostatics/
// PoissonLinearSolverFunctionFE.t.cpp:38-660) on one GPU:
//  * Jacobi diagonal: sum_e Q_e^T diag(K_e), child rows distributed to their
//    parents, constrained rows set to 1 (getDiagonal :38-87 and the
//    constructor :231-245); the preconditioner applies its inverse
//    (PreconditionerJacobi.t.cpp:44-80);
//  * right-hand side b = sum of the integrals of the input quadrature values
//    against the basis, with the homogeneous constraints' C2P
//    (integrateWithBasisValues, reinit :396-460), minus the stiffness apply
//    to the Dirichlet lift: the field constraints' parent->child step on a
//    zero vector gives u_D, K is applied to it cell by cell, and the
//    homogeneous constraints' C2P follows (d_AxContextNHDB->apply, :476-480);
//  * solution = x + u_D with the field constraints' parent->child step, which
//    also becomes the next initial guess (getSolution :594-621).
// The reference skips re-integrating a right-hand side whose quadrature
// values changed by at most 1e-12 (:408-455); here every reinit integrates.
#include <algorithm>

#include "api_util.hpp"
#include "kernels.hpp"

using namespace dfx;

struct dfx_poisson
{
  dfx_ctx         *ctx   = nullptr;
  dfx_problem     *prob  = nullptr;
  dfx_op          *K     = nullptr; // Ax context (homogeneous constraints)
  dfx_constraints *field = nullptr; // Dirichlet values in the inhomogeneities
  dfx_op          *pc    = nullptr; // Jacobi (owned) or NULL (identity)
  u32              N = 0, nComp = 0;
  int              pcType = DFX_PC_JACOBI;
  DevBuf           b, x, initial, uD, diagInv;
};

namespace
{
  View<double> bufView(DevBuf &b, u32 N, u32 B)
  {
    return View<double>{b.as<double>(), N, B, B};
  }

  // u_D = P2C_field(0); b = integrate(f) - C2P_hom(K u_D)
  void poissonReinit(dfx_poisson *p, dfx_mv *rhsQuad)
  {
    hipStream_t s = p->ctx->stream;
    const u32   N = p->N, B = p->nComp;
    View<double> uD = bufView(p->uD, N, B), b = bufView(p->b, N, B);
    k::fill<double>(s, uD, 0.0);
    k::p2c<double>(s, p->field, uD);
    problemIntegrate(p->prob, p->K->cons,
                     View<const double>{static_cast<const double *>(rhsQuad->d), rhsQuad->N,
                                        rhsQuad->B, rhsQuad->B},
                     b);
    DevBuf       tmpU((size_t)N * B * sizeof(double)), tmpY((size_t)N * B * sizeof(double));
    View<double> u = bufView(tmpU, N, B), y = bufView(tmpY, N, B);
    DFX_HIP(hipMemcpyAsync(u.p, uD.p, (size_t)N * B * sizeof(double), hipMemcpyDeviceToDevice,
                           s));
    applyCellwiseNoP2C<double>(p->K, u, y);
    // b = 1 b + (-1) rhsNHDB
    k::axpby<double>(s, b, 1.0, View<const double>{b.p, N, B, B}, -1.0,
                     View<const double>{y.p, N, B, B});
    DFX_HIP(hipStreamSynchronize(s));
  }

  void copyOut(dfx_ctx *ctx, DevBuf &src, dfx_mv *dst, u32 N, u32 B)
  {
    DFX_REQUIRE(dst && dst->dtype == DFX_F64 && dst->N == N && dst->B == B,
                "an FP64 nDofs x nComponents multivector is expected");
    DFX_HIP(hipMemcpyAsync(dst->d, src.p, (size_t)N * B * sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream));
    DFX_HIP(hipStreamSynchronize(ctx->stream));
  }
} // namespace

extern "C"
{
  int dfx_poisson_create(dfx_problem *prob, dfx_op *K, dfx_constraints *field,
                         dfx_mv *rhsQuad, int pcType, dfx_poisson **out)
  {
    if (!prob || !K || !field || !rhsQuad || !out)
      return DFX_E_ARG;
    dfx_ctx *ctx = K->ctx;
    return guard(ctx, [&] {
      DFX_REQUIRE(K->kind == OP_CELLWISE, "the stiffness operator must be cell-wise");
      DFX_REQUIRE(field->nDofs == K->nDofs, "field constraints built for another N");
      DFX_REQUIRE(pcType == DFX_PC_JACOBI || pcType == DFX_PC_NONE, "Unknown PreConditionerType");
      DFX_REQUIRE(rhsQuad->dtype == DFX_F64 && rhsQuad->B > 0,
                  "the right-hand side must be FP64 quadrature values");
      std::unique_ptr<dfx_poisson> p(new dfx_poisson());
      p->ctx    = ctx;
      p->prob   = prob;
      p->K      = K;
      p->field  = field;
      p->N      = K->nDofs;
      p->nComp  = rhsQuad->B;
      p->pcType = pcType;
      const size_t bytes = (size_t)p->N * p->nComp * sizeof(double);
      p->b.alloc(bytes);
      p->x.alloc(bytes);
      p->initial.alloc(bytes);
      p->uD.alloc(bytes);
      hipStream_t s = ctx->stream;
      k::fill<double>(s, bufView(p->x, p->N, p->nComp), 0.0);
      k::fill<double>(s, bufView(p->initial, p->N, p->nComp), 0.0);
      // Jacobi: diag = C2P(sum_e diag K_e), constrained rows 1, applied as
      // its inverse
      DevBuf diag((size_t)p->N * sizeof(double));
      std::vector<double> h(p->N, 1.0);
      if (pcType == DFX_PC_JACOBI)
        {
          cellDiagonal(K, diag.as<double>());
          View<double> d = bufView(diag, p->N, 1);
          if (K->cons)
            {
              k::c2p<double>(s, K->cons, d);
              k::setRows<double>(s, K->cons, d, 1.0);
            }
          DFX_HIP(hipMemcpyAsync(h.data(), diag.p, h.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
          DFX_HIP(hipStreamSynchronize(s));
          for (double &v : h)
            v = 1.0 / v;
        }
      dfx_op *pc = nullptr;
      int     rc = dfx_op_create_diagonal(ctx, p->N, h.data(), 0, nullptr, nullptr, &pc);
      if (rc)
        throw Error(rc, ctx->err);
      p->pc = pc;
      poissonReinit(p.get(), rhsQuad);
      ctxRetain(ctx);
      consRetain(field);
      *out = p.release();
      return 0;
    });
  }

  int dfx_poisson_reinit(dfx_poisson *p, dfx_constraints *field, dfx_mv *rhsQuad)
  {
    if (!p || !rhsQuad)
      return DFX_E_ARG;
    return guard(p->ctx, [&] {
      DFX_REQUIRE(rhsQuad->dtype == DFX_F64 && rhsQuad->B == p->nComp,
                  "input field has different components than that when constructed");
      if (field && field != p->field)
        {
          DFX_REQUIRE(field->nDofs == p->N, "field constraints built for another N");
          consRetain(field);
          consRelease(p->field);
          p->field = field;
        }
      poissonReinit(p, rhsQuad);
      return 0;
    });
  }

  int dfx_poisson_get_ax(dfx_poisson *p, dfx_op **A)
  {
    if (!p || !A)
      return DFX_E_ARG;
    *A = p->K;
    return 0;
  }

  int dfx_poisson_get_pc(dfx_poisson *p, dfx_op **pc)
  {
    if (!p || !pc)
      return DFX_E_ARG;
    *pc = p->pc;
    return 0;
  }

  int dfx_poisson_get_rhs(dfx_poisson *p, dfx_mv *b)
  {
    if (!p)
      return DFX_E_ARG;
    return guard(p->ctx, [&] {
      copyOut(p->ctx, p->b, b, p->N, p->nComp);
      return 0;
    });
  }

  int dfx_poisson_get_initial_guess(dfx_poisson *p, dfx_mv *x)
  {
    if (!p)
      return DFX_E_ARG;
    return guard(p->ctx, [&] {
      copyOut(p->ctx, p->initial, x, p->N, p->nComp);
      return 0;
    });
  }

  int dfx_poisson_set_solution(dfx_poisson *p, const dfx_mv *x)
  {
    if (!p || !x)
      return DFX_E_ARG;
    return guard(p->ctx, [&] {
      DFX_REQUIRE(x->dtype == DFX_F64 && x->N == p->N && x->B == p->nComp,
                  "an FP64 nDofs x nComponents multivector is expected");
      DFX_HIP(hipMemcpyAsync(p->x.p, x->d, (size_t)p->N * p->nComp * sizeof(double),
                             hipMemcpyDeviceToDevice, p->ctx->stream));
      DFX_HIP(hipStreamSynchronize(p->ctx->stream));
      return 0;
    });
  }

  int dfx_poisson_get_solution(dfx_poisson *p, dfx_mv *solution)
  {
    if (!p || !solution)
      return DFX_E_ARG;
    return guard(p->ctx, [&] {
      DFX_REQUIRE(solution->dtype == DFX_F64 && solution->N == p->N &&
                    solution->B == p->nComp,
                  "an FP64 nDofs x nComponents multivector is expected");
      hipStream_t  s   = p->ctx->stream;
      const u32    N   = p->N, B = p->nComp;
      View<double> sol{static_cast<double *>(solution->d), N, B, B};
      // solution = x + u_D, then the field constraints' parent->child step
      k::axpby<double>(s, sol, 1.0, View<const double>{p->x.as<double>(), N, B, B}, 1.0,
                       View<const double>{p->uD.as<double>(), N, B, B});
      k::p2c<double>(s, p->field, sol);
      DFX_HIP(hipMemcpyAsync(p->initial.p, sol.p, (size_t)N * B * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
      DFX_HIP(hipStreamSynchronize(s));
      return 0;
    });
  }

  // CGLinearSolver::solve(linearSolverFunction) (src/linearAlgebra/
  // CGLinearSolver.t.cpp:68-313): x0 = the initial guess, PCG with the Ax and
  // preconditioner contexts on the right-hand side, setSolution(x), then
  // getSolution(solution)
  int dfx_poisson_solve(dfx_poisson *p, uint32_t maxIter, double absTol, double relTol,
                        double divTol, dfx_mv *solution, uint32_t *iterations,
                        double *residualNorms)
  {
    if (!p || !solution)
      return DFX_E_ARG;
    dfx_ctx *ctx = p->ctx;
    int      rc  = 0;
    int      g   = guard(ctx, [&] {
      const u32 N = p->N, B = p->nComp;
      dfx_mv    bm{ctx, N, B, DFX_F64, p->b.p, false};
      dfx_mv    xm{ctx, N, B, DFX_F64, p->x.p, false};
      DFX_HIP(hipMemcpyAsync(p->x.p, p->initial.p, (size_t)N * B * sizeof(double),
                             hipMemcpyDeviceToDevice, ctx->stream));
      rc = dfx_cg_solve(ctx, p->K, p->pc, &bm, &xm, absTol, relTol, divTol, maxIter,
                        iterations, residualNorms);
      if (rc < 0)
        return rc;
      const int r2 = dfx_poisson_get_solution(p, solution);
      return r2 < 0 ? r2 : 0;
    });
    return g != 0 ? g : rc;
  }

  int dfx_poisson_destroy(dfx_poisson *p)
  {
    if (!p)
      return DFX_E_ARG;
    dfx_ctx *ctx = p->ctx;
    if (ctx && ctx->stream)
      (void)hipStreamSynchronize(ctx->stream);
    if (p->pc)
      dfx_op_destroy(p->pc);
    consRelease(p->field);
    delete p;
    ctxRelease(ctx);
    return 0;
  }
}
