// Compiles include/dfx_dropin.hpp against stand-ins of the reference's
// interface headers (tests/dropin/ref, restating the cited signatures) and
// exercises it.
//
// Without a GPU: the process-wide libdfx context cannot be created and the
// shim reports it as the reference's utils::throwException does (exit 0 when
// the message names libdfx).
// With a GPU: a small problem built through the C ABI (3^3 cells, p = 3, one
// smoothed Coulomb centre) is solved through the shim classes
// (DfxChebyshevFilteredEigenSolver with stored intermediate subspaces,
// DfxPoissonLinearSolverFunction with DfxCGLinearSolver) and the results are
// compared bit for bit with the same calls made directly on the C ABI.
// Prints "dropin ok" on success.
#include <dfx_dropin.hpp>

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace dftefe;
using namespace dftefe::linearAlgebra;

template class dftefe::linearAlgebra::DfxCellOperatorContext<double, double>;
template class dftefe::linearAlgebra::DfxCellOperatorContext<float, float>;
template class dftefe::linearAlgebra::DfxDiagonalOperatorContext<double, double>;
template class dftefe::linearAlgebra::DfxChebyshevFilteredEigenSolver<double, double>;

namespace
{
  int fails = 0;
  void check(bool ok, const char *what)
  {
    if (!ok)
      {
        std::printf("FAIL: %s\n", what);
        ++fails;
      }
  }

  template <class T>
  std::vector<T> get(dfx_problem *P, int which, size_t n)
  {
    std::vector<T> v(n);
    if (n && dfx_problem_get(P, which, v.data()) != 0)
      throw std::runtime_error("dfx_problem_get");
    return v;
  }

  std::vector<double> download(const MultiVector<double, dfxMem> &X)
  {
    std::vector<double> h((size_t)X.locallyOwnedSize() * X.getNumberComponents());
    if (hipMemcpy(h.data(), X.data(), h.size() * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess)
      throw std::runtime_error("hipMemcpy");
    return h;
  }

  int cpuPath()
  {
    try
      {
        dfxContext(0);
      }
    catch (const std::logic_error &e)
      {
        const bool named = std::strstr(e.what(), "libdfx") != nullptr;
        std::printf("no GPU: %s\n", e.what());
        if (named)
          std::printf("dropin ok (no device)\n");
        return named ? 0 : 1;
      }
    std::printf("FAIL: a context was created without a device\n");
    return 1;
  }

  int gpuPath()
  {
    dfx_ctx         *ctx = dfxContext(0);
    dfx_problem_spec s{};
    s.p = 3;
    s.n[0] = s.n[1] = s.n[2] = 3;
    for (int d = 0; d < 3; ++d)
      {
        s.lo[d] = -6.0;
        s.hi[d] = 6.0;
      }
    const double xyz[3] = {0.2, -0.1, 0.3};
    const int    Z[1]   = {1};
    s.nAtoms            = 1;
    s.atomXYZ           = xyz;
    s.atomZ             = Z;
    s.vSmear            = 0.5;
    s.rcut              = 4.0;
    s.smooth            = 2.0;
    dfx_problem *P      = nullptr;
    dfxCheck(ctx, dfx_problem_build(ctx, &s, &P));
    uint64_t sz[8];
    dfx_problem_sizes(P, sz);
    const size_type nCells = (size_type)sz[0], N = (size_type)sz[1], nC = (size_type)sz[2];
    auto nd    = get<size_type>(P, DFX_PROB_CELL_NDOFS, nCells);
    auto map   = get<size_type>(P, DFX_PROB_CELL_MAP, sz[4]);
    auto H     = get<double>(P, DFX_PROB_H, sz[5]);
    auto M     = get<double>(P, DFX_PROB_M, sz[5]);
    auto dinv  = get<double>(P, DFX_PROB_DINV, N);
    auto mgll  = get<double>(P, DFX_PROB_MGLL, N);
    auto rows  = get<uint32_t>(P, DFX_PROB_ROWS, sz[6]);
    auto rsz   = get<uint32_t>(P, DFX_PROB_ROW_SIZES, sz[6]);
    auto roff  = get<uint32_t>(P, DFX_PROB_ROW_OFFSETS, sz[6]);
    auto cols  = get<uint32_t>(P, DFX_PROB_COLS, sz[7]);
    auto wts   = get<double>(P, DFX_PROB_WEIGHTS, sz[7]);
    auto inhom = get<double>(P, DFX_PROB_INHOM, sz[6]);
    dfx_constraints *cons = nullptr;
    dfxCheck(ctx, dfx_constraints_create(ctx, N, (uint32_t)rows.size(), rows.data(), rsz.data(),
                                         roff.data(), cols.data(), wts.data(), inhom.data(),
                                         &cons));
    {
      DfxCellOperatorContext<double, double>     Hop(nd, map, H, N, nC, cons, true);
      DfxCellOperatorContext<double, double>     Mop(nd, map, M, N, nC, cons);
      DfxDiagonalOperatorContext<double, double> Minv(dinv, 0, {}, cons);
      DfxDiagonalOperatorContext<double, double> Mlump(mgll, 0, {}, cons);

      // the subspace guess and the Lanczos vector, constrained, on the device
      const size_type nv = 4;
      MultiVector<double, dfxMem> G(N, nv), L(N, 1), X(N, nv), G2(N, nv), X2(N, nv);
      {
        DfxView<double> g(ctx, G), l(ctx, L);
        dfxCheck(ctx, dfx_mv_fill_random(g.get(), 1, -1.0, 1.0));
        dfxCheck(ctx, dfx_mv_fill_random(l.get(), 2, -1.0, 1.0));
        dfxCheck(ctx, dfx_constraints_distribute_parent_to_child(cons, g.get()));
        dfxCheck(ctx, dfx_constraints_distribute_parent_to_child(cons, l.get()));
        DfxView<double> g2(ctx, G2);
        dfxCheck(ctx, dfx_mv_copy(g2.get(), g.get()));
      }
      double lo = 0, hi = 0, beta = 0;
      {
        DfxView<double> l(ctx, L);
        dfxCheck(ctx, dfx_lanczos_extreme(ctx, Hop.handle(), Mlump.handle(), Minv.handle(),
                                          l.get(), 20, &lo, &hi, &beta));
      }
      const double a = lo, c = hi + beta / 10, b = a + 0.1 * (c - a);

      // through the shim: reference constructor, stored subspaces
      ElpaScalapackManager elpa;
      DfxChebyshevFilteredEigenSolver<double, double> chfsi(
        a, b, c, 8, 1e-14, G, elpa, true, 0, OrthogonalizationType::CHOLESKY_GRAMSCHMIDT, true);
      std::vector<double> eig(nv);
      for (size_type i = 0; i < nv; ++i)
        eig[i] = a + (b - a) * i / (nv - 1);
      std::vector<double> eig2 = eig;
      EigenSolverError    err  = chfsi.solve(Hop, eig, X, true, Mop, Minv);
      check(err.isSuccess && err.err == EigenSolverErrorCode::SUCCESS, "shim ChFSI solve");
      check(chfsi.getFilteredSubspace().getNumberComponents() == nv, "filtered subspace");
      check(chfsi.getOrthogonalizedFilteredSubspace().locallyOwnedSize() == N,
            "orthogonalized filtered subspace");
      // the same solve on the C ABI directly
      {
        dfx_chfsi_params p{};
        p.wantedSpectrumLowerBound   = a;
        p.wantedSpectrumUpperBound   = b;
        p.unWantedSpectrumUpperBound = c;
        p.polynomialDegree           = 8;
        p.illConditionTolerance      = 1e-14;
        p.isResidualChebyshevFilter  = 1;
        DfxView<double> g2(ctx, G2), x2(ctx, X2);
        dfx_chfsi      *cs = nullptr;
        dfxCheck(ctx, dfx_chfsi_create(ctx, &p, g2.get(), &cs));
        dfxCheck(ctx, dfx_chfsi_solve(cs, Hop.handle(), eig2.data(), x2.get(), 1,
                                      Mop.handle(), Minv.handle()));
        dfx_chfsi_destroy(cs);
      }
      check(eig == eig2, "shim eigenvalues equal the C-ABI ones bit for bit");
      check(download(X) == download(X2), "shim eigenvectors equal the C-ABI ones");
      std::printf("shim ChFSI eigenvalues %.12f %.12f %.12f %.12f\n", eig[0], eig[1], eig[2],
                  eig[3]);
      // the other orthogonalization types are refused
      bool refused = false;
      try
        {
          DfxChebyshevFilteredEigenSolver<double, double> bad(
            a, b, c, 8, 1e-14, G, elpa, true, 0, OrthogonalizationType::MULTIPASS_LOWDIN);
        }
      catch (const std::logic_error &e)
        {
          refused = std::strstr(e.what(), "Orthogonalization type not present") != nullptr;
        }
      check(refused, "MULTIPASS_LOWDIN refused with the reference's message");
      // a mismatched eigenvalue array is an error, not a host overflow
      bool sized = false;
      try
        {
          std::vector<double> shortEig(2);
          chfsi.solve(Hop, shortEig, X, true, Mop, Minv);
        }
      catch (const std::logic_error &)
        {
          sized = true;
        }
      check(sized, "eigenValues size checked");

      // Poisson: the function object and the CG solver through the shim
      dfx_op *Kop = nullptr;
      dfxCheck(ctx, dfx_problem_create_operators(P, 0, nullptr, nullptr, &Kop, nullptr,
                                                 nullptr, nullptr));
      std::vector<uint64_t> qoff(nCells + 1);
      dfx_problem_get(P, DFX_PROB_QUAD_OFFSETS, qoff.data());
      const uint32_t nQ   = (uint32_t)qoff.back();
      dfx_mv        *rhsQ = nullptr;
      dfxCheck(ctx, dfx_mv_create(ctx, nQ, 2, DFX_F64, &rhsQ));
      dfxCheck(ctx, dfx_mv_fill_random(rhsQ, 7, 0.0, 1.0));
      MultiVector<double, dfxMem> rhs(N, 2), init(N, 2), sol(N, 2), sol2(N, 2);
      DfxPoissonLinearSolverFunction fn(P, Kop, cons, rhsQ, DFX_PC_JACOBI, rhs, init, 0);
      DfxCGLinearSolver              cg(2000, 1e-30, 1e-10, 1e10);
      LinearSolverError              lerr = cg.solve(fn);
      check(lerr.isSuccess, "shim CG solve");
      fn.getSolution(sol);
      {
        dfx_poisson *p2 = nullptr;
        dfxCheck(ctx, dfx_poisson_create(P, Kop, cons, rhsQ, DFX_PC_JACOBI, &p2));
        DfxView<double> s2(ctx, sol2);
        uint32_t        it2 = 0;
        double          res2[2];
        dfxCheck(ctx, dfx_poisson_solve(p2, 2000, 1e-30, 1e-10, 1e10, s2.get(), &it2, res2));
        check(it2 == cg.iterations(), "same CG iteration count");
        dfx_poisson_destroy(p2);
      }
      check(download(sol) == download(sol2), "shim Poisson solution equals the C-ABI one");
      std::printf("shim CG: %u iterations\n", cg.iterations());
      dfx_mv_destroy(rhsQ);
      dfx_op_destroy(Kop);
    }
    dfx_constraints_destroy(cons);
    dfx_problem_destroy(P);
    if (fails)
      return 1;
    std::printf("dropin ok\n");
    return 0;
  }
} // namespace

int main()
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return cpuPath();
  try
    {
      return gpuPath();
    }
  catch (const std::exception &e)
    {
      std::printf("FAIL: %s\n", e.what());
      return 1;
    }
}
