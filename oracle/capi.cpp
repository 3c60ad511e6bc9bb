// C entry points of the CPU oracle (loaded by tests/ and bench.py's
// cpu_baseline leg through ctypes).  TEST INFRASTRUCTURE ONLY.
#include "oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

using namespace oracle;

namespace
{
  thread_local std::string g_err;

  template <class F>
  int guarded(F &&f)
  {
    try
      {
        return f();
      }
    catch (const std::exception &e)
      {
        g_err = e.what();
        return -1;
      }
  }
} // namespace

extern "C"
{
  const char *orc_error() { return g_err.c_str(); }

  void *orc_build_ex(int p, const int *n, const double *lo, const double *hi,
                     const int *refLo, const int *refHi, int nq, int nqEnr,
                     int nAtoms, const double *atomXYZ, const int *atomZ,
                     double vSmear, int enrich, double rcut, double smooth,
                     int orthogonalize, int keepBasis)
  {
    try
      {
        MeshSpec s;
        s.keepBasis = keepBasis != 0;
        s.p = p;
        for (int d = 0; d < 3; ++d)
          {
            s.n[d]     = n[d];
            s.lo[d]    = lo[d];
            s.hi[d]    = hi[d];
            s.refLo[d] = refLo ? refLo[d] : 0;
            s.refHi[d] = refHi ? refHi[d] : 0;
          }
        s.nq    = nq;
        s.nqEnr = nqEnr;
        for (int a = 0; a < nAtoms; ++a)
          s.atoms.push_back({{atomXYZ[3 * a], atomXYZ[3 * a + 1], atomXYZ[3 * a + 2]},
                             atomZ[a]});
        s.vSmear = vSmear;
        s.enrich = enrich != 0;
        s.rcut   = rcut;
        s.smooth = smooth;
        s.orthogonalize = orthogonalize != 0;
        return new Problem(buildProblem(s));
      }
    catch (const std::exception &e)
      {
        g_err = e.what();
        return nullptr;
      }
  }

  // a classical mesh (no atoms, no refinement) whose H_e carries the affine
  // test potential V(x) = v[0] + v[1..3] . x
  void *orc_build_affine(int p, const int *n, const double *lo, const double *hi, int nq,
                         const double *v)
  {
    try
      {
        MeshSpec s;
        s.p = p;
        for (int d = 0; d < 3; ++d)
          {
            s.n[d]  = n[d];
            s.lo[d] = lo[d];
            s.hi[d] = hi[d];
          }
        s.nq      = nq;
        s.affineV = true;
        for (int i = 0; i < 4; ++i)
          s.vAffine[i] = v[i];
        return new Problem(buildProblem(s));
      }
    catch (const std::exception &e)
      {
        g_err = e.what();
        return nullptr;
      }
  }

  void *orc_build(int p, const int *n, const double *lo, const double *hi,
                  const int *refLo, const int *refHi, int nq, int nqEnr,
                  int nAtoms, const double *atomXYZ, const int *atomZ,
                  double vSmear, int enrich, double rcut, double smooth,
                  int orthogonalize)
  {
    return orc_build_ex(p, n, lo, hi, refLo, refHi, nq, nqEnr, nAtoms, atomXYZ, atomZ,
                        vSmear, enrich, rcut, smooth, orthogonalize, 0);
  }

  // quadrature layout of a problem built with keepBasis: cell e's points are
  // [quadOff[e], quadOff[e + 1]); quadOff (nCells + 1) and JxW (quadOff[nCells])
  int orc_quadrature(void *h, uint64_t *quadOff, double *quadW)
  {
    const Problem &P = *static_cast<Problem *>(h);
    if (P.quadOff.empty())
      return -1;
    if (quadOff)
      std::copy(P.quadOff.begin(), P.quadOff.end(), quadOff);
    if (quadW)
      std::copy(P.quadW.begin(), P.quadW.end(), quadW);
    return 0;
  }

  int orc_density(void *h, const double *X, uint32_t B, const double *occ, double *rho)
  {
    try
      {
        density(*static_cast<Problem *>(h), X, B, occ, rho);
        return 0;
      }
    catch (const std::exception &e)
      {
        g_err = e.what();
        return -1;
      }
  }

  int orc_interpolate(void *h, const double *X, uint32_t B, double *out)
  {
    return guarded([&] {
      interpolate(*static_cast<Problem *>(h), X, B, out);
      return 0;
    });
  }

  // integrateWithBasisValues with the problem's constraints' child->parent
  // step when c2p != 0
  int orc_integrate(void *h, const double *f, uint32_t B, double *out, int c2p)
  {
    return guarded([&] {
      const Problem &P = *static_cast<Problem *>(h);
      integrate(P, f, B, out);
      if (c2p)
        distributeChildToParent(P.cons, out, B);
      return 0;
    });
  }

  // PoissonLinearSolverFunctionFE (src/electrostatics/
  // PoissonLinearSolverFunctionFE.t.cpp) on the problem's K, with the
  // problem's constraints as the homogeneous ones and the same rows carrying
  // the inhomogeneities `inhom` (nRows) as the field constraints:
  //   uD   = P2C_field(0)                               (ctor :188-190)
  //   b    = C2P(integrate(f)) - C2P(sum_e Q_e^T K_e Q_e uD)  (reinit :396-480;
  //          LaplaceOperatorContextFE::apply with X = field, Y = homogeneous
  //          managers, LaplaceOperatorContextFE.t.cpp:397-467)
  //   diag = C2P(sum_e Q_e^T diag K_e), constrained rows 1  (:38-87, :231-245)
  int orc_poisson(void *h, const double *inhom, const double *f, uint32_t B, double *b,
                  double *diag, double *uD)
  {
    return guarded([&] {
      const Problem &P   = *static_cast<Problem *>(h);
      Constraints    fld = P.cons;
      std::copy(inhom, inhom + fld.inhom.size(), fld.inhom.begin());
      const size_t n = (size_t)P.nDofs * B;
      std::fill(uD, uD + n, 0.0);
      distributeParentToChild(fld, uD, B);
      integrate(P, f, B, b);
      distributeChildToParent(P.cons, b, B);
      // K uD cell by cell (X as given: the field's parent->child step is done)
      std::vector<double> y(n, 0.0), xc, yc;
      for (u32 e = 0; e < P.nCells; ++e)
        {
          const u32 ne = P.cellNdofs[e];
          xc.resize((size_t)ne * B);
          yc.resize((size_t)ne * B);
          for (u32 i = 0; i < ne; ++i)
            std::memcpy(&xc[(size_t)i * B], uD + (size_t)P.cellMap[P.cellMapOffset[e] + i] * B,
                        sizeof(double) * B);
          gemm('N', 'N', (int)B, (int)ne, (int)ne, 1.0, xc.data(), (int)B,
               P.K.data() + P.cellMatOffset[e], (int)ne, 0.0, yc.data(), (int)B);
          for (u32 i = 0; i < ne; ++i)
            for (u32 j = 0; j < B; ++j)
              y[(size_t)P.cellMap[P.cellMapOffset[e] + i] * B + j] += yc[(size_t)i * B + j];
        }
      distributeChildToParent(P.cons, y.data(), B);
      for (size_t t = 0; t < n; ++t)
        b[t] = 1.0 * b[t] + (-1.0) * y[t];
      std::fill(diag, diag + P.nDofs, 0.0);
      for (u32 e = 0; e < P.nCells; ++e)
        {
          const u32 ne = P.cellNdofs[e];
          for (u32 i = 0; i < ne; ++i)
            diag[P.cellMap[P.cellMapOffset[e] + i]] +=
              P.K[P.cellMatOffset[e] + (size_t)i * ne + i];
        }
      distributeChildToParent(P.cons, diag, 1);
      setConstrainedNodes(P.cons, diag, 1, 1.0);
      return 0;
    });
  }

  // A problem assembled from caller arrays (maps, constraints, cell matrices
  // and diagonals), used to run the oracle on data produced elsewhere.
  void *orc_from_arrays(uint32_t nCells, uint32_t nDofs, uint32_t nClassical,
                        const uint32_t *cellNdofs, const uint32_t *cellMap,
                        const double *H, const double *M, const double *K,
                        uint32_t nRows, const uint32_t *rows,
                        const uint32_t *rowSizes, const uint32_t *rowOffsets,
                        const uint32_t *cols, const double *weights,
                        const double *inhom, const double *dinv,
                        const double *mgll, int p)
  {
    try
      {
        Problem *P    = new Problem();
        P->p          = p;
        P->nCells     = nCells;
        P->nDofs      = nDofs;
        P->nClassical = nClassical;
        P->nEnriched  = nDofs - nClassical;
        P->cellNdofs.assign(cellNdofs, cellNdofs + nCells);
        P->cellNclassical.assign(nCells, 0);
        P->cellMapOffset.assign(nCells + 1, 0);
        P->cellMatOffset.assign(nCells + 1, 0);
        for (uint32_t e = 0; e < nCells; ++e)
          {
            P->cellMapOffset[e + 1] = P->cellMapOffset[e] + cellNdofs[e];
            P->cellMatOffset[e + 1] =
              P->cellMatOffset[e] + (uint64_t)cellNdofs[e] * cellNdofs[e];
          }
        P->cellMap.assign(cellMap, cellMap + P->cellMapOffset[nCells]);
        const size_t nm = P->cellMatOffset[nCells];
        if (H)
          P->H.assign(H, H + nm);
        if (M)
          P->M.assign(M, M + nm);
        if (K)
          P->K.assign(K, K + nm);
        uint32_t nEntries = 0;
        for (uint32_t i = 0; i < nRows; ++i)
          nEntries += rowSizes[i];
        P->cons.rows.assign(rows, rows + nRows);
        P->cons.rowSizes.assign(rowSizes, rowSizes + nRows);
        P->cons.rowOffsets.assign(rowOffsets, rowOffsets + nRows);
        P->cons.cols.assign(cols, cols + nEntries);
        P->cons.weights.assign(weights, weights + nEntries);
        P->cons.inhom.assign(inhom, inhom + nRows);
        if (dinv)
          P->dinv.assign(dinv, dinv + nDofs);
        else
          P->dinv.assign(nDofs, 0.0);
        if (mgll)
          P->mgll.assign(mgll, mgll + nDofs);
        else
          P->mgll.assign(nDofs, 0.0);
        return P;
      }
    catch (const std::exception &e)
      {
        g_err = e.what();
        return nullptr;
      }
  }

  void orc_free(void *h) { delete static_cast<Problem *>(h); }

  // sizes: nCells, nDofs, nClassical, nEnriched, mapLen, matLen, nRows, nEntries
  void orc_sizes(void *h, uint64_t *out)
  {
    const Problem &P = *static_cast<Problem *>(h);
    out[0]           = P.nCells;
    out[1]           = P.nDofs;
    out[2]           = P.nClassical;
    out[3]           = P.nEnriched;
    out[4]           = P.cellMap.size();
    out[5]           = P.cellMatOffset.empty() ? 0 : P.cellMatOffset.back();
    out[6]           = P.cons.rows.size();
    out[7]           = P.cons.cols.size();
  }

  // which: 0 cellNdofs, 1 cellMap, 2 rows, 3 rowSizes, 4 rowOffsets, 5 cols,
  //        6 isHanging (as u32), 7 isDirichlet (as u32), 8 cellLevel
  int orc_get_u32(void *h, int which, uint32_t *out)
  {
    const Problem &P = *static_cast<Problem *>(h);
    auto cp = [&](const std::vector<uint32_t> &v) {
      std::memcpy(out, v.data(), v.size() * sizeof(uint32_t));
      return 0;
    };
    switch (which)
      {
        case 0: return cp(P.cellNdofs);
        case 1: return cp(P.cellMap);
        case 2: return cp(P.cons.rows);
        case 3: return cp(P.cons.rowSizes);
        case 4: return cp(P.cons.rowOffsets);
        case 5: return cp(P.cons.cols);
        case 6:
          for (size_t i = 0; i < P.isHanging.size(); ++i)
            out[i] = (uint32_t)P.isHanging[i];
          return 0;
        case 7:
          for (size_t i = 0; i < P.isDirichlet.size(); ++i)
            out[i] = (uint32_t)P.isDirichlet[i];
          return 0;
        case 8:
          for (size_t i = 0; i < P.cellLevel.size(); ++i)
            out[i] = (uint32_t)P.cellLevel[i];
          return 0;
      }
    return -1;
  }

  // which: 0 H, 1 M, 2 K, 3 weights, 4 inhom, 5 dinv, 6 mgll, 7 nodeX,
  //        8 enrMinv, 9 cellLo, 10 cellH
  int orc_get_f64(void *h, int which, double *out)
  {
    const Problem &P = *static_cast<Problem *>(h);
    auto cp = [&](const std::vector<double> &v) {
      std::memcpy(out, v.data(), v.size() * sizeof(double));
      return 0;
    };
    switch (which)
      {
        case 0: return cp(P.H);
        case 1: return cp(P.M);
        case 2: return cp(P.K);
        case 3: return cp(P.cons.weights);
        case 4: return cp(P.cons.inhom);
        case 5: return cp(P.dinv);
        case 6: return cp(P.mgll);
        case 7: return cp(P.nodeX);
        case 8: return cp(P.enrMinv);
        case 9: return cp(P.cellLo);
        case 10: return cp(P.cellH);
        case 11: return cp(P.alpha);
      }
    return -1;
  }

  int orc_p2c(void *h, double *X, uint32_t B)
  {
    return guarded([&] {
      distributeParentToChild(static_cast<Problem *>(h)->cons, X, B);
      return 0;
    });
  }
  int orc_c2p(void *h, double *Y, uint32_t B)
  {
    return guarded([&] {
      distributeChildToParent(static_cast<Problem *>(h)->cons, Y, B);
      return 0;
    });
  }
  int orc_gather(void *h, const double *x, uint32_t B, double *xCell)
  {
    return guarded([&] {
      gatherCells(*static_cast<Problem *>(h), x, B, xCell);
      return 0;
    });
  }
  int orc_scatter_add(void *h, const double *yCell, uint32_t B, double *y)
  {
    return guarded([&] {
      scatterAddCells(*static_cast<Problem *>(h), yCell, B, y);
      return 0;
    });
  }
  int orc_apply(void *h, int kind, double *X, double *Y, uint32_t B)
  {
    return guarded([&] {
      applyOp(*static_cast<Problem *>(h), kind, X, Y, B);
      return 0;
    });
  }
  int orc_gemm(char ta, char tb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb, double beta,
               double *C, int ldc)
  {
    gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return 0;
  }
  int orc_syev(int n, double *A, double *w, double *Z)
  {
    return syevJacobi(n, A, n, w, Z, n);
  }
  int orc_potrf(int n, double *A) { return potrfLower(n, A, n); }
  int orc_trtri(int n, double *A) { return trtriLower(n, A, n); }
  int orc_gauss(int n, double *x, double *w)
  {
    std::vector<double> xv, wv;
    gaussLegendre(n, xv, wv);
    std::memcpy(x, xv.data(), n * sizeof(double));
    std::memcpy(w, wv.data(), n * sizeof(double));
    return 0;
  }
  int orc_gll(int n, double *x, double *w)
  {
    std::vector<double> xv, wv;
    gaussLobatto(n, xv, wv);
    std::memcpy(x, xv.data(), n * sizeof(double));
    std::memcpy(w, wv.data(), n * sizeof(double));
    return 0;
  }
  void orc_rng_uniform(uint64_t seed, double a, double b, uint64_t n, double *out)
  {
    Rng r(seed);
    for (uint64_t i = 0; i < n; ++i)
      out[i] = r.uniform(a, b);
  }
  int orc_enrichment(const double *atom, int Z, int k, double rcut,
                     double smooth, const double *x, double *val, double *grad)
  {
    Atom a{{atom[0], atom[1], atom[2]}, Z};
    enrichmentValueGrad(a, k, rcut, smooth, x, *val, grad);
    return 0;
  }

  int orc_chebyshev_filter(void *h, double *X, uint32_t B, int m, double a,
                           double b, double c, double *Y)
  {
    return guarded([&] {
      chebyshevFilter(*static_cast<Problem *>(h), X, B, m, a, b, c, Y);
      return 0;
    });
  }
  int orc_residual_filter(void *h, const double *lam, double *X, uint32_t B,
                          int m, double a, double b, double c, double *Y)
  {
    return guarded([&] {
      residualChebyshevFilter(*static_cast<Problem *>(h), lam, X, B, m, a, b, c, Y);
      return 0;
    });
  }
  int orc_cholgs(void *h, double *X, uint32_t n, double *Xo)
  {
    return guarded([&] { return choleskyGramSchmidt(*static_cast<Problem *>(h), X, n, Xo); });
  }
  int orc_rr(void *h, double *X, uint32_t n, double *eig, double *Xo)
  {
    return guarded([&] { return rayleighRitz(*static_cast<Problem *>(h), X, n, eig, Xo); });
  }
  int orc_chfsi(void *h, double a, double b, double c, int m, int residual,
                uint32_t batch, double *eig, double *guess, double *X, uint32_t n)
  {
    return guarded([&] {
      return chfsiSolve(*static_cast<Problem *>(h), a, b, c, m, residual != 0,
                        batch, eig, guess, X, n);
    });
  }
  int orc_lanczos(void *h, const double *guess, int maxKrylov, double *out3)
  {
    return guarded([&] {
      return lanczosExtreme(*static_cast<Problem *>(h), guess, maxKrylov, &out3[0],
                            &out3[1], &out3[2]);
    });
  }
  int orc_residual_norms(void *h, const double *X, const double *eig, uint32_t n,
                         double *res)
  {
    return guarded([&] {
      eigenResidualNorms(*static_cast<Problem *>(h), X, eig, n, res);
      return 0;
    });
  }
  int orc_cg(void *h, int opKind, const double *diagInv, const double *b,
             double *x, uint32_t nRhs, double absTol, double relTol,
             double divTol, uint32_t maxIter, uint32_t *iters)
  {
    return guarded([&] {
      return cgSolve(*static_cast<Problem *>(h), opKind, diagInv, b, x, nRhs,
                     absTol, relTol, divTol, maxIter, iters);
    });
  }
}
