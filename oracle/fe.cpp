// CPU oracle, part 1: RNG, quadrature, structured hex mesh with one level of
// local refinement (hanging nodes), FE_Q(p) on GLL support points, cell
// matrices, flattened constraints, and the analytic Slater enrichment.
//
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// What it restates (dft-efe file:line):
//  * cell -> dof maps: src/basis/FEBasisManager.t.cpp:118-147, 228-260; EFE
//    cells list classical dofs first, then enriched ids + nClassical:
//    src/basis/EFEBasisDofHandlerDealii.t.cpp:944-949, 1159-1190.
//  * hanging-node + Dirichlet constraints, flattened with rows ascending:
//    src/basis/CFEBasisDofHandlerDealii.t.cpp:215-229,
//    src/basis/FEBasisManager.t.cpp:181-207,
//    src/basis/CFEConstraintsLocalDealii.t.cpp:290-460.
//  * cell matrices: kinetic 1/2 int grad N_i . grad N_j
//    (src/ksdft/KineticFE.t.cpp:90-100), local potential int N_i V N_j
//    (src/ksdft/ElectrostaticLocalFE.t.cpp:1625-1660 ->
//    src/basis/FEBasisOperations.t.cpp:41-433), summed per cell
//    (src/ksdft/KohnShamOperatorContextFE.t.cpp:1264-1282); consistent mass
//    int N_i N_j (src/basis/CFEOverlapOperatorContext.t.cpp).
//  * GLL-lumped overlap inverse diagonal:
//    src/basis/CFEOverlapInverseOpContextGLL.t.cpp:380-439.
//  * enrichment: real Y_lm without Condon-Shortley phase
//    (src/atoms/SphericalHarmonicFunctions.cpp:25-68, 379-429), smooth cutoff
//    (src/atoms/SmoothCutoffFunctions.cpp:12-59), support radius
//    cutoff + cutoff/smoothness with a bounding-box cell test
//    (src/basis/EnrichmentIdsPartition.t.cpp:273-331).
//
// Documented differences from deal.II (parity unpinned against the reference
// build, which cannot be compiled here): global dof numbering is "first
// appearance" in cell order; the cell-local order is tensor-lexicographic
// (x fastest) instead of deal.II's hierarchical order.
#include "oracle.hpp"

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <map>
#include <stdexcept>

namespace oracle
{
  // ------------------------------------------------------------------ RNG
  u64 Rng::next()
  {
    s += 0x9E3779B97F4A7C15ULL;
    u64 z = s;
    z     = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z     = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double Rng::u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

  // ------------------------------------------------------------ quadrature
  static void legendreP(int n, double x, double &pn, double &pnm1)
  {
    double p0 = 1.0, p1 = x;
    if (n == 0)
      {
        pn   = 1.0;
        pnm1 = 0.0;
        return;
      }
    for (int k = 2; k <= n; ++k)
      {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0        = p1;
        p1        = p2;
      }
    pn   = p1;
    pnm1 = p0;
  }

  // Gauss-Legendre: Newton on P_n (src/quadrature/QuadratureRuleGauss.cpp).
  void gaussLegendre(int n, std::vector<double> &x, std::vector<double> &w)
  {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i)
      {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it)
          {
            double pn, pnm1;
            legendreP(n, z, pn, pnm1);
            double dp = n * (z * pn - pnm1) / (z * z - 1.0);
            double dz = pn / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16)
              break;
          }
        double pn, pnm1;
        legendreP(n, z, pn, pnm1);
        double dp    = n * (z * pn - pnm1) / (z * z - 1.0);
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
      }
  }

  // Gauss-Lobatto-Legendre: endpoints + roots of P'_{n-1}
  // (src/quadrature/QuadratureRuleGLL.cpp); weights 2/(N(N+1) P_N(x)^2).
  void gaussLobatto(int n, std::vector<double> &x, std::vector<double> &w)
  {
    const int N = n - 1;
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i)
      {
        double z = -std::cos(M_PI * i / N);
        if (i > 0 && i < N)
          for (int it = 0; it < 100; ++it)
            {
              double pn, pnm1;
              legendreP(N, z, pn, pnm1);
              // f = (1-z^2) P_N'(z)  ~  N (P_{N-1} - z P_N)
              double f  = N * (pnm1 - z * pn);
              // f' = -N(N+1) P_N
              double fp = -N * (N + 1.0) * pn;
              double dz = f / fp;
              z -= dz;
              if (std::fabs(dz) < 1e-16)
                break;
            }
        x[i] = z;
      }
    x[0] = -1.0;
    x[N] = 1.0;
    if (n % 2 == 1)
      x[N / 2] = 0.0;
    for (int i = 0; i < n; ++i)
      {
        double pn, pnm1;
        legendreP(N, x[i], pn, pnm1);
        w[i] = 2.0 / (N * (N + 1.0) * pn * pn);
      }
  }

  double lagrange(const std::vector<double> &nodes, int k, double x)
  {
    double v = 1.0;
    for (int m = 0; m < (int)nodes.size(); ++m)
      if (m != k)
        v *= (x - nodes[m]) / (nodes[k] - nodes[m]);
    return v;
  }

  double lagrangeDeriv(const std::vector<double> &nodes, int k, double x)
  {
    double s = 0.0;
    for (int j = 0; j < (int)nodes.size(); ++j)
      {
        if (j == k)
          continue;
        double t = 1.0 / (nodes[k] - nodes[j]);
        for (int m = 0; m < (int)nodes.size(); ++m)
          if (m != k && m != j)
            t *= (x - nodes[m]) / (nodes[k] - nodes[m]);
        s += t;
      }
    return s;
  }

  // ------------------------------------------------------------ enrichment
  // smoothCutoffValue / Derivative: src/atoms/SmoothCutoffFunctions.cpp:12-59
  static double f1(double x) { return x <= 0.0 ? 0.0 : std::exp(-1.0 / x); }
  static double f1Der(double x) { return x <= 0.0 ? 0.0 : f1(x) / (x * x); }
  static double f2(double x) { return f1(x) / (f1(x) + f1(1.0 - x)); }
  static double f2Der(double x)
  {
    double a = f1(x) + f1(1.0 - x);
    return (f1Der(x) * f1(1.0 - x) + f1(x) * f1Der(1.0 - x)) / (a * a);
  }
  static void cutoff(double r, double rc, double d, double &v, double &dv)
  {
    double y = 1.0 - d * (r - rc) / rc;
    v        = f2(y);
    dv       = f2Der(y) * (-d / rc);
  }

  // Slater-type radial parts with zeta = Z: 1s e^{-Zr}, 2s r e^{-Zr},
  // 2p r e^{-Zr} (times the real Y_1m below).  Real Y_lm without the
  // Condon-Shortley phase: Y_00 = 1/(2 sqrt(pi)), Y_1,-1 ~ y/r, Y_10 ~ z/r,
  // Y_11 ~ x/r with prefactor sqrt(3/(4 pi))  (SphericalHarmonicFunctions.cpp).
  void enrichmentValueGrad(const Atom &a, int k, double rcut, double smooth,
                           const double *x, double &val, double *grad)
  {
    const double dx[3] = {x[0] - a.r[0], x[1] - a.r[1], x[2] - a.r[2]};
    const double r     = std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    const double Z     = a.Z;
    double       fc, dfc;
    cutoff(r, rcut, smooth, fc, dfc);
    const double ez = std::exp(-Z * r);
    const double y00 = 0.5 / std::sqrt(M_PI);
    const double c1  = std::sqrt(3.0 / (4.0 * M_PI));
    // g(r) radial incl. cutoff, value and derivative
    if (k == 0 || k == 1)
      {
        double g, dg;
        if (k == 0)
          {
            g  = ez;
            dg = -Z * ez;
          }
        else
          {
            g  = r * ez;
            dg = ez - Z * r * ez;
          }
        val             = y00 * g * fc;
        const double dr = y00 * (dg * fc + g * dfc);
        for (int d = 0; d < 3; ++d)
          grad[d] = r > 1e-14 ? dr * dx[d] / r : 0.0;
        return;
      }
    // 2p: c1 * x_i * e^{-Zr} * fc(r)
    const int    comp = (k == 2) ? 1 : (k == 3 ? 2 : 0);
    const double h    = ez * fc;
    const double dh   = -Z * ez * fc + ez * dfc;
    val               = c1 * dx[comp] * h;
    for (int d = 0; d < 3; ++d)
      {
        double g = (r > 1e-14) ? dx[comp] * dh * dx[d] / r : 0.0;
        if (d == comp)
          g += h;
        grad[d] = c1 * g;
      }
  }

  double potentialAt(const std::vector<Atom> &atoms, double smear, const double *x)
  {
    double v = 0.0;
    for (const Atom &a : atoms)
      {
        const double dx = x[0] - a.r[0], dy = x[1] - a.r[1], dz = x[2] - a.r[2];
        const double r  = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (r < 1e-12)
          v += -a.Z * 2.0 / (smear * std::sqrt(M_PI));
        else
          v += -a.Z * std::erf(r / smear) / r;
      }
    return v;
  }

  // ------------------------------------------------------------------ mesh
  namespace
  {
    using Key = std::array<long long, 3>;

    long long keyInt(long long v) { return 2 * v; }
    long long keyFrac(int level, long long idx)
    {
      return 2 * (((long long)(level + 1) << 30) + idx) + 1;
    }
    bool isIntKey(long long k) { return (k & 1) == 0; }

    struct CellRec
    {
      int level;
      int c[3];
    };

    long long key1d(int p, int level, int c, int a)
    {
      const int sh = 1 - level;
      if (a == 0)
        return keyInt((long long)c << sh);
      if (a == p)
        return keyInt((long long)(c + 1) << sh);
      if (level == 0 && 2 * a == p)
        return keyInt(2LL * c + 1);
      return keyFrac(level, (long long)c * p + a);
    }
  } // namespace

  // Enrichment orthogonalization (the OrthoEFE construction,
  // src/basis/EnrichmentClassicalInterfaceSpherical.h:76-97): for each atom,
  // the classical L2 projection of its enrichment functions onto the nodes
  // interior to the enrichment support (every incident cell in the support,
  // unconstrained), M_II alpha = d with d_c = int N_c phi, solved per function
  // by Jacobi-PCG (the CGLinearSolver iteration, stopped at ||r|| < 1e-13 ||d||).
  static void orthogonalizeEnrichment(Problem &P,
                                      const std::vector<std::vector<int>> &cellAtoms,
                                      int nfun)
  {
    const int nl = (P.p + 1) * (P.p + 1) * (P.p + 1);
    P.alpha.assign((size_t)P.nEnriched * P.nClassical, 0.0);
    std::vector<u32> inc(P.nClassical, 0);
    for (u32 e = 0; e < P.nCells; ++e)
      for (int l = 0; l < nl; ++l)
        ++inc[P.cellMap[P.cellMapOffset[e] + l]];
    std::vector<char> constrained(P.nClassical, 0);
    for (u32 r : P.cons.rows)
      constrained[r] = 1;
    const int nAtoms = (int)(P.nEnriched / nfun);
    for (int ai = 0; ai < nAtoms; ++ai)
      {
        std::vector<u32> cells, colStart;
        for (u32 e = 0; e < P.nCells; ++e)
          for (size_t q = 0; q < cellAtoms[e].size(); ++q)
            if (cellAtoms[e][q] == ai)
              {
                cells.push_back(e);
                colStart.push_back((u32)(nl + nfun * q));
              }
        std::vector<u32> loc(P.nClassical, UINT32_MAX), glob, cnt;
        std::vector<u32> lmap(cells.size() * nl);
        for (size_t i = 0; i < cells.size(); ++i)
          for (int l = 0; l < nl; ++l)
            {
              const u32 g = P.cellMap[P.cellMapOffset[cells[i]] + l];
              if (loc[g] == UINT32_MAX)
                {
                  loc[g] = (u32)glob.size();
                  glob.push_back(g);
                  cnt.push_back(0);
                }
              ++cnt[loc[g]];
              lmap[i * nl + l] = loc[g];
            }
        const size_t      n = glob.size();
        std::vector<char> interior(n);
        for (size_t i = 0; i < n; ++i)
          interior[i] = cnt[i] == inc[glob[i]] && !constrained[glob[i]];
        auto Mop = [&](const std::vector<double> &x, std::vector<double> &y) {
          std::fill(y.begin(), y.end(), 0.0);
          for (size_t i = 0; i < cells.size(); ++i)
            {
              const u32     e  = cells[i];
              const u32     ne = P.cellNdofs[e];
              const double *Me = &P.M[P.cellMatOffset[e]];
              for (int r = 0; r < nl; ++r)
                {
                  double acc = 0.0;
                  for (int c = 0; c < nl; ++c)
                    acc += Me[r + (size_t)ne * c] * x[lmap[i * nl + c]];
                  y[lmap[i * nl + r]] += acc;
                }
            }
          for (size_t i = 0; i < n; ++i)
            if (!interior[i])
              y[i] = 0.0;
        };
        std::vector<double> diag(n, 0.0);
        for (size_t i = 0; i < cells.size(); ++i)
          {
            const u32 e = cells[i], ne = P.cellNdofs[e];
            for (int l = 0; l < nl; ++l)
              diag[lmap[i * nl + l]] += P.M[P.cellMatOffset[e] + l + (size_t)ne * l];
          }
        for (int k = 0; k < nfun; ++k)
          {
            std::vector<double> d(n, 0.0), x(n, 0.0), r(n), z(n), p(n), w(n);
            for (size_t i = 0; i < cells.size(); ++i)
              {
                const u32 e = cells[i], ne = P.cellNdofs[e];
                for (int l = 0; l < nl; ++l)
                  d[lmap[i * nl + l]] +=
                    P.M[P.cellMatOffset[e] + l + (size_t)ne * (colStart[i] + k)];
              }
            for (size_t i = 0; i < n; ++i)
              if (!interior[i])
                d[i] = 0.0;
            double bn = 0.0;
            for (double v : d)
              bn += v * v;
            bn = std::sqrt(bn);
            r  = d;
            for (size_t i = 0; i < n; ++i)
              z[i] = interior[i] ? r[i] / diag[i] : r[i];
            p = z;
            double zr = 0.0;
            for (size_t i = 0; i < n; ++i)
              zr += z[i] * r[i];
            for (int it = 0; it < 10000; ++it)
              {
                double rn = 0.0;
                for (double v : r)
                  rn += v * v;
                if (std::sqrt(rn) < 1e-13 * bn)
                  break;
                Mop(p, w);
                double pw = 0.0;
                for (size_t i = 0; i < n; ++i)
                  pw += p[i] * w[i];
                const double al = zr / pw;
                for (size_t i = 0; i < n; ++i)
                  {
                    x[i] += al * p[i];
                    r[i] -= al * w[i];
                  }
                for (size_t i = 0; i < n; ++i)
                  z[i] = interior[i] ? r[i] / diag[i] : r[i];
                double zrn = 0.0;
                for (size_t i = 0; i < n; ++i)
                  zrn += z[i] * r[i];
                const double be = zrn / zr;
                zr              = zrn;
                for (size_t i = 0; i < n; ++i)
                  p[i] = z[i] + be * p[i];
              }
            const size_t g = (size_t)ai * nfun + k;
            for (size_t i = 0; i < n; ++i)
              if (interior[i])
                P.alpha[g * P.nClassical + glob[i]] = x[i];
          }
      }
  }

  Problem buildProblem(const MeshSpec &S)
  {
    Problem P;
    const int p  = S.p;
    const int np = p + 1;
    const int nl = np * np * np;
    if (p < 1 || p > 6)
      throw std::invalid_argument("p must be in 1..6");
    P.p = p;
    std::vector<double> gll, gllw;
    gaussLobatto(np, gll, gllw);

    bool refined = S.refLo[0] < S.refHi[0] && S.refLo[1] < S.refHi[1] &&
                   S.refLo[2] < S.refHi[2];
    auto inRef = [&](int cx, int cy, int cz) {
      return refined && cx >= S.refLo[0] && cx < S.refHi[0] &&
             cy >= S.refLo[1] && cy < S.refHi[1] && cz >= S.refLo[2] &&
             cz < S.refHi[2];
    };

    double h0[3];
    for (int d = 0; d < 3; ++d)
      h0[d] = (S.hi[d] - S.lo[d]) / S.n[d];

    // cell list: coarse lexicographic (x fastest); refined coarse cells are
    // replaced in place by their 8 children (child x fastest)
    std::vector<CellRec> cells;
    for (int cz = 0; cz < S.n[2]; ++cz)
      for (int cy = 0; cy < S.n[1]; ++cy)
        for (int cx = 0; cx < S.n[0]; ++cx)
          {
            if (!inRef(cx, cy, cz))
              cells.push_back({0, {cx, cy, cz}});
            else
              for (int k = 0; k < 2; ++k)
                for (int j = 0; j < 2; ++j)
                  for (int i = 0; i < 2; ++i)
                    cells.push_back({1, {2 * cx + i, 2 * cy + j, 2 * cz + k}});
          }
    P.nCells = (u32)cells.size();

    // classical nodes, numbered by first appearance
    std::map<Key, u32>         nodeId;
    std::vector<Key>           nodeKey;
    std::vector<char>          nodeFine, nodeCoarse;
    std::vector<u32>           classicalMap(P.nCells * (size_t)nl);
    for (u32 e = 0; e < P.nCells; ++e)
      {
        const CellRec &C = cells[e];
        for (int c = 0; c < np; ++c)
          for (int b = 0; b < np; ++b)
            for (int a = 0; a < np; ++a)
              {
                Key k = {key1d(p, C.level, C.c[0], a),
                         key1d(p, C.level, C.c[1], b),
                         key1d(p, C.level, C.c[2], c)};
                auto it = nodeId.find(k);
                u32  id;
                if (it == nodeId.end())
                  {
                    id = (u32)nodeKey.size();
                    nodeId.emplace(k, id);
                    nodeKey.push_back(k);
                    nodeFine.push_back(0);
                    nodeCoarse.push_back(0);
                  }
                else
                  id = it->second;
                (C.level ? nodeFine : nodeCoarse)[id] = 1;
                classicalMap[e * (size_t)nl + a + np * (b + np * c)] = id;
              }
      }
    P.nClassical = (u32)nodeKey.size();

    // node positions from keys (exact per key)
    auto keyPos = [&](int d, long long k) -> double {
      if (isIntKey(k))
        return S.lo[d] + (double)(k / 2) * (0.5 * h0[d]);
      long long v     = (k - 1) / 2;
      int       level = (int)(v >> 30) - 1;
      long long idx   = v & ((1LL << 30) - 1);
      long long c     = idx / p;
      int       a     = (int)(idx % p);
      double    hl    = level ? 0.5 * h0[d] : h0[d];
      return S.lo[d] + hl * (double)c + hl * (gll[a] + 1.0) * 0.5;
    };
    P.nodeX.resize(3 * (size_t)P.nClassical);
    for (u32 i = 0; i < P.nClassical; ++i)
      for (int d = 0; d < 3; ++d)
        P.nodeX[3 * (size_t)i + d] = keyPos(d, nodeKey[i][d]);

    // ------------------------------------------------ constraints
    P.isDirichlet.assign(P.nClassical, 0);
    P.isHanging.assign(P.nClassical, 0);
    for (u32 i = 0; i < P.nClassical; ++i)
      for (int d = 0; d < 3; ++d)
        {
          long long k = nodeKey[i][d];
          if (isIntKey(k) && (k / 2 == 0 || k / 2 == 2LL * S.n[d]))
            P.isDirichlet[i] = 1;
        }

    // coarse (level-0) cell index lookup, for hanging parents
    std::map<std::array<int, 3>, u32> coarseCell;
    for (u32 e = 0; e < P.nCells; ++e)
      if (cells[e].level == 0)
        coarseCell[{cells[e].c[0], cells[e].c[1], cells[e].c[2]}] = e;

    std::vector<std::vector<std::pair<u32, double>>> hang(P.nClassical);
    for (u32 i = 0; i < P.nClassical; ++i)
      {
        if (!nodeFine[i] || P.isDirichlet[i] || nodeCoarse[i])
          continue;
        const Key &k = nodeKey[i];
        // candidate coarse indices per dim containing this node
        std::vector<int> cand[3];
        for (int d = 0; d < 3; ++d)
          {
            long long kk = k[d];
            if (isIntKey(kk))
              {
                long long v = kk / 2;
                if (v % 2 == 0)
                  {
                    if (v / 2 - 1 >= 0)
                      cand[d].push_back((int)(v / 2 - 1));
                    if (v / 2 < S.n[d])
                      cand[d].push_back((int)(v / 2));
                  }
                else
                  cand[d].push_back((int)((v - 1) / 2));
              }
            else
              {
                long long idx = ((kk - 1) / 2) & ((1LL << 30) - 1);
                cand[d].push_back((int)((idx / p) / 2));
              }
          }
        // first unrefined coarse cell in cell order (z, y, x lexicographic)
        bool found = false;
        int  cc[3] = {0, 0, 0};
        for (int z : cand[2])
          {
            for (int y : cand[1])
              {
                for (int x : cand[0])
                  if (coarseCell.count({x, y, z}))
                    {
                      cc[0] = x;
                      cc[1] = y;
                      cc[2] = z;
                      found = true;
                      break;
                    }
                if (found)
                  break;
              }
            if (found)
              break;
          }
        if (!found)
          continue;
        // local coordinate in the coarse cell, per dim
        double xi[3];
        for (int d = 0; d < 3; ++d)
          {
            long long kk = k[d];
            if (isIntKey(kk))
              xi[d] = (double)(kk / 2 - 2LL * cc[d] - 1);
            else
              {
                long long idx = ((kk - 1) / 2) & ((1LL << 30) - 1);
                long long cf  = idx / p;
                int       a   = (int)(idx % p);
                xi[d] = (cf - 2LL * cc[d] == 0) ? (gll[a] - 1.0) * 0.5
                                                : (gll[a] + 1.0) * 0.5;
              }
          }
        const u32 ec = coarseCell[{cc[0], cc[1], cc[2]}];
        P.isHanging[i] = 1;
        std::vector<double> L[3];
        for (int d = 0; d < 3; ++d)
          {
            L[d].resize(np);
            for (int a = 0; a < np; ++a)
              L[d][a] = lagrange(gll, a, xi[d]);
          }
        for (int c = 0; c < np; ++c)
          for (int b = 0; b < np; ++b)
            for (int a = 0; a < np; ++a)
              {
                double w = (L[0][a] * L[1][b]) * L[2][c];
                if (w == 0.0)
                  continue;
                u32 par = classicalMap[ec * (size_t)nl + a + np * (b + np * c)];
                if (P.isDirichlet[par])
                  continue; // closure: homogeneous Dirichlet parent contributes 0
                hang[i].push_back({par, w});
              }
        std::sort(hang[i].begin(), hang[i].end(),
                  [](auto &x, auto &y) { return x.first < y.first; });
      }

    for (u32 i = 0; i < P.nClassical; ++i)
      {
        if (!P.isDirichlet[i] && !P.isHanging[i])
          continue;
        P.cons.rows.push_back(i);
        P.cons.rowOffsets.push_back((u32)P.cons.cols.size());
        P.cons.inhom.push_back(0.0);
        if (P.isDirichlet[i])
          {
            P.cons.rowSizes.push_back(0);
            continue;
          }
        P.cons.rowSizes.push_back((u32)hang[i].size());
        for (auto &pw : hang[i])
          {
            P.cons.cols.push_back(pw.first);
            P.cons.weights.push_back(pw.second);
          }
      }

    // ------------------------------------------------ enrichment ids
    const int nfun   = 5;
    const double rs  = S.rcut + S.rcut / S.smooth;
    P.nEnriched      = S.enrich ? (u32)(S.atoms.size() * nfun) : 0;
    P.nDofs          = P.nClassical + P.nEnriched;
    P.cellLo.resize(3 * (size_t)P.nCells);
    P.cellH.resize(3 * (size_t)P.nCells);
    P.cellLevel.resize(P.nCells);
    std::vector<std::vector<int>> cellAtoms(P.nCells);
    for (u32 e = 0; e < P.nCells; ++e)
      {
        const CellRec &C = cells[e];
        P.cellLevel[e]   = C.level;
        for (int d = 0; d < 3; ++d)
          {
            double hl = C.level ? 0.5 * h0[d] : h0[d];
            P.cellH[3 * (size_t)e + d]  = hl;
            P.cellLo[3 * (size_t)e + d] = S.lo[d] + hl * C.c[d];
          }
        if (!S.enrich)
          continue;
        for (int ai = 0; ai < (int)S.atoms.size(); ++ai)
          {
            bool flag = true;
            for (int d = 0; d < 3; ++d)
              {
                double a = P.cellLo[3 * (size_t)e + d];
                double b = a + P.cellH[3 * (size_t)e + d];
                double c = S.atoms[ai].r[d] - rs, dd = S.atoms[ai].r[d] + rs;
                if ((c < a && dd < a) || (c > b && dd > b))
                  {
                    flag = false;
                    break;
                  }
              }
            if (flag)
              cellAtoms[e].push_back(ai);
          }
      }

    // ------------------------------------------------ maps
    P.cellNdofs.resize(P.nCells);
    P.cellNclassical.assign(P.nCells, (u32)nl);
    P.cellMapOffset.resize(P.nCells + 1);
    P.cellMatOffset.resize(P.nCells + 1);
    P.cellMapOffset[0] = 0;
    P.cellMatOffset[0] = 0;
    for (u32 e = 0; e < P.nCells; ++e)
      {
        u32 ne          = (u32)nl + (u32)(cellAtoms[e].size() * nfun);
        P.cellNdofs[e]  = ne;
        P.cellMapOffset[e + 1] = P.cellMapOffset[e] + ne;
        P.cellMatOffset[e + 1] = P.cellMatOffset[e] + (u64)ne * ne;
        for (int l = 0; l < nl; ++l)
          P.cellMap.push_back(classicalMap[e * (size_t)nl + l]);
        for (int ai : cellAtoms[e])
          for (int k = 0; k < nfun; ++k)
            P.cellMap.push_back(P.nClassical + (u32)(ai * nfun + k));
      }

    // ------------------------------------------------ cell matrices
    P.H.assign(P.cellMatOffset[P.nCells], 0.0);
    P.M.assign(P.cellMatOffset[P.nCells], 0.0);
    P.K.assign(P.cellMatOffset[P.nCells], 0.0);
    const int nqC = S.nq ? S.nq : np;
    const int nqE = S.nqEnr ? S.nqEnr : 2 * np;
    std::vector<double> qxC, qwC, qxE, qwE;
    gaussLegendre(nqC, qxC, qwC);
    gaussLegendre(nqE, qxE, qwE);

    if (S.keepBasis)
      {
        P.basisOff.assign(P.nCells + 1, 0);
        P.quadOff.assign(P.nCells + 1, 0);
        for (u32 e = 0; e < P.nCells; ++e)
          {
            const u64 nq = P.cellNdofs[e] > (u32)nl ? qxE.size() : qxC.size();
            P.quadOff[e + 1]  = P.quadOff[e] + nq * nq * nq;
            P.basisOff[e + 1] = P.basisOff[e] + nq * nq * nq * P.cellNdofs[e];
          }
        P.basisF.assign(P.basisOff[P.nCells], 0.0);
        P.quadW.assign(P.quadOff[P.nCells], 0.0);
      }

    // cell matrices of the cells in `todo`; with `alpha`, the enriched
    // functions are phi_g - sum_l alpha[g][map(l)] N_l (orthogonalized basis)
    auto computeCells = [&](const std::vector<u32> &todo, const std::vector<double> *alpha) {
    std::vector<double> f, gx, gy, gz, V;
    for (u32 e : todo)
      {
        const u32   ne  = P.cellNdofs[e];
        const bool  enr = ne > (u32)nl;
        const auto &qx  = enr ? qxE : qxC;
        const auto &qw  = enr ? qwE : qwC;
        const int   nq  = (int)qx.size();
        const int   nQ  = nq * nq * nq;
        const double *lo = &P.cellLo[3 * (size_t)e];
        const double *h  = &P.cellH[3 * (size_t)e];
        const double detJ = h[0] * h[1] * h[2] / 8.0;
        // tabulate 1-D basis at quadrature points
        std::vector<double> L(np * nq), D(np * nq);
        for (int a = 0; a < np; ++a)
          for (int q = 0; q < nq; ++q)
            {
              L[a * nq + q] = lagrange(gll, a, qx[q]);
              D[a * nq + q] = lagrangeDeriv(gll, a, qx[q]);
            }
        f.assign((size_t)ne * nQ, 0.0);
        gx.assign((size_t)ne * nQ, 0.0);
        gy.assign((size_t)ne * nQ, 0.0);
        gz.assign((size_t)ne * nQ, 0.0);
        V.assign(nQ, 0.0);
        std::vector<double> W(nQ);
        for (int qc = 0; qc < nq; ++qc)
          for (int qb = 0; qb < nq; ++qb)
            for (int qa = 0; qa < nq; ++qa)
              {
                const int q = qa + nq * (qb + nq * qc);
                W[q]        = ((qw[qa] * qw[qb]) * qw[qc]) * detJ;
                double X[3] = {lo[0] + h[0] * (qx[qa] + 1.0) * 0.5,
                               lo[1] + h[1] * (qx[qb] + 1.0) * 0.5,
                               lo[2] + h[2] * (qx[qc] + 1.0) * 0.5};
                V[q]        = S.affineV ? S.vAffine[0] + S.vAffine[1] * X[0] +
                                         S.vAffine[2] * X[1] + S.vAffine[3] * X[2]
                                      : potentialAt(S.atoms, S.vSmear, X);
                for (int c = 0; c < np; ++c)
                  for (int b = 0; b < np; ++b)
                    for (int a = 0; a < np; ++a)
                      {
                        const int    l  = a + np * (b + np * c);
                        const double La = L[a * nq + qa], Lb = L[b * nq + qb],
                                     Lc = L[c * nq + qc];
                        f[(size_t)l * nQ + q]  = (La * Lb) * Lc;
                        gx[(size_t)l * nQ + q] = (D[a * nq + qa] * Lb) * Lc * (2.0 / h[0]);
                        gy[(size_t)l * nQ + q] = (La * D[b * nq + qb]) * Lc * (2.0 / h[1]);
                        gz[(size_t)l * nQ + q] = (La * Lb) * D[c * nq + qc] * (2.0 / h[2]);
                      }
                for (u32 l = nl; l < ne; ++l)
                  {
                    const u32 gid = P.cellMap[P.cellMapOffset[e] + l] - P.nClassical;
                    double    val, grad[3];
                    enrichmentValueGrad(S.atoms[gid / nfun], gid % nfun, S.rcut,
                                        S.smooth, X, val, grad);
                    if (alpha)
                      for (int lc = 0; lc < nl; ++lc)
                        {
                          const double a = (*alpha)[(size_t)gid * P.nClassical +
                                                    P.cellMap[P.cellMapOffset[e] + lc]];
                          if (a == 0.0)
                            continue;
                          val -= a * f[(size_t)lc * nQ + q];
                          grad[0] -= a * gx[(size_t)lc * nQ + q];
                          grad[1] -= a * gy[(size_t)lc * nQ + q];
                          grad[2] -= a * gz[(size_t)lc * nQ + q];
                        }
                    f[(size_t)l * nQ + q]  = val;
                    gx[(size_t)l * nQ + q] = grad[0];
                    gy[(size_t)l * nQ + q] = grad[1];
                    gz[(size_t)l * nQ + q] = grad[2];
                  }
              }
        if (S.keepBasis)
          {
            std::copy(f.begin(), f.end(), P.basisF.begin() + P.basisOff[e]);
            std::copy(W.begin(), W.end(), P.quadW.begin() + P.quadOff[e]);
          }
        double *He = &P.H[P.cellMatOffset[e]];
        double *Me = &P.M[P.cellMatOffset[e]];
        double *Ke = &P.K[P.cellMatOffset[e]];
        for (u32 j = 0; j < ne; ++j)
          for (u32 i = 0; i < ne; ++i)
            {
              double k = 0.0, m = 0.0, v = 0.0;
              const double *fi = &f[(size_t)i * nQ], *fj = &f[(size_t)j * nQ];
              const double *xi = &gx[(size_t)i * nQ], *xj = &gx[(size_t)j * nQ];
              const double *yi = &gy[(size_t)i * nQ], *yj = &gy[(size_t)j * nQ];
              const double *zi = &gz[(size_t)i * nQ], *zj = &gz[(size_t)j * nQ];
              for (int q = 0; q < nQ; ++q)
                {
                  k += (xi[q] * xj[q] + yi[q] * yj[q] + zi[q] * zj[q]) * W[q];
                  m += (fi[q] * fj[q]) * W[q];
                  v += (fi[q] * V[q] * fj[q]) * W[q];
                }
              Ke[i + (size_t)ne * j] = k;
              Me[i + (size_t)ne * j] = m;
              He[i + (size_t)ne * j] = 0.5 * k + v;
            }
      }
    };
    std::vector<u32> allCells(P.nCells);
    for (u32 e = 0; e < P.nCells; ++e)
      allCells[e] = e;
    computeCells(allCells, nullptr);
    if (P.nEnriched && S.orthogonalize)
      {
        orthogonalizeEnrichment(P, cellAtoms, nfun);
        std::vector<u32> enrCells;
        for (u32 e = 0; e < P.nCells; ++e)
          if (P.cellNdofs[e] > (u32)nl)
            enrCells.push_back(e);
        computeCells(enrCells, &P.alpha);
      }

    // ------------------------------------------------ GLL-lumped inverse
    // diagonal of the GLL mass (cell-wise diagonal), assembled, condensed to
    // parents (C2P), constrained entries set to 1, reciprocal, constrained
    // entries set to 0: CFEOverlapInverseOpContextGLL.t.cpp:380-439.
    P.mgll.assign(P.nDofs, 0.0);
    for (u32 e = 0; e < P.nCells; ++e)
      {
        const double *h    = &P.cellH[3 * (size_t)e];
        const double  detJ = h[0] * h[1] * h[2] / 8.0;
        for (int c = 0; c < np; ++c)
          for (int b = 0; b < np; ++b)
            for (int a = 0; a < np; ++a)
              {
                const int l = a + np * (b + np * c);
                P.mgll[P.cellMap[P.cellMapOffset[e] + l]] +=
                  ((gllw[a] * gllw[b]) * gllw[c]) * detJ;
              }
      }
    std::vector<double> dcond(P.mgll);
    distributeChildToParent(P.cons, dcond.data(), 1);
    setConstrainedNodes(P.cons, dcond.data(), 1, 1.0);
    P.dinv.assign(P.nDofs, 0.0);
    for (u32 i = 0; i < P.nClassical; ++i)
      P.dinv[i] = 1.0 / dcond[i];
    for (u32 r : P.cons.rows)
      P.dinv[r] = 0.0;

    // ------------------------------------------------ enriched block inverse
    // (M^ee)^-1 of the assembled consistent-mass enriched block, used by the
    // enriched overlap inverse (OrthoEFEOverlapInverseOpContextGLL.t.cpp:
    // 1182-1305: classical diagonal part + dense enriched-block GEMM).
    if (P.nEnriched)
      {
        const u32 nE = P.nEnriched;
        std::vector<double> Mee((size_t)nE * nE, 0.0);
        for (u32 e = 0; e < P.nCells; ++e)
          {
            const u32 ne = P.cellNdofs[e];
            const double *Me = &P.M[P.cellMatOffset[e]];
            for (u32 j = nl; j < ne; ++j)
              for (u32 i = nl; i < ne; ++i)
                {
                  u32 gi = P.cellMap[P.cellMapOffset[e] + i] - P.nClassical;
                  u32 gj = P.cellMap[P.cellMapOffset[e] + j] - P.nClassical;
                  Mee[gi + (size_t)nE * gj] += Me[i + (size_t)ne * j];
                }
          }
        // inverse via Cholesky: L L^T = Mee, inv = L^-T L^-1
        std::vector<double> Lm(Mee);
        if (potrfLower((int)nE, Lm.data(), (int)nE) != 0)
          throw std::runtime_error("enriched mass block not SPD");
        for (u32 j = 0; j < nE; ++j)
          for (u32 i = 0; i < j; ++i)
            Lm[i + (size_t)nE * j] = 0.0;
        trtriLower((int)nE, Lm.data(), (int)nE);
        P.enrMinv.assign((size_t)nE * nE, 0.0);
        gemm('T', 'N', (int)nE, (int)nE, (int)nE, 1.0, Lm.data(), (int)nE,
             Lm.data(), (int)nE, 0.0, P.enrMinv.data(), (int)nE);
      }
    return P;
  }
} // namespace oracle

namespace oracle
{
  // ksdft::DensityCalculator::computeRho (src/ksdft/DensityCalculator.t.cpp:
  // 288-420) on one batch holding every state: the wavefunctions are
  // interpolated to the quadrature points cell by cell, with no constraint
  // distribution (FEBasisOperations::interpolate, src/basis/
  // FEBasisOperations.t.cpp:996-1277: gather, then C_e = A_e B_e^T), and
  // rho(q) = sum_i 2 |psi_i(q)|^2 f_i summed in state order
  // (computeRhoInBatch, DensityCalculator.t.cpp:80-101).
  void density(const Problem &P, const double *X, u32 B, const double *occ, double *rho)
  {
    if (P.quadOff.empty())
      throw std::runtime_error("density needs the basis values (build with keepBasis)");
    std::vector<double> psi;
    for (u32 e = 0; e < P.nCells; ++e)
      {
        const u32     ne  = P.cellNdofs[e];
        const u64     nQ  = P.quadOff[e + 1] - P.quadOff[e];
        const double *f   = &P.basisF[P.basisOff[e]];
        const u32    *map = &P.cellMap[P.cellMapOffset[e]];
        psi.assign((size_t)nQ * B, 0.0);
        for (u64 q = 0; q < nQ; ++q)
          for (u32 i = 0; i < B; ++i)
            {
              double v = 0.0;
              for (u32 k = 0; k < ne; ++k)
                v += X[(size_t)map[k] * B + i] * f[(size_t)k * nQ + q];
              psi[q * B + i] = v;
            }
        for (u64 q = 0; q < nQ; ++q)
          {
            double b = 0.0;
            for (u32 i = 0; i < B; ++i)
              b += 2.0 * (psi[q * B + i] * psi[q * B + i]) * occ[i];
            rho[P.quadOff[e] + q] = b;
          }
      }
  }

  // FEBasisOperations::interpolate (src/basis/FEBasisOperations.t.cpp:996-1277)
  // with no constraint step: out[(quadOff[e] + q) B + c] = sum_k N_k(x_q)
  // X[map_e[k], c], k in order (the gather and per-cell gemm of density())
  void interpolate(const Problem &P, const double *X, u32 B, double *out)
  {
    if (P.quadOff.empty())
      throw std::runtime_error("interpolate needs the basis values (build with keepBasis)");
    for (u32 e = 0; e < P.nCells; ++e)
      {
        const u32     ne  = P.cellNdofs[e];
        const u64     q0  = P.quadOff[e], nQ = P.quadOff[e + 1] - q0;
        const double *bf  = &P.basisF[P.basisOff[e]];
        const u32    *map = &P.cellMap[P.cellMapOffset[e]];
        for (u64 q = 0; q < nQ; ++q)
          for (u32 c = 0; c < B; ++c)
            {
              double v = 0.0;
              for (u32 k = 0; k < ne; ++k)
                v += X[(size_t)map[k] * B + c] * bf[(size_t)k * nQ + q];
              out[(q0 + q) * B + c] = v;
            }
      }
  }

  // FEBasisOperations::integrateWithBasisValues (src/basis/FEBasisOperations.t.cpp:
  // 1846-2190) before its constraint step: per cell, inpJxW(q, c) = JxW(q) f(q, c)
  // (the khatriRao product, :2043-2059), C(c, k) = sum_q inpJxW(q, c) N_k(x_q)
  // in q order (the per-cell gemm, :2144-2163), added into out in cell order
  // (addCellWiseDataToFieldData, FECellWiseDataOperations.t.cpp:89-151).
  // out is overwritten (setValue(0), :1952).
  void integrate(const Problem &P, const double *f, u32 B, double *out)
  {
    if (P.quadOff.empty())
      throw std::runtime_error("integrate needs the basis values (build with keepBasis)");
    std::fill(out, out + (size_t)P.nDofs * B, 0.0);
    std::vector<double> fw, c;
    for (u32 e = 0; e < P.nCells; ++e)
      {
        const u32     ne  = P.cellNdofs[e];
        const u64     q0  = P.quadOff[e], nQ = P.quadOff[e + 1] - q0;
        const double *bf  = &P.basisF[P.basisOff[e]];
        const u32    *map = &P.cellMap[P.cellMapOffset[e]];
        fw.resize((size_t)nQ * B);
        for (u64 q = 0; q < nQ; ++q)
          for (u32 j = 0; j < B; ++j)
            fw[q * B + j] = P.quadW[q0 + q] * f[(q0 + q) * B + j];
        c.assign((size_t)ne * B, 0.0);
        for (u32 k = 0; k < ne; ++k)
          for (u32 j = 0; j < B; ++j)
            {
              double v = 0.0;
              for (u64 q = 0; q < nQ; ++q)
                v += fw[q * B + j] * bf[(size_t)k * nQ + q];
              c[(size_t)k * B + j] = v;
            }
        for (u32 k = 0; k < ne; ++k)
          for (u32 j = 0; j < B; ++j)
            out[(size_t)map[k] * B + j] += c[(size_t)k * B + j];
      }
  }
} // namespace oracle
