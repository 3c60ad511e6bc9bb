// Problem builder of libdfx: the structured-mesh FE setup that feeds the hot
// path (maps, Dirichlet constraints, cell matrices), built on the GPU.
//
// Host: lattice-indexed node numbering (first appearance over cells, cell-local
// tensor-lexicographic order), Dirichlet rows, enrichment supports.
// GPU:  basis values / gradients at Gauss points (classical tensor products and
// analytic Slater x real Y_lm x smooth-cutoff enrichment), then the cell
// matrices as batched GEMMs  M_e = F^T W F,  K_e = G^T W G,  H_e = K_e / 2 +
// F^T (W V) F  (FEBasisOperations::computeFEMatrices, src/basis/
// FEBasisOperations.t.cpp:41-433; KineticFE.t.cpp:90-100;
// ElectrostaticLocalFE.t.cpp:1625-1660).
// Enrichment orthogonalization (the OrthoEFE construction,
// src/basis/EnrichmentClassicalInterfaceSpherical.h:76-97): per atom, the
// classical L2 projection of its enrichment functions on the nodes interior
// to the enrichment support, M_II alpha = d, solved with this library's own
// cell-wise mass operator and PCG; the enriched basis is phi - sum alpha N.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <numeric>

#include "api_util.hpp"
#include "atom.hpp"
#include "kernels.hpp"

using namespace dfx;

namespace
{
  constexpr int NFUN = 5; // analytic set: 1s, 2s, 2p_y, 2p_z, 2p_x

  // ----------------------------------------------------------- quadrature
  void legendre(int n, double x, double &p, double &dp)
  {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k)
      {
        const double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0              = p1;
        p1              = p2;
      }
    p  = n == 0 ? 1.0 : p1;
    dp = n == 0 ? 0.0 : n * (x * p1 - p0) / (x * x - 1.0);
  }

  void gaussRule(int n, std::vector<double> &x, std::vector<double> &w)
  {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i)
      {
        double z = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 200; ++it)
          {
            double p, dp;
            legendre(n, z, p, dp);
            const double dz = p / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16)
              break;
          }
        double p, dp;
        legendre(n, z, p, dp);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
      }
  }

  // GLL points: +-1 and the roots of P'_{n-1} (Newton on (1 - x^2) P'_{n-1})
  void gllRule(int n, std::vector<double> &x, std::vector<double> &w)
  {
    const int N = n - 1;
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i)
      {
        double z = -std::cos(M_PI * i / N);
        if (i > 0 && i < N)
          for (int it = 0; it < 200; ++it)
            {
              double p, dp, pm, dpm;
              legendre(N, z, p, dp);
              legendre(N - 1, z, pm, dpm);
              // (1-z^2) P_N' = N (P_{N-1} - z P_N); derivative -N(N+1) P_N
              const double f = N * (pm - z * p), fp = -N * (N + 1.0) * p;
              const double dz = f / fp;
              z -= dz;
              if (std::fabs(dz) < 1e-16)
                break;
            }
        x[i] = (i == 0) ? -1.0 : (i == N ? 1.0 : z);
      }
    if (n % 2 == 1)
      x[N / 2] = 0.0;
    for (int i = 0; i < n; ++i)
      {
        double p, dp;
        legendre(N, x[i], p, dp);
        w[i] = 2.0 / (N * (N + 1.0) * p * p);
      }
  }

  double lag(const std::vector<double> &nd, int k, double x)
  {
    double v = 1.0;
    for (int m = 0; m < (int)nd.size(); ++m)
      if (m != k)
        v *= (x - nd[m]) / (nd[k] - nd[m]);
    return v;
  }
  double lagD(const std::vector<double> &nd, int k, double x)
  {
    double s = 0.0;
    for (int j = 0; j < (int)nd.size(); ++j)
      {
        if (j == k)
          continue;
        double t = 1.0 / (nd[k] - nd[j]);
        for (int m = 0; m < (int)nd.size(); ++m)
          if (m != k && m != j)
            t *= (x - nd[m]) / (nd[k] - nd[m]);
        s += t;
      }
    return s;
  }

  // ----------------------------------------------------------- device
  struct TabCell
  {
    double lo[3], h[3];
    u32    n;       // functions of the cell
    u32    nq;      // Gauss points per direction
    u32    tabOff;  // offset of this rule's 1-D tables
    u32    mapOff;  // into the (device) cell map
    u64    bufOff;  // into the tabulation buffers (elements)
  };

  struct AtomDev
  {
    double r[3];
    int    Z;
    int    pad;
  };

  // the local potential of one atom: the smeared Coulomb -Z erf(r / s) / r,
  // or (pseudo set) a pseudopotential's mixed local part Y_00 (spline(r) up to
  // the last knot, tail / r beyond)
  struct PotDev
  {
    double r[3];
    int    Z;
    int    pseudo;
    u32    splOff, splN;
    double tail, cl;
  };

  __device__ double cutoffDev(double r, double rc, double d, double &dv)
  {
    auto f1 = [](double x) { return x <= 0.0 ? 0.0 : exp(-1.0 / x); };
    auto f1d = [&](double x) { return x <= 0.0 ? 0.0 : f1(x) / (x * x); };
    const double y = 1.0 - d * (r - rc) / rc;
    const double a = f1(y), b = f1(1.0 - y);
    const double v = a / (a + b);
    const double s = a + b;
    dv = ((f1d(y) * b + a * f1d(1.0 - y)) / (s * s)) * (-d / rc);
    return v;
  }

  // value and gradient of enrichment function k of atom A at x
  __device__ void enrichDev(const AtomDev &A, int k, double rc, double sm,
                            const double *x, double &val, double *g)
  {
    const double dx[3] = {x[0] - A.r[0], x[1] - A.r[1], x[2] - A.r[2]};
    const double r     = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    const double Z     = A.Z;
    double       dfc;
    const double fc  = cutoffDev(r, rc, sm, dfc);
    const double ez  = exp(-Z * r);
    const double y00 = 0.28209479177387814;  // 1 / (2 sqrt(pi))
    const double c1  = 0.48860251190291992;  // sqrt(3 / (4 pi))
    if (k < 2)
      {
        const double gr  = k == 0 ? ez : r * ez;
        const double dgr = k == 0 ? -Z * ez : ez - Z * r * ez;
        val              = y00 * gr * fc;
        const double dr  = y00 * (dgr * fc + gr * dfc);
        for (int d = 0; d < 3; ++d)
          g[d] = r > 1e-14 ? dr * dx[d] / r : 0.0;
        return;
      }
    const int    comp = k == 2 ? 1 : (k == 3 ? 2 : 0);
    const double hh = ez * fc, dh = -Z * ez * fc + ez * dfc;
    val             = c1 * dx[comp] * hh;
    for (int d = 0; d < 3; ++d)
      {
        double gg = r > 1e-14 ? dx[comp] * dh * dx[d] / r : 0.0;
        if (d == comp)
          gg += hh;
        g[d] = c1 * gg;
      }
  }

  // one enriched basis function: analytic Slater function `kind` (< NFUN)
  // of a nucleus of charge Z, or (kind == NFUN) R(r) Y_lm f_c(r) with R the
  // spline at spl + splOff (knots, y, b, c, d: splN values each)
  struct EnrFun
  {
    double r[3];
    double rc, sm;
    int    Z, kind, l, m;
    u32    splOff, splN;
  };

  __device__ void enrichFun(const EnrFun &E, const double *spl, const double *x, double &val,
                            double *g)
  {
    if (E.kind < NFUN)
      {
        const AtomDev A{{E.r[0], E.r[1], E.r[2]}, E.Z, 0};
        enrichDev(A, E.kind, E.rc, E.sm, x, val, g);
        return;
      }
    const double *b = spl + E.splOff;
    const SplineDev S{b, b + E.splN, b + 2 * (u64)E.splN, b + 3 * (u64)E.splN,
                      b + 4 * (u64)E.splN, E.splN};
    const double dx[3] = {x[0] - E.r[0], x[1] - E.r[1], x[2] - E.r[2]};
    radialYlm(S, E.l, E.m, E.rc, E.sm, dx, val, g);
  }

  // One block per cell; threads over Gauss points.  Writes, column-major in
  // (point, function):  F, FW = W F, FV = W V F  (nQ x n) and G, GW (3 nQ x n);
  // valuesOnly: F alone (no potential, weights or gradients are formed; the
  // other pointers are not touched)
  __global__ void tabulateK(const TabCell *cells, const u32 *map, u32 nClassical,
                            const double *tabL, const double *tabD, const double *tabX,
                            const double *tabW, int np, const PotDev *atoms, u32 nAtoms,
                            double smear, const EnrFun *funs, const double *spl,
                            const double *alpha, bool valuesOnly,
                            double *F, double *FW, double *FV, double *G, double *GW)
  {
    const TabCell C  = cells[blockIdx.x];
    const u32     nq = C.nq, nQ = nq * nq * nq, n = C.n;
    const int     ncl = np * np * np;
    const double  detJ = C.h[0] * C.h[1] * C.h[2] / 8.0;
    const double *L = tabL + C.tabOff, *D = tabD + C.tabOff;
    const double *X = tabX + (C.tabOff / np), *Wt = tabW + (C.tabOff / np);
    const u32    *cmap = map + C.mapOff;
    for (u32 q = threadIdx.x; q < nQ; q += blockDim.x)
      {
        const u32 qa = q % nq, qb = (q / nq) % nq, qc = q / (nq * nq);
        const double x[3] = {C.lo[0] + C.h[0] * (X[qa] + 1.0) * 0.5,
                             C.lo[1] + C.h[1] * (X[qb] + 1.0) * 0.5,
                             C.lo[2] + C.h[2] * (X[qc] + 1.0) * 0.5};
        const double w = ((Wt[qa] * Wt[qb]) * Wt[qc]) * detJ;
        double v = 0.0;
        for (u32 a = 0; a < (valuesOnly ? 0u : nAtoms); ++a)
          {
            const double dx = x[0] - atoms[a].r[0], dy = x[1] - atoms[a].r[1],
                         dz = x[2] - atoms[a].r[2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (atoms[a].pseudo)
              {
                // SphericalDataMixed::getValue (src/atoms/SphericalDataMixed.cpp:43-69)
                const PotDev &A  = atoms[a];
                const double *b  = spl + A.splOff;
                const SplineDev S{b, b + A.splN, b + 2 * (u64)A.splN, b + 3 * (u64)A.splN,
                                  b + 4 * (u64)A.splN, A.splN, A.cl};
                double R, dR;
                if (r <= S.x[S.n - 1])
                  splineEval(S, r, R, dR);
                else
                  R = A.tail / r;
                v += 0.28209479177387814 * R;
              }
            else
              v += r < 1e-12 ? -atoms[a].Z * 2.0 / (smear * 1.7724538509055160)
                             : -atoms[a].Z * erf(r / smear) / r;
          }
        auto put = [&](u32 j, double f, double gx, double gy, double gz) {
          const u64 o = C.bufOff;
          F[o + q + (u64)nQ * j]  = f;
          if (valuesOnly)
            return;
          FW[o + q + (u64)nQ * j] = w * f;
          FV[o + q + (u64)nQ * j] = (w * v) * f;
          const u64 og = 3 * C.bufOff + (u64)3 * nQ * j;
          G[og + q]            = gx;
          G[og + nQ + q]       = gy;
          G[og + 2 * nQ + q]   = gz;
          GW[og + q]           = w * gx;
          GW[og + nQ + q]      = w * gy;
          GW[og + 2 * nQ + q]  = w * gz;
        };
        const double sx = 2.0 / C.h[0], sy = 2.0 / C.h[1], sz = 2.0 / C.h[2];
        for (int c = 0; c < np; ++c)
          for (int b = 0; b < np; ++b)
            for (int a = 0; a < np; ++a)
              {
                const double La = L[a * nq + qa], Lb = L[b * nq + qb], Lc = L[c * nq + qc];
                if (valuesOnly)
                  put(a + np * (b + np * c), (La * Lb) * Lc, 0.0, 0.0, 0.0);
                else
                  put(a + np * (b + np * c), (La * Lb) * Lc,
                      (D[a * nq + qa] * Lb) * Lc * sx, (La * D[b * nq + qb]) * Lc * sy,
                      (La * Lb) * D[c * nq + qc] * sz);
              }
        for (u32 j = ncl; j < n; ++j)
          {
            const u32 g = cmap[j] - nClassical;
            double    f, gr[3];
            enrichFun(funs[g], spl, x, f, gr);
            if (alpha)
              {
                const double *al = alpha + (u64)g * nClassical;
                for (int c = 0; c < np; ++c)
                  for (int b = 0; b < np; ++b)
                    for (int a = 0; a < np; ++a)
                      {
                        const double s = al[cmap[a + np * (b + np * c)]];
                        if (s == 0.0)
                          continue;
                        const double La = L[a * nq + qa], Lb = L[b * nq + qb],
                                     Lc = L[c * nq + qc];
                        f -= s * ((La * Lb) * Lc);
                        if (valuesOnly)
                          continue;
                        gr[0] -= s * ((D[a * nq + qa] * Lb) * Lc * sx);
                        gr[1] -= s * ((La * D[b * nq + qb]) * Lc * sy);
                        gr[2] -= s * ((La * Lb) * D[c * nq + qc] * sz);
                      }
              }
            put(j, f, gr[0], gr[1], gr[2]);
          }
      }
  }

  // H = 0.5 K + V assembled in place: H already holds F^T W V F
  __global__ void addHalfK(double *const *H, double *const *K, const u32 *n, u32 count)
  {
    const u32 e = blockIdx.x;
    if (e >= count)
      return;
    const u64 nn = (u64)n[e] * n[e];
    for (u64 t = threadIdx.x; t < nn; t += blockDim.x)
      H[e][t] += 0.5 * K[e][t];
  }

  // copy a column block: out[e-slot rows, cols] from raw cell matrices
  __global__ void extractClassicalBlock(const double *raw, const u64 *rawOff,
                                        const u32 *ne, const u32 *cellList, u32 ncl,
                                        double *out)
  {
    const u32 i = blockIdx.x, e = cellList[i];
    const u64 nn = (u64)ncl * ncl;
    for (u64 t = threadIdx.x; t < nn; t += blockDim.x)
      {
        const u32 r = (u32)(t % ncl), c = (u32)(t / ncl);
        out[(u64)i * nn + t] = raw[rawOff[e] + r + (u64)ne[e] * c];
      }
  }

  // rhs slots: out[(i*ncl + r)*nf + k] = raw_e[r, col_k]
  __global__ void extractMixedBlock(const double *raw, const u64 *rawOff, const u32 *ne,
                                    const u32 *cellList, const u32 *colStart, u32 ncl,
                                    u32 nf, double *out)
  {
    const u32 i = blockIdx.x, e = cellList[i];
    for (u32 t = threadIdx.x; t < ncl * nf; t += blockDim.x)
      {
        const u32 r = t / nf, k = t % nf;
        out[(u64)i * ncl * nf + t] =
          raw[rawOff[e] + r + (u64)ne[e] * (colStart[i] + k)];
      }
  }

  // enriched x enriched sub-block of each listed cell's (column-major) mass
  // matrix, packed: out[off[i] + a + m_i * b] = M_e[ncl + a, ncl + b],
  // m_i = n_e - ncl
  __global__ void extractEnrBlockK(const double *M, const u64 *matOff, const u32 *ne,
                                   const u32 *cellList, const u64 *outOff, u32 ncl,
                                   double *out)
  {
    const u32 i = blockIdx.x, e = cellList[i];
    const u32 n = ne[e], m = n - ncl;
    for (u32 t = threadIdx.x; t < m * m; t += blockDim.x)
      {
        const u32 a = t % m, b = t / m;
        out[outOff[i] + t] = M[matOff[e] + (ncl + a) + (u64)n * (ncl + b)];
      }
  }

  // the wavefunction rows of each cell of a batch, B x n column-major per
  // cell (the interpolation GEMM's A operand): Xc[t][k B + i] = X[map_c[k], i]
  __global__ void gatherCellRowsK(const double *X, u32 ldx, u32 B, const u32 *map,
                                  const u32 *mapOff, const u32 *cells, u32 n, double *Xc)
  {
    const u32 t = blockIdx.x, c = cells[t];
    for (u32 e = threadIdx.x; e < n * B; e += blockDim.x)
      {
        const u32 k = e / B, i = e - k * B;
        Xc[(u64)t * n * B + e] = X[(u64)map[mapOff[c] + k] * ldx + i];
      }
  }

  // rho[quadOff[c] + q] = sum_i 2 psi_i(q)^2 occ[i], states in order
  // (DensityCalculator.t.cpp:87-98); psi of cell t at psi[(t nQ + q) B + i]
  __global__ void rhoFromPsiK(const double *psi, u32 B, u32 nQ, const u32 *cells,
                              const u64 *quadOff, const double *occ, double *rho)
  {
    const u32 t = blockIdx.y, c = cells[t];
    for (u32 q = blockIdx.x * blockDim.x + threadIdx.x; q < nQ; q += gridDim.x * blockDim.x)
      {
        const double *p = psi + ((u64)t * nQ + q) * B;
        double        b = 0.0;
        for (u32 i = 0; i < B; ++i)
          b += 2.0 * (p[i] * p[i]) * occ[i];
        rho[quadOff[c] + q] = b;
      }
  }

  __global__ void scatterAlpha(const double *x, const u32 *glob, const u32 *interior,
                               u32 nLocal, u32 g0, u32 nf, u32 nClassical, double *alpha)
  {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < nLocal * nf;
         t += gridDim.x * blockDim.x)
      {
        const u32 i = t / nf, k = t % nf;
        if (interior[i])
          alpha[(u64)(g0 + k) * nClassical + glob[i]] = x[t];
      }
  }
} // namespace

// ------------------------------------------------------------------ problem
struct dfx_problem
{
  dfx_ctx         *ctx = nullptr;
  dfx_problem_spec spec{};
  std::vector<double> atomXYZ;
  std::vector<int>    atomZ;
  u32 p = 0, nCells = 0, nClassical = 0, nEnriched = 0, nDofs = 0;
  std::vector<u32>    cellNdofs, cellMap, mapOff;
  std::vector<u64>    matOff;
  std::vector<double> cellLo, cellH, nodeXYZ;
  std::vector<u32>    rows, rowSizes, rowOffsets, cols;
  std::vector<double> weights, inhom, mgll, dinv, enrM, enrMinv;
  std::vector<std::vector<u32>> atomCells; // enrichment supports
  std::vector<const dfx_atom *> atomData;  // numerical enrichment (empty: analytic)
  std::string                   enrichField;
  std::vector<u32>    funOff;               // atom a: enriched ids funOff[a] .. funOff[a+1]-1
  std::vector<EnrFun> funs;                 // per enriched id
  std::vector<double> splines;              // numerical functions' knots and coefficients
  std::vector<const dfx_atom *> localPot;  // pseudopotential local parts (empty: Coulomb)
  std::vector<PotDev> pots;                 // per atom
  DevBuf H, M, K, alpha;                    // raw (caller-layout) cell matrices
};

namespace
{
  void buildHost(dfx_problem *P)
  {
    const auto &S  = P->spec;
    const u32   p  = S.p, np = p + 1, ncl = np * np * np;
    const u32   nx = S.n[0], ny = S.n[1], nz = S.n[2];
    DFX_REQUIRE(p >= 1 && p <= 6, "p must be in 1..6");
    DFX_REQUIRE(nx && ny && nz, "empty mesh");
    P->p      = p;
    P->nCells = nx * ny * nz;
    const u64 NX = (u64)p * nx + 1, NY = (u64)p * ny + 1, NZ = (u64)p * nz + 1;
    DFX_REQUIRE(NX * NY * NZ < (1ull << 32), "mesh too large for 32-bit dof ids");
    std::vector<u32> lat(NX * NY * NZ, UINT32_MAX);
    std::vector<u32> cmapC((size_t)P->nCells * ncl);
    u32 next = 0;
    double h[3];
    for (int d = 0; d < 3; ++d)
      h[d] = (S.hi[d] - S.lo[d]) / S.n[d];
    P->cellLo.resize(3 * (size_t)P->nCells);
    P->cellH.resize(3 * (size_t)P->nCells);
    u32 e = 0;
    for (u32 cz = 0; cz < nz; ++cz)
      for (u32 cy = 0; cy < ny; ++cy)
        for (u32 cx = 0; cx < nx; ++cx, ++e)
          {
            const u32 cc[3] = {cx, cy, cz};
            for (int d = 0; d < 3; ++d)
              {
                P->cellLo[3 * (size_t)e + d] = S.lo[d] + h[d] * cc[d];
                P->cellH[3 * (size_t)e + d]  = h[d];
              }
            for (u32 c = 0; c < np; ++c)
              for (u32 b = 0; b < np; ++b)
                for (u32 a = 0; a < np; ++a)
                  {
                    const u64 li = (cx * p + a) + NX * ((cy * p + b) + NY * (u64)(cz * p + c));
                    if (lat[li] == UINT32_MAX)
                      lat[li] = next++;
                    cmapC[(size_t)e * ncl + a + np * (b + np * c)] = lat[li];
                  }
          }
    P->nClassical = next;
    {
      // node coordinates: lattice index cx p + a lies at lo + h (cx + (x_a + 1) / 2)
      std::vector<double> gx0, gw0;
      gllRule(np, gx0, gw0);
      P->nodeXYZ.assign(3 * (size_t)P->nClassical, 0.0);
      const u64 NN[3] = {NX, NY, NZ};
      for (u64 k = 0; k < NZ; ++k)
        for (u64 j = 0; j < NY; ++j)
          for (u64 i = 0; i < NX; ++i)
            {
              const u64 ijk[3] = {i, j, k};
              const u32 id     = lat[i + NX * (j + NY * k)];
              for (int d = 0; d < 3; ++d)
                {
                  const u64 c = ijk[d] / p, a = ijk[d] % p;
                  (void)NN;
                  P->nodeXYZ[3 * (size_t)id + d] = S.lo[d] + h[d] * ((double)c + 0.5 * (gx0[a] + 1.0));
                }
            }
    }
    // Dirichlet: every lattice point on the domain boundary, rows ascending
    for (u64 k = 0; k < NZ; ++k)
      for (u64 j = 0; j < NY; ++j)
        for (u64 i = 0; i < NX; ++i)
          if (i == 0 || j == 0 || k == 0 || i == NX - 1 || j == NY - 1 || k == NZ - 1)
            P->rows.push_back(lat[i + NX * (j + NY * k)]);
    std::sort(P->rows.begin(), P->rows.end());
    P->rowSizes.assign(P->rows.size(), 0);
    P->rowOffsets.assign(P->rows.size(), 0);
    P->inhom.assign(P->rows.size(), 0.0);

    // enrichment supports: cells whose box meets the atom's support cube
    // (per atom its functions: the analytic set with the spec's cutoff, or
    // the atom data's functions, each with its own; the support is the
    // largest of the atom's function supports)
    const u32 nA = S.enrich ? S.nAtoms : 0;
    std::vector<double> rsA(nA);
    P->funOff.assign(nA + 1, 0);
    P->funs.clear();
    P->splines.clear();
    std::map<std::pair<const dfx_atom *, u32>, u32> splAt; // shared by atoms of one file
    for (u32 a = 0; a < nA; ++a)
      {
        u32 nf;
        atomEnrichment(S, a, nf, rsA[a]);
        P->funOff[a + 1] = P->funOff[a] + nf;
        for (u32 k = 0; k < nf; ++k)
          {
            EnrFun E{};
            for (int d = 0; d < 3; ++d)
              E.r[d] = S.atomXYZ[3 * a + d];
            E.Z = S.atomZ[a];
            if (!S.atomData)
              {
                E.kind = (int)k;
                E.rc   = S.rcut;
                E.sm   = S.smooth;
              }
            else
              {
                const dfx_atom::Function &F = atomFunctions(S, a)[k];
                E.kind = NFUN;
                E.l    = F.q[1];
                E.m    = F.q[2];
                E.rc   = F.cutoff;
                E.sm   = F.smoothness;
                const auto key = std::make_pair(S.atomData[a], k);
                auto       it  = splAt.find(key);
                if (it == splAt.end())
                  {
                    it = splAt.emplace(key, (u32)P->splines.size()).first;
                    for (const auto *v : {&F.knots(S.atomData[a]->r), &F.y, &F.b, &F.c, &F.d})
                      P->splines.insert(P->splines.end(), v->begin(), v->end());
                  }
                E.splOff = it->second;
                E.splN   = (u32)F.y.size();
              }
            P->funs.push_back(E);
          }
      }
    // local potentials: smeared Coulomb, or the atom's pseudopotential vlocal
    P->pots.assign(S.nAtoms, PotDev{});
    for (u32 a = 0; a < S.nAtoms; ++a)
      {
        PotDev &V = P->pots[a];
        for (int d = 0; d < 3; ++d)
          V.r[d] = S.atomXYZ[3 * a + d];
        V.Z = S.atomZ[a];
        if (S.localPotential && S.localPotential[a])
          {
            const auto it = S.localPotential[a]->fields.find("vlocal");
            DFX_REQUIRE(it != S.localPotential[a]->fields.end() && it->second.size() == 1 &&
                          it->second[0].mixed,
                        "local potential of an atom: a UPF file read with the field vlocal");
            const dfx_atom::Function &F = it->second[0];
            V.pseudo = 1;
            V.splOff = (u32)P->splines.size();
            V.splN   = (u32)F.x.size();
            V.tail   = F.tail;
            V.cl     = F.cLeft;
            for (const auto *v : {&F.x, &F.y, &F.b, &F.c, &F.d})
              P->splines.insert(P->splines.end(), v->begin(), v->end());
          }
      }
    P->nEnriched    = P->funOff[nA];
    P->nDofs        = P->nClassical + P->nEnriched;
    P->atomCells.assign(nA, {});
    std::vector<std::vector<u32>> cellAtoms(P->nCells);
    for (u32 a = 0; a < nA; ++a)
      for (u32 c = 0; c < P->nCells; ++c)
        {
          const double rs = rsA[a];
          bool         in = true;
          for (int d = 0; d < 3 && in; ++d)
            {
              const double lo = P->cellLo[3 * (size_t)c + d], hi = lo + P->cellH[3 * (size_t)c + d];
              const double s0 = S.atomXYZ[3 * a + d] - rs, s1 = S.atomXYZ[3 * a + d] + rs;
              in = !((s0 < lo && s1 < lo) || (s0 > hi && s1 > hi));
            }
          if (in)
            {
              P->atomCells[a].push_back(c);
              cellAtoms[c].push_back(a);
            }
        }
    P->cellNdofs.resize(P->nCells);
    P->mapOff.assign(P->nCells + 1, 0);
    P->matOff.assign(P->nCells + 1, 0);
    for (u32 c = 0; c < P->nCells; ++c)
      {
        u32 n = ncl;
        for (u32 a : cellAtoms[c])
          n += P->funOff[a + 1] - P->funOff[a];
        P->cellNdofs[c]     = n;
        P->mapOff[c + 1]    = P->mapOff[c] + n;
        P->matOff[c + 1]    = P->matOff[c] + (u64)n * n;
        P->cellMap.insert(P->cellMap.end(), cmapC.begin() + (size_t)c * ncl,
                          cmapC.begin() + (size_t)(c + 1) * ncl);
        for (u32 a : cellAtoms[c])
          for (u32 k = P->funOff[a]; k < P->funOff[a + 1]; ++k)
            P->cellMap.push_back(P->nClassical + k);
      }

    // GLL lumped mass, C2P (no hanging rows here), constrained -> 1, inverse,
    // constrained -> 0  (CFEOverlapInverseOpContextGLL.t.cpp:380-439)
    std::vector<double> gx, gw;
    gllRule(np, gx, gw);
    P->mgll.assign(P->nDofs, 0.0);
    for (u32 c = 0; c < P->nCells; ++c)
      {
        const double *hh   = &P->cellH[3 * (size_t)c];
        const double  detJ = hh[0] * hh[1] * hh[2] / 8.0;
        for (u32 k = 0; k < np; ++k)
          for (u32 j = 0; j < np; ++j)
            for (u32 i = 0; i < np; ++i)
              P->mgll[cmapC[(size_t)c * ncl + i + np * (j + np * k)]] +=
                ((gw[i] * gw[j]) * gw[k]) * detJ;
      }
    P->dinv.assign(P->nDofs, 0.0);
    std::vector<char> isRow(P->nClassical, 0);
    for (u32 r : P->rows)
      isRow[r] = 1;
    for (u32 i = 0; i < P->nClassical; ++i)
      P->dinv[i] = isRow[i] ? 0.0 : 1.0 / P->mgll[i];
  }

  // Tabulate + batched GEMMs for a list of cells (caller indices), writing
  // their raw blocks of H, M, K on the device.
  // Tabulates the basis of `cellList` at the quadrature points (values F,
  // weighted values FW, FV = F W V, gradients G, GW; nQ x n column-major per
  // cell) in batches of one shape (n, nq) within a buffer budget, and calls
  // batch(cells, cnt, n, nQ, F, FW, FV, G, GW) for each batch (cells: the
  // batch's cells, in the batch's order).  extraPerCell: bytes the caller
  // allocates per cell of a batch.
  // valuesOnly: F alone is tabulated (the other buffers passed to batch are F
  // as well and must not be read)
  template <class Batch>
  void tabulateBatches(dfx_problem *P, const std::vector<u32> &cellList, const double *alpha,
                       size_t extraPerCell, Batch batch, bool valuesOnly = false)
  {
    if (cellList.empty())
      return;
    dfx_ctx    *ctx = P->ctx;
    hipStream_t s   = ctx->stream;
    const auto &S   = P->spec;
    const u32   np  = P->p + 1, ncl = np * np * np;
    const u32   nqC = S.nq ? S.nq : np, nqE = S.nqEnr ? S.nqEnr : 2 * np;
    // 1-D tables: [classical rule | enriched rule]
    std::vector<double> tabL, tabD, tabX, tabW, gll, gllw;
    gllRule(np, gll, gllw);
    auto addRule = [&](u32 nq) {
      std::vector<double> x, w;
      gaussRule(nq, x, w);
      const u32 off = (u32)tabL.size();
      for (u32 a = 0; a < np; ++a)
        for (u32 q = 0; q < nq; ++q)
          {
            tabL.push_back(lag(gll, a, x[q]));
            tabD.push_back(lagD(gll, a, x[q]));
          }
      tabX.insert(tabX.end(), x.begin(), x.end());
      tabW.insert(tabW.end(), w.begin(), w.end());
      return off;
    };
    const u32 offC = addRule(nqC), offE = addRule(nqE);
    DevBuf dL, dD, dX, dW, dAtoms, dMap;
    upload(dL, tabL, s);
    upload(dD, tabD, s);
    upload(dX, tabX, s);
    upload(dW, tabW, s);
    upload(dMap, P->cellMap, s);
    upload(dAtoms, P->pots, s);
    DevBuf dFuns, dSpl;
    upload(dFuns, P->funs, s);
    upload(dSpl, P->splines, s);

    // group by (n, nq) so each group is one batched GEMM shape
    std::vector<u32> order(cellList);
    auto nqOf = [&](u32 c) { return P->cellNdofs[c] > ncl ? nqE : nqC; };
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) {
      if (P->cellNdofs[a] != P->cellNdofs[b])
        return P->cellNdofs[a] < P->cellNdofs[b];
      return nqOf(a) < nqOf(b);
    });
    const size_t budget = (size_t)6 << 30; // bytes of tabulation buffers per batch
    size_t i0 = 0;
    while (i0 < order.size())
      {
        // batch: cells of one shape, within the memory budget
        const u32 n = P->cellNdofs[order[i0]], nq = nqOf(order[i0]);
        const u64 nQ = (u64)nq * nq * nq;
        const size_t per = (size_t)(valuesOnly ? 1 : 9) * nQ * n * sizeof(double) + extraPerCell;
        size_t i1 = i0;
        // (at most 65535 cells: one grid row per cell in the batch kernels)
        while (i1 < order.size() && P->cellNdofs[order[i1]] == n && nqOf(order[i1]) == nq &&
               (i1 - i0 + 1) * per <= budget && i1 - i0 < 65535)
          ++i1;
        if (i1 == i0)
          ++i1;
        const u32 cnt = (u32)(i1 - i0);
        std::vector<TabCell> tc(cnt);
        for (u32 t = 0; t < cnt; ++t)
          {
            const u32 c = order[i0 + t];
            TabCell   T{};
            for (int d = 0; d < 3; ++d)
              {
                T.lo[d] = P->cellLo[3 * (size_t)c + d];
                T.h[d]  = P->cellH[3 * (size_t)c + d];
              }
            T.n      = n;
            T.nq     = nq;
            T.tabOff = nq == nqC && P->cellNdofs[c] <= ncl ? offC : offE;
            T.mapOff = P->mapOff[c];
            T.bufOff = (u64)t * nQ * n;
            tc[t]    = T;
          }
        DevBuf dTc;
        upload(dTc, tc, s);
        DevBuf &bF  = work(ctx, "setup.F", (size_t)cnt * nQ * n * 8);
        DevBuf &bFW = valuesOnly ? bF : work(ctx, "setup.FW", (size_t)cnt * nQ * n * 8);
        DevBuf &bFV = valuesOnly ? bF : work(ctx, "setup.FV", (size_t)cnt * nQ * n * 8);
        DevBuf &bG  = valuesOnly ? bF : work(ctx, "setup.G", (size_t)cnt * 3 * nQ * n * 8);
        DevBuf &bGW = valuesOnly ? bF : work(ctx, "setup.GW", (size_t)cnt * 3 * nQ * n * 8);
        tabulateK<<<cnt, 256, 0, s>>>(dTc.as<TabCell>(), dMap.as<u32>(), P->nClassical,
                                      dL.as<double>(), dD.as<double>(), dX.as<double>(),
                                      dW.as<double>(), (int)np, dAtoms.as<PotDev>(),
                                      S.nAtoms, S.vSmear, dFuns.as<EnrFun>(),
                                      dSpl.as<double>(), alpha, valuesOnly,
                                      bF.as<double>(), bFW.as<double>(), bFV.as<double>(),
                                      bG.as<double>(), bGW.as<double>());
        checkLaunch("tabulate");
        batch(order.data() + i0, cnt, n, nQ, bF, bFW, bFV, bG, bGW);
        DFX_HIP(hipStreamSynchronize(s));
        i0 = i1;
      }
  }

  // cell matrices of `cellList` from the tabulated basis: M = F^T W F,
  // K = G^T W G, H = F^T W V F + K / 2 (rocBLAS batched GEMMs)
  void assembleCells(dfx_problem *P, const std::vector<u32> &cellList, const double *alpha)
  {
    dfx_ctx    *ctx = P->ctx;
    hipStream_t s   = ctx->stream;
    tabulateBatches(P, cellList, alpha, 0,
                    [&](const u32 *cells, u32 cnt, u32 n, u64 nQ, DevBuf &bF, DevBuf &bFW,
                        DevBuf &bFV, DevBuf &bG, DevBuf &bGW) {
        // pointer arrays: A (F or G), B (FW / FV / GW), C (H / M / K)
        std::vector<const double *> pF(cnt), pFW(cnt), pFV(cnt), pG(cnt), pGW(cnt);
        std::vector<double *>       pH(cnt), pM(cnt), pK(cnt);
        std::vector<u32>            nv(cnt, n);
        for (u32 t = 0; t < cnt; ++t)
          {
            const u32 c = cells[t];
            pF[t]  = bF.as<double>() + (u64)t * nQ * n;
            pFW[t] = bFW.as<double>() + (u64)t * nQ * n;
            pFV[t] = bFV.as<double>() + (u64)t * nQ * n;
            pG[t]  = bG.as<double>() + (u64)t * 3 * nQ * n;
            pGW[t] = bGW.as<double>() + (u64)t * 3 * nQ * n;
            pH[t]  = P->H.as<double>() + P->matOff[c];
            pM[t]  = P->M.as<double>() + P->matOff[c];
            pK[t]  = P->K.as<double>() + P->matOff[c];
          }
        DevBuf dpF, dpFW, dpFV, dpG, dpGW, dpH, dpM, dpK, dnv;
        upload(dpF, pF, s);
        upload(dpFW, pFW, s);
        upload(dpFV, pFV, s);
        upload(dpG, pG, s);
        upload(dpGW, pGW, s);
        upload(dpH, pH, s);
        upload(dpM, pM, s);
        upload(dpK, pK, s);
        upload(dnv, nv, s);
        const double one = 1.0, zero = 0.0;
        auto gemm = [&](const DevBuf &A, const DevBuf &B, const DevBuf &Cc, rocblas_int kdim) {
          CaptureLock       lk(captureMutex());
          rocblas_status st = rocblas_dgemm_batched(
            ctx->blas, rocblas_operation_transpose, rocblas_operation_none,
            (rocblas_int)n, (rocblas_int)n, kdim, &one, A.as<const double *>(), kdim,
            B.as<const double *>(), kdim, &zero, Cc.as<double *>(), (rocblas_int)n,
            (rocblas_int)cnt);
          if (st != rocblas_status_success)
            throw Error(DFX_E_HIP, "rocblas_dgemm_batched failed");
        };
        rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
        gemm(dpF, dpFW, dpM, (rocblas_int)nQ);        // M = F^T W F
        gemm(dpG, dpGW, dpK, (rocblas_int)(3 * nQ));  // K = G^T W G
        gemm(dpF, dpFV, dpH, (rocblas_int)nQ);        // H = F^T W V F ...
        addHalfK<<<cnt, 256, 0, s>>>(dpH.as<double *>(), dpK.as<double *>(), dnv.as<u32>(),
                                     cnt);             // ... + K / 2
        checkLaunch("addHalfK");
        // the pointer arrays are released when this batch returns
        DFX_HIP(hipStreamSynchronize(s));
                    });
  }

  // quadrature layout of the problem's cells: cell c's points are
  // [off[c], off[c + 1]), x fastest; W (optional) = JxW of every point,
  // XYZ (optional) = coordinates of every point (x, y, z per point)
  void quadLayout(const dfx_problem *P, std::vector<u64> &off, std::vector<double> *W,
                  std::vector<double> *XYZ = nullptr)
  {
    const auto &S   = P->spec;
    const u32   np  = P->p + 1, ncl = np * np * np;
    const u32   nqC = S.nq ? S.nq : np, nqE = S.nqEnr ? S.nqEnr : 2 * np;
    std::vector<double> xC, wC, xE, wE;
    gaussRule(nqC, xC, wC);
    gaussRule(nqE, xE, wE);
    off.assign(P->nCells + 1, 0);
    if (W)
      W->clear();
    if (XYZ)
      XYZ->clear();
    for (u32 c = 0; c < P->nCells; ++c)
      {
        const bool enr = P->cellNdofs[c] > ncl;
        const u32  nq  = enr ? nqE : nqC;
        off[c + 1]     = off[c] + (u64)nq * nq * nq;
        if (XYZ)
          {
            // reference point x_q in [-1, 1] maps to lo + h (x_q + 1) / 2
            const auto   &x  = enr ? xE : xC;
            const double *lo = &P->cellLo[3 * (size_t)c], *h = &P->cellH[3 * (size_t)c];
            for (u32 qc = 0; qc < nq; ++qc)
              for (u32 qb = 0; qb < nq; ++qb)
                for (u32 qa = 0; qa < nq; ++qa)
                  {
                    XYZ->push_back(lo[0] + h[0] * (0.5 * (x[qa] + 1.0)));
                    XYZ->push_back(lo[1] + h[1] * (0.5 * (x[qb] + 1.0)));
                    XYZ->push_back(lo[2] + h[2] * (0.5 * (x[qc] + 1.0)));
                  }
          }
        if (!W)
          continue;
        const auto  &w    = enr ? wE : wC;
        const double detJ = P->cellH[3 * (size_t)c] * P->cellH[3 * (size_t)c + 1] *
                            P->cellH[3 * (size_t)c + 2] / 8.0;
        for (u32 qc = 0; qc < nq; ++qc)
          for (u32 qb = 0; qb < nq; ++qb)
            for (u32 qa = 0; qa < nq; ++qa)
              W->push_back(((w[qa] * w[qb]) * w[qc]) * detJ);
      }
  }

  // DensityCalculator::computeRho (src/ksdft/DensityCalculator.t.cpp:288-420)
  // with the interpolation of FEBasisOperations::interpolate
  // (src/basis/FEBasisOperations.t.cpp:996-1277): per batch of cells of one
  // shape, the basis re-tabulated at the quadrature points, the cells'
  // wavefunction rows gathered, psi = Xc F^T by a batched GEMM, and rho
  // reduced over the states.  One batch of states (all of X).
  void density(dfx_problem *P, const double *occupation, View<const double> X,
               double *rho)
  {
    dfx_ctx    *ctx = P->ctx;
    hipStream_t s   = ctx->stream;
    const u32   B   = X.B;
    std::vector<u64> qoff;
    quadLayout(P, qoff, nullptr);
    DevBuf dQoff, dOcc, dMap, dMapOff;
    upload(dQoff, qoff, s);
    upload(dOcc, occupation, B, s);
    upload(dMap, P->cellMap, s);
    upload(dMapOff, P->mapOff, s);
    std::vector<u32> all(P->nCells);
    std::iota(all.begin(), all.end(), 0u);
    u32 maxN = 0;
    u64 maxQ = 0;
    for (u32 c = 0; c < P->nCells; ++c)
      {
        maxN = std::max(maxN, P->cellNdofs[c]);
        maxQ = std::max(maxQ, qoff[c + 1] - qoff[c]);
      }
    const double *alpha = P->alpha.p ? P->alpha.as<double>() : nullptr;
    tabulateBatches(P, all, alpha, (size_t)(maxN + maxQ) * B * sizeof(double),
                    [&](const u32 *cells, u32 cnt, u32 n, u64 nQ, DevBuf &bF, DevBuf &,
                        DevBuf &, DevBuf &, DevBuf &) {
                      DevBuf &bXc  = work(ctx, "dens.xc", (size_t)cnt * n * B * 8);
                      DevBuf &bPsi = work(ctx, "dens.psi", (size_t)cnt * nQ * B * 8);
                      DevBuf  dCells;
                      upload(dCells, cells, cnt, s);
                      gatherCellRowsK<<<cnt, 256, 0, s>>>(X.p, X.ld, B, dMap.as<u32>(),
                                                          dMapOff.as<u32>(), dCells.as<u32>(),
                                                          n, bXc.as<double>());
                      checkLaunch("gatherCellRows");
                      std::vector<const double *> pA(cnt), pB(cnt);
                      std::vector<double *>       pC(cnt);
                      for (u32 t = 0; t < cnt; ++t)
                        {
                          pA[t] = bXc.as<double>() + (u64)t * n * B;
                          pB[t] = bF.as<double>() + (u64)t * nQ * n;
                          pC[t] = bPsi.as<double>() + (u64)t * nQ * B;
                        }
                      DevBuf dpA, dpB, dpC;
                      upload(dpA, pA, s);
                      upload(dpB, pB, s);
                      upload(dpC, pC, s);
                      const double one = 1.0, zero = 0.0;
                      {
                        CaptureLock lk(captureMutex());
                        rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
                        // psi (B x nQ) = Xc (B x n) F^T (F: nQ x n)
                        if (rocblas_dgemm_batched(ctx->blas, rocblas_operation_none,
                                                  rocblas_operation_transpose, (rocblas_int)B,
                                                  (rocblas_int)nQ, (rocblas_int)n, &one,
                                                  dpA.as<const double *>(), (rocblas_int)B,
                                                  dpB.as<const double *>(), (rocblas_int)nQ,
                                                  &zero, dpC.as<double *>(), (rocblas_int)B,
                                                  (rocblas_int)cnt) != rocblas_status_success)
                          throw Error(DFX_E_HIP, "rocblas_dgemm_batched failed");
                      }
                      rhoFromPsiK<<<dim3((u32)((nQ + 255) / 256), cnt), 256, 0, s>>>(
                        bPsi.as<double>(), B, (u32)nQ, dCells.as<u32>(), dQoff.as<u64>(),
                        dOcc.as<double>(), rho);
                      checkLaunch("rhoFromPsi");
                      // the batch's pointer arrays are released when it returns
                      DFX_HIP(hipStreamSynchronize(s));
                    },
                    true);
  }

  // fw[(t nQ + q) B + j] = JxW[q] f[(quadOff[c] + q) ldf + j] for the cells
  // c = cells[t] of one batch: the inp o JxW product of
  // FEBasisOperations::integrateWithBasisValues (FEBasisOperations.t.cpp:
  // 2043-2059), JxW first
  __global__ void weightQuadRowsK(const double *f, u32 ldf, u32 B, const double *W,
                                  const u64 *quadOff, const u32 *cells, u32 nQ, double *fw)
  {
    const u32 t = blockIdx.y, c = cells[t];
    const u64 q0 = quadOff[c];
    for (u32 e = blockIdx.x * blockDim.x + threadIdx.x; e < nQ * B;
         e += gridDim.x * blockDim.x)
      {
        const u32 q = e / B, j = e - q * B;
        fw[((u64)t * nQ + q) * B + j] = W[q0 + q] * f[(q0 + q) * ldf + j];
      }
  }
} // namespace

namespace dfx
{
  // FEBasisOperations::integrateWithBasisValues (src/basis/FEBasisOperations.t.cpp:
  // 1846-2190): out = C2P(sum_e Q_e^T N_e^T (JxW o f_e)) with the problem's
  // basis (orthogonalized enrichment included) re-tabulated at its quadrature
  // points.  Per batch of cells of one shape the weighted quadrature rows are
  // formed and C_e (B x n_e) = fw_e (B x nQ) F_e (nQ x n_e) is one batched
  // GEMM written straight into the cell's slots of the cell->dof map; the
  // slots are then added into `out` dof by dof in cell order
  // (addCellWiseDataToFieldData, FECellWiseDataOperations.t.cpp:89-151) and
  // the constraints (when given) distribute child rows to their parents.
  void problemIntegrate(dfx_problem *P, const dfx_constraints *cons, View<const double> f,
                        View<double> out)
  {
    dfx_ctx    *ctx = P->ctx;
    hipStream_t s   = ctx->stream;
    const u32   B   = f.B;
    std::vector<u64>    qoff;
    std::vector<double> W;
    quadLayout(P, qoff, &W);
    DFX_REQUIRE(f.N == qoff.back(), "integrate: one row per quadrature point expected");
    DFX_REQUIRE(out.N == P->nDofs && out.B == B, "integrate: output must be nDofs x B");
    const u64 nSlots = P->cellMap.size();
    // dof -> slots of the cell->dof map, in cell order
    std::vector<u32> sOff(P->nDofs + 1, 0), sList(nSlots);
    for (u64 t = 0; t < nSlots; ++t)
      ++sOff[P->cellMap[t] + 1];
    for (u32 d = 0; d < P->nDofs; ++d)
      sOff[d + 1] += sOff[d];
    {
      std::vector<u32> fp(sOff.begin(), sOff.end() - 1);
      for (u64 t = 0; t < nSlots; ++t)
        sList[fp[P->cellMap[t]]++] = (u32)t;
    }
    DevBuf dQoff, dW, dSOff, dSList;
    upload(dQoff, qoff, s);
    upload(dW, W, s);
    upload(dSOff, sOff, s);
    upload(dSList, sList, s);
    DevBuf &bSlots = work(ctx, "integ.slots", nSlots * B * sizeof(double));
    u64 maxQ = 0;
    for (u32 c = 0; c < P->nCells; ++c)
      maxQ = std::max(maxQ, qoff[c + 1] - qoff[c]);
    std::vector<u32> all(P->nCells);
    std::iota(all.begin(), all.end(), 0u);
    const double *alpha = P->alpha.p ? P->alpha.as<double>() : nullptr;
    tabulateBatches(P, all, alpha, (size_t)maxQ * B * sizeof(double),
                    [&](const u32 *cells, u32 cnt, u32 n, u64 nQ, DevBuf &bF, DevBuf &,
                        DevBuf &, DevBuf &, DevBuf &) {
                      DevBuf &bFw = work(ctx, "integ.fw", (size_t)cnt * nQ * B * 8);
                      DevBuf  dCells;
                      upload(dCells, cells, cnt, s);
                      weightQuadRowsK<<<dim3((u32)((nQ * B + 255) / 256), cnt), 256, 0, s>>>(
                        f.p, f.ld, B, dW.as<double>(), dQoff.as<u64>(), dCells.as<u32>(),
                        (u32)nQ, bFw.as<double>());
                      checkLaunch("weightQuadRows");
                      std::vector<const double *> pA(cnt), pB(cnt);
                      std::vector<double *>       pC(cnt);
                      for (u32 t = 0; t < cnt; ++t)
                        {
                          pA[t] = bFw.as<double>() + (u64)t * nQ * B;
                          pB[t] = bF.as<double>() + (u64)t * nQ * n;
                          pC[t] = bSlots.as<double>() + (u64)P->mapOff[cells[t]] * B;
                        }
                      DevBuf dpA, dpB, dpC;
                      upload(dpA, pA, s);
                      upload(dpB, pB, s);
                      upload(dpC, pC, s);
                      const double one = 1.0, zero = 0.0;
                      {
                        CaptureLock lk(captureMutex());
                        rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host);
                        // C (B x n, the cell's slot rows) = fw (B x nQ) F (nQ x n)
                        if (rocblas_dgemm_batched(ctx->blas, rocblas_operation_none,
                                                  rocblas_operation_none, (rocblas_int)B,
                                                  (rocblas_int)n, (rocblas_int)nQ, &one,
                                                  dpA.as<const double *>(), (rocblas_int)B,
                                                  dpB.as<const double *>(), (rocblas_int)nQ,
                                                  &zero, dpC.as<double *>(), (rocblas_int)B,
                                                  (rocblas_int)cnt) != rocblas_status_success)
                          throw Error(DFX_E_HIP, "rocblas_dgemm_batched failed");
                      }
                      // the batch's pointer arrays are released when it returns
                      DFX_HIP(hipStreamSynchronize(s));
                    },
                    true);
    k::fill<double>(s, out, 0.0);
    k::scatterAdd<double>(s, dSOff.as<u32>(), dSList.as<u32>(), P->nDofs,
                          View<const double>{bSlots.as<double>(), (u32)nSlots, B, B}, out);
    if (cons)
      k::c2p<double>(s, cons, out);
    DFX_HIP(hipStreamSynchronize(s));
    ctx->work.erase("integ.fw");
    ctx->work.erase("integ.slots");
  }
} // namespace dfx

namespace
{
  // Per-atom L2 projection of the enrichment functions on the classical
  // nodes interior to the atom's support, with a cell-wise mass operator over
  // the support cells (non-interior nodes as zero Dirichlet rows) and PCG.
  void orthogonalize(dfx_problem *P)
  {
    dfx_ctx    *ctx = P->ctx;
    hipStream_t s   = ctx->stream;
    const u32   np = P->p + 1, ncl = np * np * np;
    const u32   nA = (u32)P->atomCells.size();
    P->alpha.alloc((size_t)P->nEnriched * P->nClassical * sizeof(double));
    DFX_HIP(hipMemsetAsync(P->alpha.p, 0, P->alpha.bytes, s));
    // global incidence count of each classical node
    std::vector<u32> inc(P->nClassical, 0);
    for (u32 c = 0; c < P->nCells; ++c)
      for (u32 l = 0; l < ncl; ++l)
        ++inc[P->cellMap[P->mapOff[c] + l]];
    std::vector<char> isRow(P->nClassical, 0);
    for (u32 r : P->rows)
      isRow[r] = 1;
    DevBuf dRawOff, dNe;
    upload(dRawOff, P->matOff, s);
    upload(dNe, P->cellNdofs, s);
    for (u32 a = 0; a < nA; ++a)
      {
        const auto &cells = P->atomCells[a];
        const u32   nc    = (u32)cells.size();
        // local numbering of the support's classical nodes
        std::vector<u32> loc(P->nClassical, UINT32_MAX), glob, lmap((size_t)nc * ncl), cnt;
        for (u32 i = 0; i < nc; ++i)
          for (u32 l = 0; l < ncl; ++l)
            {
              const u32 g = P->cellMap[P->mapOff[cells[i]] + l];
              if (loc[g] == UINT32_MAX)
                {
                  loc[g] = (u32)glob.size();
                  glob.push_back(g);
                  cnt.push_back(0);
                }
              ++cnt[loc[g]];
              lmap[(size_t)i * ncl + l] = loc[g];
            }
        const u32 nl = (u32)glob.size();
        std::vector<u32> interior(nl), crow, csz, coff;
        for (u32 i = 0; i < nl; ++i)
          {
            interior[i] = (cnt[i] == inc[glob[i]] && !isRow[glob[i]]) ? 1u : 0u;
            if (!interior[i])
              {
                crow.push_back(i);
                csz.push_back(0);
                coff.push_back(0);
              }
          }
        std::vector<double> cinh(crow.size(), 0.0);
        // column of this atom's first function in each support cell
        std::vector<u32> colStart(nc);
        for (u32 i = 0; i < nc; ++i)
          {
            const u32 c = cells[i];
            u32       j = ncl;
            while (P->cellMap[P->mapOff[c] + j] != P->nClassical + P->funOff[a])
              ++j;
            colStart[i] = j;
          }
        DevBuf dCells, dCol, blocks, rhsSlots;
        upload(dCells, cells, s);
        upload(dCol, colStart, s);
        blocks.alloc((size_t)nc * ncl * ncl * sizeof(double));
        const u32 nf = P->funOff[a + 1] - P->funOff[a];
        rhsSlots.alloc((size_t)nc * ncl * nf * sizeof(double));
        extractClassicalBlock<<<nc, 256, 0, s>>>(P->M.as<double>(), dRawOff.as<u64>(),
                                                 dNe.as<u32>(), dCells.as<u32>(), ncl,
                                                 blocks.as<double>());
        extractMixedBlock<<<nc, 256, 0, s>>>(P->M.as<double>(), dRawOff.as<u64>(),
                                             dNe.as<u32>(), dCells.as<u32>(), dCol.as<u32>(),
                                             ncl, nf, rhsSlots.as<double>());
        checkLaunch("extract");
        std::vector<double> hBlocks((size_t)nc * ncl * ncl);
        DFX_HIP(hipMemcpyAsync(hBlocks.data(), blocks.p, blocks.bytes, hipMemcpyDeviceToHost, s));
        DFX_HIP(hipStreamSynchronize(s));
        std::vector<u32> cn(nc, ncl);
        dfx_constraints *cons = nullptr;
        dfx_op *op = nullptr, *pc = nullptr;
        dfx_mv *rs = nullptr, *rhs = nullptr, *x = nullptr;
        auto cleanup = [&] {
          if (op) dfx_op_destroy(op);
          if (pc) dfx_op_destroy(pc);
          if (rs) dfx_mv_destroy(rs);
          if (rhs) dfx_mv_destroy(rhs);
          if (x) dfx_mv_destroy(x);
          if (cons) dfx_constraints_destroy(cons);
        };
        auto chk = [&](int rc, const char *w) {
          if (rc < 0)
            {
              std::string m = std::string(w) + ": " + ctx->err;
              cleanup();
              throw Error(rc, m);
            }
        };
        chk(dfx_constraints_create(ctx, nl, (u32)crow.size(), crow.data(), csz.data(),
                                   coff.data(), nullptr, nullptr, cinh.data(), &cons),
            "orthogonalize constraints");
        chk(dfx_op_create_cellwise(ctx, nc, cn.data(), lmap.data(), hBlocks.data(), nl, nl,
                                   cons, 0, &op),
            "orthogonalize mass operator");
        // rhs d = sum_e Q_e^T (mixed block), then zero the non-interior rows
        chk(dfx_mv_wrap(ctx, rhsSlots.p, nc * ncl, nf, DFX_F64, &rs), "wrap");
        chk(dfx_mv_create(ctx, nl, nf, DFX_F64, &rhs), "rhs");
        chk(dfx_op_scatter_add(op, rs, rhs), "rhs assembly");
        chk(dfx_constraints_set_constrained_nodes(cons, rhs, 0.0), "rhs rows");
        // Jacobi: assembled diagonal of the local mass
        std::vector<double> diag(nl, 0.0);
        for (u32 i = 0; i < nc; ++i)
          for (u32 l = 0; l < ncl; ++l)
            diag[lmap[(size_t)i * ncl + l]] += hBlocks[(size_t)i * ncl * ncl + l + (size_t)ncl * l];
        for (u32 i = 0; i < nl; ++i)
          diag[i] = interior[i] ? 1.0 / diag[i] : 1.0;
        chk(dfx_op_create_diagonal(ctx, nl, diag.data(), 0, nullptr, nullptr, &pc), "pc");
        chk(dfx_mv_create(ctx, nl, nf, DFX_F64, &x), "x");
        u32 iters = 0;
        int rc = dfx_cg_solve(ctx, op, pc, rhs, x, 1e-300, 1e-13, 1e10, 10000, &iters, nullptr);
        if (rc != 0)
          {
            cleanup();
            throw Error(DFX_E_STATE, "enrichment orthogonalization: PCG did not converge");
          }
        DevBuf dGlob, dInt;
        upload(dGlob, glob, s);
        upload(dInt, interior, s);
        scatterAlpha<<<256, 256, 0, s>>>(static_cast<double *>(x->d), dGlob.as<u32>(),
                                         dInt.as<u32>(), nl, P->funOff[a], nf, P->nClassical,
                                         P->alpha.as<double>());
        checkLaunch("scatterAlpha");
        DFX_HIP(hipStreamSynchronize(s));
        cleanup();
      }
  }

  // M^ee assembled in cell order from the final cell matrices, and its inverse
  void enrichedBlock(dfx_problem *P)
  {
    const u32 nE = P->nEnriched, nC = P->nClassical;
    if (!nE)
      return;
    const u32 ncl = (P->p + 1) * (P->p + 1) * (P->p + 1);
    P->enrM.assign((size_t)nE * nE, 0.0);
    // the enriched sub-blocks of the enriched cells, packed on the device and
    // copied in one transfer; summed on the host in cell order
    std::vector<u32> cells;
    std::vector<u64> offs;
    u64              total = 0;
    for (u32 c = 0; c < P->nCells; ++c)
      if (P->cellNdofs[c] != ncl)
        {
          const u64 m = P->cellNdofs[c] - ncl;
          cells.push_back(c);
          offs.push_back(total);
          total += m * m;
        }
    if (!cells.empty())
      {
        hipStream_t s = P->ctx->stream;
        DevBuf      dCells, dOffs, dMatOff, dNe, packed((size_t)total * 8);
        upload(dCells, cells, s);
        upload(dOffs, offs, s);
        upload(dMatOff, P->matOff, s);
        upload(dNe, P->cellNdofs, s);
        extractEnrBlockK<<<(u32)cells.size(), 256, 0, s>>>(
          P->M.as<double>(), dMatOff.as<u64>(), dNe.as<u32>(), dCells.as<u32>(),
          dOffs.as<u64>(), ncl, packed.as<double>());
        checkLaunch("extractEnrBlock");
        std::vector<double> h((size_t)total);
        DFX_HIP(hipMemcpyAsync(h.data(), packed.p, (size_t)total * 8, hipMemcpyDeviceToHost, s));
        DFX_HIP(hipStreamSynchronize(s));
        for (size_t q = 0; q < cells.size(); ++q)
          {
            const u32    c = cells[q], n = P->cellNdofs[c], m = n - ncl;
            const double *blk = h.data() + offs[q];
            for (u32 j = ncl; j < n; ++j)
              for (u32 i = ncl; i < n; ++i)
                P->enrM[(P->cellMap[P->mapOff[c] + i] - nC) +
                        (size_t)nE * (P->cellMap[P->mapOff[c] + j] - nC)] +=
                  blk[(i - ncl) + (size_t)m * (j - ncl)];
          }
      }
    // inverse through the Cholesky factor: L L^T = M, inv = L^-T L^-1
    std::vector<double> L(P->enrM);
    for (u32 j = 0; j < nE; ++j)
      {
        double d = L[j + (size_t)nE * j];
        for (u32 k = 0; k < j; ++k)
          d -= L[j + (size_t)nE * k] * L[j + (size_t)nE * k];
        DFX_REQUIRE(d > 0.0, "enriched overlap block is not positive definite");
        d                        = std::sqrt(d);
        L[j + (size_t)nE * j]    = d;
        for (u32 i = j + 1; i < nE; ++i)
          {
            double t = L[i + (size_t)nE * j];
            for (u32 k = 0; k < j; ++k)
              t -= L[i + (size_t)nE * k] * L[j + (size_t)nE * k];
            L[i + (size_t)nE * j] = t / d;
          }
      }
    std::vector<double> Li((size_t)nE * nE, 0.0);
    for (u32 j = 0; j < nE; ++j)
      {
        Li[j + (size_t)nE * j] = 1.0 / L[j + (size_t)nE * j];
        for (u32 i = j + 1; i < nE; ++i)
          {
            double t = 0.0;
            for (u32 k = j; k < i; ++k)
              t += L[i + (size_t)nE * k] * Li[k + (size_t)nE * j];
            Li[i + (size_t)nE * j] = -t / L[i + (size_t)nE * i];
          }
      }
    P->enrMinv.assign((size_t)nE * nE, 0.0);
    for (u32 j = 0; j < nE; ++j)
      for (u32 i = 0; i < nE; ++i)
        {
          double t = 0.0;
          for (u32 k = std::max(i, j); k < nE; ++k)
            t += Li[k + (size_t)nE * i] * Li[k + (size_t)nE * j];
          P->enrMinv[i + (size_t)nE * j] = t;
        }
  }
} // namespace

extern "C"
{
  int dfx_problem_build(dfx_ctx *ctx, const dfx_problem_spec *spec, dfx_problem **out)
  {
    return guard(ctx, [&] {
      DFX_REQUIRE(spec && out, "null argument");
      DFX_REQUIRE(!spec->nAtoms || (spec->atomXYZ && spec->atomZ), "atoms missing");
      std::unique_ptr<dfx_problem> P(new dfx_problem());
      P->ctx  = ctx;
      P->spec = *spec;
      P->atomXYZ.assign(spec->atomXYZ, spec->atomXYZ + 3 * (size_t)spec->nAtoms);
      P->atomZ.assign(spec->atomZ, spec->atomZ + spec->nAtoms);
      P->spec.atomXYZ = P->atomXYZ.data();
      P->spec.atomZ   = P->atomZ.data();
      if (spec->nAtoms && spec->localPotential)
        {
          P->localPot.assign(spec->localPotential, spec->localPotential + spec->nAtoms);
          P->spec.localPotential = P->localPot.data();
        }
      else
        P->spec.localPotential = nullptr;
      if (spec->enrich && spec->nAtoms && spec->atomData)
        {
          P->atomData.assign(spec->atomData, spec->atomData + spec->nAtoms);
          P->enrichField      = spec->enrichField ? spec->enrichField : "orbital";
          P->spec.atomData    = P->atomData.data();
          P->spec.enrichField = P->enrichField.c_str();
          for (u32 a = 0; a < spec->nAtoms; ++a)
            {
              DFX_REQUIRE(P->atomData[a], "atom data missing for an atom");
              const auto &F = atomFunctions(P->spec, a);
              DFX_REQUIRE(!F.empty(), "atom data with no enrichment function");
              for (const auto &f : F)
                {
                  if (f.q[1] > 2)
                    throw Error(DFX_E_UNSUPPORTED, "radial functions with l > 2 are not supported");
                  DFX_REQUIRE(!f.mixed, "a mixed (local potential) field cannot enrich the basis");
                }
            }
        }
      else
        {
          P->spec.atomData    = nullptr;
          P->spec.enrichField = nullptr;
        }
      buildHost(P.get());
      const u64 nm = P->matOff.back();
      P->H.alloc(nm * 8);
      P->M.alloc(nm * 8);
      P->K.alloc(nm * 8);
      std::vector<u32> all(P->nCells);
      std::iota(all.begin(), all.end(), 0u);
      assembleCells(P.get(), all, nullptr);
      if (P->nEnriched && spec->orthogonalize)
        {
          orthogonalize(P.get());
          std::vector<u32> enr;
          for (u32 c = 0; c < P->nCells; ++c)
            if (P->cellNdofs[c] > (P->p + 1) * (P->p + 1) * (P->p + 1))
              enr.push_back(c);
          assembleCells(P.get(), enr, P->alpha.as<double>());
        }
      enrichedBlock(P.get());
      ctx->work.erase("setup.F");
      ctx->work.erase("setup.FW");
      ctx->work.erase("setup.FV");
      ctx->work.erase("setup.G");
      ctx->work.erase("setup.GW");
      ctxRetain(ctx);
      *out = P.release();
      return 0;
    });
  }

  int dfx_problem_sizes(const dfx_problem *P, uint64_t *s)
  {
    if (!P || !s)
      return DFX_E_ARG;
    s[0] = P->nCells;
    s[1] = P->nDofs;
    s[2] = P->nClassical;
    s[3] = P->nEnriched;
    s[4] = P->cellMap.size();
    s[5] = P->matOff.back();
    s[6] = P->rows.size();
    s[7] = P->cols.size();
    return 0;
  }

  int dfx_problem_get(const dfx_problem *P, int which, void *host)
  {
    if (!P || !host)
      return DFX_E_ARG;
    return guard(P->ctx, [&] {
      auto cpu = [&](const void *src, size_t bytes) {
        if (bytes)
          std::memcpy(host, src, bytes);
        return 0;
      };
      auto dev = [&](const DevBuf &b, size_t bytes) {
        if (bytes)
          {
            DFX_HIP(hipMemcpyAsync(host, b.p, bytes, hipMemcpyDeviceToHost, P->ctx->stream));
            DFX_HIP(hipStreamSynchronize(P->ctx->stream));
          }
        return 0;
      };
      const size_t nm = P->matOff.back() * 8;
      switch (which)
        {
          case DFX_PROB_CELL_NDOFS: return cpu(P->cellNdofs.data(), P->cellNdofs.size() * 4);
          case DFX_PROB_CELL_MAP: return cpu(P->cellMap.data(), P->cellMap.size() * 4);
          case DFX_PROB_ROWS: return cpu(P->rows.data(), P->rows.size() * 4);
          case DFX_PROB_ROW_SIZES: return cpu(P->rowSizes.data(), P->rowSizes.size() * 4);
          case DFX_PROB_ROW_OFFSETS: return cpu(P->rowOffsets.data(), P->rowOffsets.size() * 4);
          case DFX_PROB_COLS: return cpu(P->cols.data(), P->cols.size() * 4);
          case DFX_PROB_WEIGHTS: return cpu(P->weights.data(), P->weights.size() * 8);
          case DFX_PROB_INHOM: return cpu(P->inhom.data(), P->inhom.size() * 8);
          case DFX_PROB_H: return dev(P->H, nm);
          case DFX_PROB_M: return dev(P->M, nm);
          case DFX_PROB_K: return dev(P->K, nm);
          case DFX_PROB_DINV: return cpu(P->dinv.data(), P->dinv.size() * 8);
          case DFX_PROB_MGLL: return cpu(P->mgll.data(), P->mgll.size() * 8);
          case DFX_PROB_ENR_MINV: return cpu(P->enrMinv.data(), P->enrMinv.size() * 8);
          case DFX_PROB_ENR_M: return cpu(P->enrM.data(), P->enrM.size() * 8);
          case DFX_PROB_NODE_XYZ: return cpu(P->nodeXYZ.data(), P->nodeXYZ.size() * 8);
          case DFX_PROB_ALPHA:
            return P->alpha.p ? dev(P->alpha, P->alpha.bytes)
                              : cpu(nullptr, 0);
          case DFX_PROB_QUAD_OFFSETS:
            {
              std::vector<u64> off;
              quadLayout(P, off, nullptr);
              return cpu(off.data(), off.size() * 8);
            }
          case DFX_PROB_QUAD_W:
            {
              std::vector<u64>    off;
              std::vector<double> w;
              quadLayout(P, off, &w);
              return cpu(w.data(), w.size() * 8);
            }
          case DFX_PROB_QUAD_XYZ:
            {
              std::vector<u64>    off;
              std::vector<double> xyz;
              quadLayout(P, off, nullptr, &xyz);
              return cpu(xyz.data(), xyz.size() * 8);
            }
        }
      throw Error(DFX_E_ARG, "unknown problem array");
    });
  }

  int dfx_problem_density(dfx_problem *P, const double *occupation, dfx_mv *X, dfx_mv *rho)
  {
    if (!P || !occupation || !X || !rho || X->dtype != DFX_F64 || rho->dtype != DFX_F64 ||
        X->N != P->nDofs || rho->B != 1 || !X->B)
      return DFX_E_ARG;
    return guard(P->ctx, [&] {
      std::vector<u64> off;
      quadLayout(P, off, nullptr);
      DFX_REQUIRE(rho->N == off.back(), "rho must hold one value per quadrature point");
      density(P, occupation,
              View<const double>{static_cast<const double *>(X->d), X->N, X->B, X->B},
              static_cast<double *>(rho->d));
      return 0;
    });
  }

  int dfx_problem_integrate(dfx_problem *P, dfx_constraints *cons, dfx_mv *f, dfx_mv *out)
  {
    if (!P || !f || !out || f->dtype != DFX_F64 || out->dtype != DFX_F64 || !f->B)
      return DFX_E_ARG;
    return guard(P->ctx, [&] {
      DFX_REQUIRE(!cons || cons->nDofs == P->nDofs, "constraints built for another N");
      problemIntegrate(P, cons,
                       View<const double>{static_cast<const double *>(f->d), f->N, f->B, f->B},
                       View<double>{static_cast<double *>(out->d), out->N, out->B, out->B});
      return 0;
    });
  }

  int dfx_problem_create_operators(dfx_problem *P, uint32_t flags, dfx_op **H, dfx_op **M,
                                   dfx_op **K, dfx_op **MInv, dfx_op **MLump,
                                   dfx_constraints **consOut)
  {
    if (!P)
      return DFX_E_ARG;
    dfx_ctx *ctx = P->ctx;
    return guard(ctx, [&] {
      dfx_constraints *cons = nullptr;
      int rc = dfx_constraints_create(ctx, P->nDofs, (u32)P->rows.size(), P->rows.data(),
                                      P->rowSizes.data(), P->rowOffsets.data(),
                                      P->cols.data(), P->weights.data(), P->inhom.data(),
                                      &cons);
      if (rc)
        throw Error(rc, ctx->err);
      auto cell = [&](const DevBuf &mats, dfx_op **o) {
        if (!o)
          return;
        std::unique_ptr<dfx_op> op(new dfx_op());
        buildCellwise(ctx, op.get(), P->nCells, P->cellNdofs.data(), P->cellMap.data(),
                      P->nDofs, P->nClassical, cons, flags);
        uploadCellMatrices(op.get(), mats.as<double>(), /*device=*/true);
        ctxRetain(ctx);
        *o = op.release();
      };
      cell(P->H, H);
      cell(P->M, M);
      cell(P->K, K);
      if (MInv)
        {
          rc = dfx_op_create_diagonal(ctx, P->nDofs, P->dinv.data(), P->nEnriched,
                                      P->nEnriched ? P->enrMinv.data() : nullptr, cons, MInv);
          if (rc)
            throw Error(rc, ctx->err);
        }
      if (MLump)
        {
          rc = dfx_op_create_diagonal(ctx, P->nDofs, P->mgll.data(), P->nEnriched,
                                      P->nEnriched ? P->enrM.data() : nullptr, cons, MLump);
          if (rc)
            throw Error(rc, ctx->err);
        }
      if (consOut)
        *consOut = cons;
      else
        dfx_constraints_destroy(cons); // operators keep their own reference
      return 0;
    });
  }

  int dfx_problem_destroy(dfx_problem *P)
  {
    if (!P)
      return DFX_E_ARG;
    dfx_ctx *ctx = P->ctx;
    delete P;
    ctxRelease(ctx);
    return 0;
  }
}
