// CPU oracle for the dft-efe Kohn-Sham eigensolver hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py use it, and only as the checker / CPU baseline.
//
// It is a plain FP64 restatement of the reference algorithm, written from the
// reference's behaviour (file:line citations on every function, relative to
// the dft-efe source tree).  Reductions use the reference's sequential order;
// the library is compiled with -ffp-contract=off so that a*b+c is never fused
// unless the code calls std::fma explicitly (the constraint kernels do, on
// both sides, so that those results are bit-comparable).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace oracle
{
  using u32 = uint32_t;
  using u64 = uint64_t;

  // ---------------------------------------------------------------- RNG
  // SplitMix64 (Steele, Lea, Flood 2014), u53 -> [0,1).  Replaces the
  // reference's time-seeded std::random_device (src/utils/RandNumGen.cpp:40-43)
  // so that inputs are reproducible.
  struct Rng
  {
    u64 s;
    explicit Rng(u64 seed) : s(seed) {}
    u64    next();
    double u01();
    double uniform(double a, double b) { return a + (b - a) * u01(); }
  };

  // ---------------------------------------------------------- quadrature
  void gaussLegendre(int n, std::vector<double> &x, std::vector<double> &w);
  void gaussLobatto(int n, std::vector<double> &x, std::vector<double> &w);
  double lagrange(const std::vector<double> &nodes, int k, double x);
  double lagrangeDeriv(const std::vector<double> &nodes, int k, double x);

  // ---------------------------------------------------------- problem
  struct Atom
  {
    double r[3];
    int    Z;
  };

  struct MeshSpec
  {
    int    p       = 3;
    int    n[3]    = {8, 8, 8};
    double lo[3]   = {-10, -10, -10};
    double hi[3]   = {10, 10, 10};
    int    refLo[3] = {0, 0, 0}; // coarse cells [refLo, refHi) refined once
    int    refHi[3] = {0, 0, 0};
    int    nq      = 0;          // Gauss points / dir, classical cells (0 -> p+1)
    int    nqEnr   = 0;          // Gauss points / dir, enriched cells (0 -> 2(p+1))
    std::vector<Atom> atoms;     // potential centres (and enrichment centres)
    double vSmear  = 0.5;        // V(r) = -Z erf(r / vSmear) / r
    bool   enrich  = false;      // attach 1s,2s,2p Slater enrichment per atom
    double rcut    = 4.0;        // smooth cutoff start
    double smooth  = 2.0;        // support ends at rcut (1 + 1/smooth)
    bool   orthogonalize = false;// L2-orthogonalize enrichment against the classical space
    bool   keepBasis = false;    // keep basis values and JxW at the quadrature points
    // test potential: V(x) = vAffine[0] + vAffine[1..3] . x in place of the
    // atoms' potential (pins the potential block of H_e in the tests)
    bool   affineV = false;
    double vAffine[4] = {0, 0, 0, 0};
  };

  // Flattened constraints, same arrays as
  // src/basis/CFEConstraintsLocalDealii.h:160-170 (rows ascending).
  struct Constraints
  {
    std::vector<u32>    rows, rowSizes, rowOffsets, cols;
    std::vector<double> weights, inhom;
  };

  struct Problem
  {
    int p = 0;
    u32 nCells = 0, nDofs = 0, nClassical = 0, nEnriched = 0;
    std::vector<u32>    cellNdofs, cellMapOffset, cellMap, cellNclassical;
    std::vector<u64>    cellMatOffset;            // offset of cell e's n_e^2 block
    std::vector<double> H, M, K;                  // per-cell, column-major
    std::vector<double> nodeX;                    // 3 * nClassical
    std::vector<double> cellLo, cellH;            // 3 * nCells each
    std::vector<int>    cellLevel;
    Constraints         cons;
    std::vector<double> mgll;                     // assembled GLL mass diag (raw)
    std::vector<double> dinv;                     // GLL overlap inverse diagonal
    std::vector<double> enrMinv;                  // nEnriched^2 dense (M^ee)^-1, col-major
    std::vector<double> alpha;                    // nEnriched x nClassical projection coefficients
    std::vector<int>    isHanging, isDirichlet;   // per classical node
    // with MeshSpec::keepBasis: per cell e, basis values N_k(x_q) at
    // basisF[basisOff[e] + k * nQ_e + q] (the final, orthogonalized basis)
    // and JxW at quadW[quadOff[e] + q]; nQ_e = quadOff[e + 1] - quadOff[e]
    std::vector<double> basisF, quadW;
    std::vector<u64>    basisOff, quadOff;
  };

  Problem buildProblem(const MeshSpec &spec);

  // enrichment function value / gradient (atom-centred, quantum index k in 0..4)
  void enrichmentValueGrad(const Atom &a, int k, double rcut, double smooth,
                           const double *x, double &val, double *grad);
  double potentialAt(const std::vector<Atom> &atoms, double smear, const double *x);

  // ---------------------------------------------------------- operators
  // A cell-wise operator in the reference's sense: Y = C^T (sum_e Q_e^T A_e Q_e) C X
  struct CellOp
  {
    const Problem *     prob;
    const double *      mats; // per-cell column-major blocks at prob->cellMatOffset
  };

  void distributeParentToChild(const Constraints &c, double *x, u32 B);
  void distributeChildToParent(const Constraints &c, double *y, u32 B);
  void setConstrainedNodes(const Constraints &c, double *x, u32 B, double v);

  // electron density at the quadrature points (needs keepBasis):
  // rho[quadOff[e] + q] = sum_i 2 occ[i] psi_i(x_q)^2,
  // psi_i(x_q) = sum_k N_k(x_q) X[map_e[k], i]
  void density(const Problem &P, const double *X, u32 B, const double *occ, double *rho);
  // integrals of quadrature values against the basis (needs keepBasis):
  // out[map_e[k], c] += sum_q JxW(q) f(q, c) N_k(x_q), cells in order; no
  // constraint step
  void integrate(const Problem &P, const double *f, u32 B, double *out);
  // values of nodal vectors at the quadrature points (needs keepBasis), no
  // constraint step
  void interpolate(const Problem &P, const double *X, u32 B, double *out);

  void gatherCells(const Problem &P, const double *x, u32 B, double *xCell);
  void scatterAddCells(const Problem &P, const double *yCell, u32 B, double *y);
  void cellApply(const Problem &P, const double *mats, double *X, double *Y, u32 B);
  void diagInvApply(const Problem &P, double *X, double *Y, u32 B);
  void enrichedInvApply(const Problem &P, double *X, double *Y, u32 B);
  void lumpedMassApply(const Problem &P, double *X, double *Y, u32 B);

  // ---------------------------------------------------------- dense LA
  void gemm(char ta, char tb, int m, int n, int k, double alpha, const double *A,
            int lda, const double *B, int ldb, double beta, double *C, int ldc);
  int  potrfLower(int n, double *A, int lda);
  int  trtriLower(int n, double *A, int lda);
  int  syevJacobi(int n, double *A, int lda, double *w, double *Z, int ldz);

  // ---------------------------------------------------------- solvers
  enum Precision
  {
    FP64 = 0
  };

  // operator selection for the solver drivers
  enum OpKind
  {
    OP_H = 0,
    OP_M = 1,
    OP_MINV = 2,
    OP_MLUMP = 3,
    OP_K = 4
  };
  void applyOp(const Problem &P, int kind, double *X, double *Y, u32 B);

  void chebyshevFilter(const Problem &P, double *X, u32 B, int m, double a,
                       double b, double c, double *Y);
  void residualChebyshevFilter(const Problem &P, const double *lambda, double *X,
                               u32 B, int m, double a, double b, double c,
                               double *Y);
  int choleskyGramSchmidt(const Problem &P, double *X, u32 n, double *Xo);
  int rayleighRitz(const Problem &P, double *X, u32 n, double *eig, double *Xo);
  int chfsiSolve(const Problem &P, double a, double b, double c, int m,
                 bool residualFilter, u32 batch, double *eig, double *guess,
                 double *X, u32 n);
  int lanczosExtreme(const Problem &P, const double *guess, int maxKrylov,
                     double *lo, double *hi, double *betaLast);
  void eigenResidualNorms(const Problem &P, const double *X, const double *eig,
                          u32 n, double *res);
  int cgSolve(const Problem &P, int opKind, const double *diagInv,
              const double *b, double *x, u32 nRhs, double absTol,
              double relTol, double divTol, u32 maxIter, u32 *iters);
} // namespace oracle
