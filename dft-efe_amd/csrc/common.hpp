// Internal declarations of libdfx (MI355X / gfx950).  Not installed.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "dfx.h"

namespace dfx
{
  using u32 = uint32_t;
  using u64 = uint64_t;

  // ------------------------------------------------------------ errors
  struct Error : std::runtime_error
  {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
  };

#define DFX_HIP(call)                                                        \
  do                                                                         \
    {                                                                        \
      hipError_t e_ = (call);                                                \
      if (e_ != hipSuccess)                                                  \
        throw ::dfx::Error(e_ == hipErrorOutOfMemory ? DFX_E_OOM : DFX_E_HIP, \
                           std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                        \
  while (0)

#define DFX_REQUIRE(cond, msg)                                               \
  do                                                                         \
    {                                                                        \
      if (!(cond))                                                           \
        throw ::dfx::Error(DFX_E_ARG, msg);                                  \
    }                                                                        \
  while (0)

  // launch check (kernel launches report errors through hipGetLastError)
  inline void checkLaunch(const char *what)
  {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      throw Error(DFX_E_HIP, std::string("launch ") + what + ": " + hipGetErrorString(e));
  }

  // ------------------------------------------------------------ captures
  // Graph captures (one host thread per serving slot may capture while
  // others work) are serialized with the calls that can reach the legacy
  // stream or synchronize the device implicitly: allocation and release,
  // rocBLAS / rocSOLVER calls, handle and stream teardown.  Recursive: a
  // capture that allocates (it should not) does not deadlock.
  inline std::recursive_mutex &captureMutex()
  {
    // never destroyed: buffers released during static destruction still
    // find it
    static std::recursive_mutex *m = new std::recursive_mutex;
    return *m;
  }
  using CaptureLock = std::lock_guard<std::recursive_mutex>;

  // ------------------------------------------------------------ buffers
  // Owning device allocation (hipMalloc / hipFree).
  struct DevBuf
  {
    void  *p     = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes)
    {
      o.p     = nullptr;
      o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
      if (this != &o)
        {
          release();
          p       = o.p;
          bytes   = o.bytes;
          o.p     = nullptr;
          o.bytes = 0;
        }
      return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n)
    {
      release();
      if (n)
        {
          CaptureLock lk(captureMutex());
          DFX_HIP(hipMalloc(&p, n));
        }
      bytes = n;
    }
    // grows the buffer (contents lost); true when it was reallocated
    bool ensure(size_t n)
    {
      if (n <= bytes)
        return false;
      alloc(n);
      return true;
    }
    void release()
    {
      if (p)
        {
          CaptureLock lk(captureMutex());
          (void)hipFree(p);
        }
      p     = nullptr;
      bytes = 0;
    }
    template <class T>
    T *as() const
    {
      return static_cast<T *>(p);
    }
  };

  template <class T>
  inline void upload(DevBuf &b, const T *h, size_t n, hipStream_t s)
  {
    b.alloc(n * sizeof(T));
    if (n)
      {
        DFX_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
        DFX_HIP(hipStreamSynchronize(s));
      }
  }
  template <class T>
  inline void upload(DevBuf &b, const std::vector<T> &h, hipStream_t s)
  {
    upload(b, h.data(), h.size(), s);
  }

  // a view of N x B values with leading dimension ld (>= B); element (i, j)
  // at ptr[i*ld + j]
  template <class T>
  struct View
  {
    T  *p;
    u32 N, B, ld;
  };
} // namespace dfx

// --------------------------------------------------------------- handles
// A context lives until its creator has destroyed it and every object made on
// it (multivector, constraints, operator, solver) is gone: children hold a
// reference, so bindings may free objects in any order.
struct dfx_ctx
{
  std::atomic<int> refs{1};
  int            device = 0;
  hipStream_t    stream = nullptr;
  // second stream for the size classes of one colour, which run side by side
  // (forked from and joined back into `stream` with the two events)
  hipStream_t    stream2 = nullptr;
  hipEvent_t     forkEv = nullptr, joinEv = nullptr;
  bool           forkClasses = true; // false: every launch on `stream` (serving slots)
  rocblas_handle blas   = nullptr;
  // hipBLASLt state of the FP64 subspace rotations (handle, workspace, one
  // plan per shape), owned here and released by ltFree
  void          *lt     = nullptr;
  void         (*ltFree)(void *) = nullptr;
  std::string    err;
  dfx::DevBuf    scratchSmall; // host<->device small transfers (reductions)
  dfx::DevBuf    reducePartials;
  std::map<std::string, dfx::DevBuf> work; // named solver scratch
  // captured degree loops of the Chebyshev filter, by name; `key` holds the
  // operators (uid and buffer generation), buffers, sizes and scalars the
  // graph was captured with; `uids` the operators it launches on, so that a
  // destroyed operator's graphs are released with it
  struct Graph
  {
    hipGraphExec_t        exec = nullptr;
    std::vector<uint64_t> key;
    std::vector<uint64_t> uids;
  };
  std::map<std::string, Graph> graphs;
  uint64_t graphCaptures = 0, graphReplays = 0;
};

struct dfx_mv
{
  dfx_ctx *ctx   = nullptr;
  dfx::u32 N     = 0, B = 0;
  int      dtype = DFX_F64;
  void    *d     = nullptr;
  bool     owns  = false;
  size_t   bytes() const { return (size_t)N * B * (dtype == DFX_F64 ? 8 : 4); }
};

struct dfx_constraints
{
  std::atomic<int> refs{1}; // creator + every operator built on it
  dfx_ctx *ctx = nullptr;
  dfx::u32 nDofs = 0, nRows = 0, nEntries = 0, nParents = 0;
  bool homDirichlet = false; // every row is x = 0 (no parents, zero inhomogeneity)
  bool homogeneous  = true;  // every inhomogeneity is zero (P2C is linear)
  std::vector<dfx::u32> hRows, hRowOffsets, hCols, hParents, hParentOffsets,
    hParentRows, hParentEntries;
  std::vector<double> hWeights, hInhom;
  // device
  dfx::DevBuf rows, rowOffsets, cols, weights, inhom;           // P2C
  dfx::DevBuf parents, parentOffsets, parentRowDofs, parentWeights; // C2P
};

namespace dfx
{
  enum OpKind
  {
    OP_CELLWISE = 0,
    OP_DIAGONAL = 1
  };

  // per-cell launch record (one per cell, in launch order)
  struct CellMeta
  {
    u32 n;       // dofs of the cell
    u32 nc;      // classical dofs (rows scattered directly)
    u32 mapOff;  // offset into the cell->dof map
    u32 enrOff;  // first row of this cell's enriched partials in the yE buffer
    u64 fragOff64; // offset (elements) of the cell's FP64 fragment-ordered matrix
    u64 fragOff32; // offset (elements) of the cell's FP32 fragment-ordered matrix
    u32 rowLo;     // smallest classical dof of the cell (buffer base row)
    u32 pad_;
  };

  // one kernel launch: cells [begin, end) of the launch order, all with the
  // same colour and the same M-tile class
  struct Launch
  {
    u32 begin, end;
    u32 mtw;    // M tiles per wave (n <= 64 * mtw)
    u32 kpadMax;
    u64 sumN2;  // sum of n_e^2 over the launch's cells
    u64 sumN;   // sum of n_e
    u64 sumNc;  // sum of classical rows
    u64 sumFirst = 0; // classical rows stored on first touch (not read)
    u32 color    = 0; // launches of one colour touch disjoint classical dofs
  };

  // one profiled kernel launch
  struct ProfRec
  {
    hipEvent_t a, b;
    double     flops, bytes;
  };
} // namespace dfx

struct dfx_op
{
  dfx_ctx     *ctx  = nullptr;
  int          kind = dfx::OP_CELLWISE;
  // unique per operator object for the life of the process (cache keys)
  uint64_t     uid  = nextUid();
  // bumped whenever a device buffer that the operator's launches read or
  // write is (re)allocated (cell matrices, enriched partials, folded
  // diagonal): a captured graph holds those addresses, so it is valid only
  // for the generation it was captured in
  uint64_t     gen  = 0;
  static uint64_t nextUid()
  {
    static std::atomic<uint64_t> n{1};
    return n.fetch_add(1);
  }
  // diagonal of a diagonal operator gathered per cell slot (d[map[s]]) for
  // folding it into this cell-wise operator's gather; built for foldedUid
  dfx::DevBuf foldedDiag;
  uint64_t    foldedUid = 0;
  dfx::u32     nDofs = 0, nClassical = 0;
  dfx_constraints *cons = nullptr;
  // ----- cell-wise
  dfx::u32 nCells = 0, nColors = 0, maxN = 0;
  dfx::u32 maxRowSpan = 0; // max over cells of (largest - smallest classical dof + 1)
  dfx::u64 sumNe = 0, sumNe2 = 0;
  std::vector<dfx::u32> hCellNdofs, hMapOffset; // caller order
  std::vector<dfx::u64> hMatOffset;             // raw (caller) layout offsets
  std::vector<dfx::u32> hOrder;                 // launch order -> caller cell
  std::vector<dfx::CellMeta> hMeta;             // launch order
  std::vector<dfx::Launch> launches64, launches32;
  dfx::u64 fragLen64 = 0, fragLen32 = 0;
  dfx::DevBuf map, meta, frag64, frag32;
  dfx::DevBuf first;     // per launch position, 8 words: first-touch bits of classical rows
  // per launch position, 8 words: last-touch bits of the unconstrained
  // classical rows (the residual step fused into the cell apply), and the
  // rows such an apply leaves to the stand-alone step (constrained,
  // enriched, in no cell)
  dfx::DevBuf last, resRows;
  dfx::u32    nResRows = 0;
  bool fillY  = false;   // some classical dof lies in no cell: Y needs a zero fill
  bool hasF32 = false;
  // enriched partial reduction
  dfx::u32 nEnrSlots = 0;
  dfx::DevBuf enrSlotOffsets, enrSlots; // CSR over enriched dofs -> yE rows
  dfx::DevBuf yE;                       // scratch, grows with B
  // standalone gather / scatter (dof -> slots CSR, slots in cell order)
  dfx::DevBuf dofSlotOffsets, dofSlots;
  // ----- diagonal
  dfx::DevBuf diag, diag32, enrBlock, enrFrag64, enrFrag32;
  dfx::u32 nEnriched = 0;
  // ----- column-chunk widths picked by dfx_op_autotune: for scalar size s
  // and exactly B columns, launches of M-tile class mtw use 16 nt columns
  struct Tuned
  {
    dfx::u32 s, B, mtw;
    int      nt;
  };
  std::vector<Tuned> tuned;
  int tunedNT(dfx::u32 s, dfx::u32 B, dfx::u32 mtw) const
  {
    for (const Tuned &t : tuned)
      if (t.s == s && t.B == B && t.mtw == mtw)
        return t.nt;
    return 0;
  }
  // ----- profiling of the cell-apply launches
  bool profiling = false;
  int  profMask  = 3; // 1: FP64 launches, 2: FP32 launches
  dfx::u32 profEvery   = 1; // every k-th apply of the profiled kind is timed
  dfx::u64 profApplies = 0;
  std::vector<dfx::ProfRec> prof;
  std::vector<hipEvent_t>   profEvents; // owned; consecutive records share one
  dfx::u64 profLaunches = 0;
  double profMs = 0, profFlops = 0, profBytes = 0;
  ~dfx_op()
  {
    for (hipEvent_t e : profEvents)
      (void)hipEventDestroy(e);
  }
};
