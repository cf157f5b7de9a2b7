// Host-side engine of libamls_hip.so: owns the device memory of one problem
// and sequences the batched kernels level by level.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "host/tree.hpp"
#include "kernels/common.hpp"
#include "kernels/launch.hpp"

namespace amls {

struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define AMLS_HIP_CHECK(expr)                                                                    \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::amls::StatusError(e_ == hipErrorOutOfMemory ? 5 : 4,                        \
                                      std::string("HIP error '") + hipGetErrorString(e_) +      \
                                          "' at " #expr);                                       \
    } while (0)

// One hipMalloc per phase; sub-allocations aligned to 256 bytes.
class DevArena {
public:
    DevArena() = default;
    DevArena(const DevArena&) = delete;
    DevArena& operator=(const DevArena&) = delete;
    ~DevArena();
    void plan(size_t bytes) { planned_ += (bytes + 255) & ~size_t(255); }
    // start planning a new layout; commit() keeps the current block when it
    // is large enough, so a phase run again does not reallocate
    void begin() { planned_ = used_ = 0; }
    void commit();   // allocate exactly what was planned
    template <class T> T* take(size_t count) {
        const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        if (used_ + bytes > cap_) throw StatusError(7, "device arena overflow");
        T* p = reinterpret_cast<T*>(base_ + used_);
        used_ += bytes;
        return p;
    }
    void release();
    size_t capacity() const { return cap_; }
    const void* base() const { return base_; }

private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0, planned_ = 0;
};

// Host staging of descriptor arrays, uploaded in one copy per phase.
class DevArena;
DevArena& fresh(std::unique_ptr<DevArena>& a);   // create or begin()

class Blob {
public:
    template <class T> size_t add(const std::vector<T>& v) {
        size_t off = (host_.size() + 63) & ~size_t(63);
        host_.resize(off + v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(host_.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
    void upload(hipStream_t s);
    void clear() { host_.clear(); }   // new contents; the device buffer is kept
    // upload into caller-owned device memory of `cap` bytes instead of an
    // allocation of the blob's own (upload then fails with AMLS_E_ARG when
    // the descriptors do not fit)
    void useExternal(void* dev, size_t cap) {
        ext_ = true;
        dev_ = static_cast<char*>(dev);
        devcap_ = cap;
    }
    size_t size() const { return host_.size(); }
    template <class T> const T* at(size_t off) const {
        return reinterpret_cast<const T*>(dev_ + off);
    }
    ~Blob();

private:
    std::vector<char> host_;
    char* dev_ = nullptr;
    size_t devcap_ = 0;
    bool ext_ = false;
};

Blob& fresh(std::unique_ptr<Blob>& b);   // create or clear()

// Split-k table of the batched GEMM: rows (output tiles, k, split), measured
// on the device by Ctx::autotune or installed by the caller; GemmBatch::add
// takes the split of the nearest row for a long-k product, or its built-in
// rule while the calling context's table is empty.
using SplitTable = std::vector<std::array<int64_t, 3>>;
extern thread_local const SplitTable* t_splitTable;   // table of the context in the current call

struct GemmBatch {
    std::vector<dev::GemmDesc> descs;
    std::vector<dev::TileRef> tiles;
    int forceSplit = 0;   // > 0: split every eligible product this many ways (measurement)
    void add(const dev::GemmDesc& d);
    bool empty() const { return tiles.empty(); }
};

// Per kernel-class launch statistics: HIP events around every launch of the
// class on the engine's stream (only while enabled), plus the algorithmic
// flops and bytes of the launch computed from its descriptors.
struct KernelStats {
    bool enabled = false;
    // byPhase: entries are keyed "phase/class" instead of "class", with the
    // phase the engine is in when the launch runs
    bool byPhase = false;
    // byShape (with byPhase): GEMM batches are keyed by their largest m, n, k,
    // descriptor count and operand form as well
    bool byShape = false;
    const char* phase = "";
    struct Entry {
        int64_t launches = 0;
        double ms = 0.0, flops = 0.0, bytes = 0.0;
    };
    std::map<std::string, Entry> byName;
    // trace: every launch is bracketed by events (as when enabled) and its
    // span kept on the host's steady-clock microseconds, through a base
    // event synchronised once; lane 1 is the engine's main stream (lane1),
    // lane 2 any other stream
    bool trace = false;
    hipStream_t lane1 = nullptr;
    hipEvent_t base = nullptr;
    int64_t baseUs = 0;
    struct Span {
        std::string cat, name, phase;
        int tid;
        int64_t t0, t1;
    };
    std::vector<Span> spans;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
        const char* phase;
        int tid;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t get();
    // record a launch; when enabled, `launch` is bracketed with events
    // `launch` may issue `count` back-to-back launches of the same kernel
    // (one event pair brackets them all)
    void run(const std::string& name, double flops, double bytes, hipStream_t s,
             const std::function<void()>& launch, int64_t count = 1);
    // a host span (lane 0) in the current phase, kept when tracing
    void hostSpan(const char* name, std::chrono::steady_clock::time_point a,
                  std::chrono::steady_clock::time_point b);
    // turns finished launches' event pairs into times (and trace spans);
    // wait: also those still running
    void resolve(bool wait);
    void reset();
    ~KernelStats();
};

// A sparse symmetric operator for Y = A X: CSR (rp, ci, vx; CSC of a
// symmetric matrix is its CSR), plus a 3 x 3 block form (nb > 0) when every
// aligned row triple shares one pattern of aligned column triples.
struct SparseOp {
    int64_t n = 0;
    const int64_t *rp = nullptr, *ci = nullptr;
    const double* vx = nullptr;
    double nnz = 0.0;
    // 3 x 3 blocks grouped by four block rows (see launch_spmm_bsr3q)
    int64_t nb = 0, nq = 0;
    const int64_t *qptr = nullptr, *qvptr = nullptr;
    const uint32_t* ucol = nullptr;
    const double* qval = nullptr;
};

// Work list of the batched tridiagonalisation (every active tile of every
// step), kept on the device between solves of the same problem: it depends
// only on the matrix orders.
struct SytrdCache {
    std::vector<int> ns;
    int4* work = nullptr;
    size_t cap = 0;
    std::vector<size_t> start;
    std::vector<int> count;
    ~SytrdCache() {
        if (work) (void)hipFree(work);
    }
};

struct Ctx {
    int device = -1;
    SytrdCache syLocal, syReduced;
    bool gpu = false;
    hipStream_t stream = nullptr;
    // second stream (default priority) for independent work forked from
    // `stream` within one plan, and the fork / join events
    hipStream_t stream2 = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    // The root separator's tridiagonalisation left running on stream2 past
    // the end of spectralPrepare (a solve's pass A takes the root's few
    // modes from a copy of its C by the Chebyshev filter meanwhile):
    // evChain marks its end; joinChain makes the main stream wait for it
    // (pass B reads the tridiagonal form), drainChain waits on the host
    // (before anything may release the spectral arena).
    hipEvent_t evChain = nullptr;
    bool chainPending = false;
    // meanwhile the other nodes' reductions run on stream3 (evBatch marks
    // their end), so the filter on the main stream overlaps them too;
    // spectralSelect joins them before it reads their tridiagonal forms
    hipStream_t stream3 = nullptr;
    hipEvent_t evFork3 = nullptr, evBatch = nullptr;
    bool batchPending = false;
    int64_t chainRoot = -1;
    double* dCroot = nullptr;
    int64_t overlapK0 = 0;       // > 0: the next spectralPrepare may leave the root's chain running
    void joinChain();   // also joins the batch
    void joinBatch();
    void drainChain();  // both, on the host
    bool silEligible(int64_t p) const;
    // stream2 is free for a plan's side work (no root reduction or batch of
    // the condensation still on the side streams).  With per-launch
    // statistics on, or serialSchedule (environment AMLS_SERIAL_SCHEDULE at
    // creation, for profilers), the panel chains and the forward sweep keep
    // to one stream, so that each launch's measured duration is its own.
    bool serialSchedule = false;
    bool pipelineOk() const {
        return !serialSchedule && !kstats.enabled && stream2 && stream3 && !chainPending &&
               !batchPending;
    }
    std::string err;

    // ---- problem (host)
    int64_t n = 0;
    double shift = 0.0;
    Symbolic S;
    std::vector<int64_t> owner;
    bool haveProblem = false;
    int partMethod = 0;   // recursiveBisection's method for the next set-up
    // host set-up of the last setProblems, seconds: canonical columns, graph,
    // nested dissection and symbolic tables; device tables (assembly maps,
    // permuted CSR, 3 x 3 block layouts); upload; total
    double setupTimes[4] = {0.0, 0.0, 0.0, 0.0};
    // independent problems solved together (a forest): problem s owns rows
    // [segRow[s], segRow[s+1]) in both orderings and post-order nodes
    // [segNode[s], segNode[s+1]); one problem is the case nseg = 1
    int64_t nseg = 1;
    std::vector<int64_t> segRow, segNode, nodeSeg;

    // ---- problem (device)
    std::unique_ptr<DevArena> probMem;
    int64_t *dsrcK = nullptr, *dtgtK = nullptr, *dsrcM = nullptr, *dtgtM = nullptr;
    int64_t nasmK = 0, nasmM = 0;
    int64_t *dperm = nullptr, *dbidx = nullptr, *drel = nullptr;
    int64_t* dSegRow = nullptr;   // segRow on the device
    // K and M as operators, original and permuted ordering (3 x 3 blocks when
    // the pattern has them)
    SparseOp opPK, opPM;   // K and M in the tree-permuted ordering
    std::unique_ptr<DevArena> bsrMem;
    // permuted CSR of K and M (for the subspace iteration)
    int64_t *dPKp = nullptr, *dPKi = nullptr, *dPMp = nullptr, *dPMi = nullptr;
    double *dPKx = nullptr, *dPMx = nullptr;

    // ---- elimination
    std::unique_ptr<DevArena> geMem;
    std::unique_ptr<Blob> geBlob;
    std::vector<int64_t> fo, xo, wo;   // element offsets per post-order node
    int64_t ftot = 0, xtot = 0, wtot = 0;
    double *dF = nullptr, *dG = nullptr, *dXb = nullptr, *dFm = nullptr, *dW = nullptr;
    double *dLinv = nullptr, *dDinv = nullptr;   // explicit L^-1 per node, D^-1
    int* dErr = nullptr;
    bool geDone = false;
    int gePrec = 0;
    // GE's launch sequence (every level) as a HIP graph: the first run for a
    // problem, precision and arena address is eager, the second captures the
    // graph, later ones replay it; runs with kernel statistics on stay eager.
    // geGraphPrec / Gen / Base: the key of the last run.
    hipGraphExec_t geGraph = nullptr;
    int geGraphPrec = -1;
    uint64_t problemGen = 0, geGraphGen = 0;
    const void* geGraphBase = nullptr;
    void dropGeGraph();
    void finishGE();

    // ---- local shift-invert Lanczos (nodes that EigenpairSolver's density
    // switch sends to sil::*, src/amls/eigenpair_solver.hpp:38-82)
    int silMode = 0;                 // silMode of EigenpairSolver: 0 switch, 1 always, -1 never
    double* dStat = nullptr;         // per node (nnz, max |diag|) of K-hat, written by GE
    std::vector<double> statHost;    // its host copy after GE
    struct SilNode {
        int64_t p;                   // post-order node
        int n, cap;                  // order, Lanczos capacity min(n, 1536)
        double sigma;                // -max(1e-10, 1e-9 max |K-hat_jj|)
        double *S, *Mf, *X, *U, *alpha, *beta, *u0, *dmhi, *dai, *dhalf, *Wp;
        double *W, *C, *part;        // Lanczos step scratch: w, coefficients, partial sums
        int32_t* state;              // (steps done, breakdown flag) on the device
        int mdone = 0;
        bool broke = false;
    };
    std::vector<SilNode> sil;        // SIL-eligible nodes of the current factorisation
    std::vector<int> silOf;          // node -> index into sil, or -1
    std::vector<int32_t> lastPath;   // per node, last spectral call: 0 dense, 1 SIL, 2 SIL fell back
    std::vector<int64_t> lastSteps;  // per node, Lanczos steps the last call used
    // the last selection: kind (-1 none, 2 pinned), its values per problem and
    // the Lanczos steps whose Ritz pairs each node kept (0: dense path)
    int lastSelKind = -1;
    std::vector<double> lastSelValue;
    std::vector<int> lastMUse;
    // smallest relative distance of a node's last kept or first dropped local
    // eigenvalue to the cut of the last threshold selection, and its node
    void truncationMargin(double& margin, int64_t& node);
    std::unique_ptr<DevArena> silMem;
    std::unique_ptr<Blob> silBlob;
    void silPrepare();               // S = R^T (K-hat - sigma M-hat)^-1 R per SIL node
    void silAdvance(const std::vector<int>& idx, const std::vector<int>& ncv);
    void silSelect(int kind, const std::vector<double>& valueSeg, std::vector<int>& mUse);

    // ---- local spectra
    std::unique_ptr<DevArena> spMem;
    std::unique_ptr<Blob> spBlob;
    // In a solve, the other nodes' reductions (stream3) and the root's
    // (stream2) as two HIP graphs, with descriptor blobs of their own: the
    // first run for a problem, arena address and root is eager, the second
    // captures them, later ones replay them (runs with statistics or tracing
    // stay eager).
    std::unique_ptr<Blob> spBlobB, spBlobC;
    hipGraphExec_t spGraphB = nullptr, spGraphC = nullptr;
    uint64_t spGraphGen = 0;
    const void* spGraphBase = nullptr;
    int64_t spGraphBig = -1;
    bool spKeySeen = false;
    void dropSpGraphs();
    void launchSpectralSide(const std::vector<dev::EigDesc>& er, const std::vector<int64_t>& sr,
                            const dev::EigDesc& eroot, int64_t rootOff, int64_t big);
    std::vector<int64_t> co, vo;       // n_p^2 and n_p offsets
    double *dC = nullptr, *dT = nullptr, *dD = nullptr, *dE = nullptr, *dTau = nullptr;
    double *dSv = nullptr, *dSp = nullptr, *dDsq = nullptr;
    int64_t* dSoff = nullptr;
    bool specReady = false;

    // ---- truncation pass
    std::unique_ptr<DevArena> passMem;
    std::unique_ptr<Blob> passBlob;
    std::vector<int64_t> k, moff, mstart;   // per node
    int64_t ptot = 0;
    std::vector<int64_t> segMoff, segP, mto;   // per problem: first mode, p, Mtilde offset
    std::vector<int64_t> zo, ro;            // Z and R-front offsets
    double *dLam = nullptr, *dY = nullptr, *dZ = nullptr, *dMu = nullptr, *dSw = nullptr;
    double *dR = nullptr, *dMt = nullptr;
    std::unique_ptr<DevArena> projMem;
    std::unique_ptr<Blob> passBlob2;
    bool selected = false, projected = false;

    // ---- reduced problem
    std::unique_ptr<DevArena> redMem;
    std::unique_ptr<Blob> redBlob;
    int64_t want = 0;
    double* dQ = nullptr;
    std::vector<double> theta;              // nseg x want, problem-major
    bool reduced = false, redVectors = false;

    // ---- recovered vectors (permuted order, n x want) and output
    std::unique_ptr<DevArena> vecMem;
    std::unique_ptr<Blob> vecBlob;
    double *dZfull = nullptr, *dXout = nullptr;
    bool recovered = false;

    // ---- subspace iteration
    std::unique_ptr<DevArena> siMem, outMem;
    int64_t outElems = 0;        // doubles outBuf holds (n x nev of the solve that sized it)
    double* outBuf = nullptr;
    std::unique_ptr<Blob> siBlob, siBlob2;
    // The steps and the set-up of the final Rayleigh-Ritz problem as a HIP
    // graph, keyed by everything their launch sequence depends on: the tree
    // (problemGen), q, nev, the step count, the start, the shift, the node
    // mode counts of a whole-basis start, and the addresses of the buffers
    // the launches touch.  First run eager, second captures, later replay.
    struct SiKey {
        uint64_t gen = 0;
        int64_t q = -1, nev = -1;
        int steps = 0;
        bool fromBasis = false;
        double shift = 0.0;
        std::vector<const void*> ptrs;
        std::vector<int64_t> k;
        bool operator==(const SiKey& o) const {
            return gen == o.gen && q == o.q && nev == o.nev && steps == o.steps &&
                   fromBasis == o.fromBasis && shift == o.shift && ptrs == o.ptrs && k == o.k;
        }
    };
    SiKey siKey;
    bool siKeySeen = false;
    hipGraphExec_t siGraph = nullptr;
    void dropSiGraph();

    // ---- top eigenpairs of the final Rayleigh-Ritz problem by Chebyshev-
    // filtered subspace iteration (chfsi.cpp); the dense path stays the
    // fallback
    std::unique_ptr<DevArena> chfMem;
    std::unique_ptr<Blob> chfBlob;
    int rrMode = 0;          // 0: filtered iteration where it applies, 1: dense only
    int lastRrIters = 0;     // filter iterations of the last final Rayleigh-Ritz, 0: dense path
    double lastRrResidual = 0.0;
    bool chfsiTop(double* C, int64_t q, int64_t m, double* lam, double* V);   // C made symmetric from its lower triangle

    // Chrome trace of host phases and GPU launches (trace.cpp)
    std::string tracePath;
    void traceHost(const char* cat, const char* name, std::chrono::steady_clock::time_point a,
                   std::chrono::steady_clock::time_point b);
    int64_t traceWrite(const std::string& path);

    SplitTable splitTable;   // empty: the built-in split-k rule
    void autotune();

    std::map<std::string, double> timing;
    int64_t lastWant = 0;
    bool lastVectors = false;    // dXout holds lastWant M-normalised eigenvectors
    KernelStats kstats;

    // phases
    struct ProblemIn {
        int64_t n;
        const int64_t *Kp, *Ki;
        const double* Kx;
        const int64_t *Mp, *Mi;
        const double* Mx;
        const double* coords;
    };
    void setProblems(const std::vector<ProblemIn>& probs, int nLevels, int64_t leafSize,
                     double shift);
    void setProblem(int64_t n, const int64_t* Kp, const int64_t* Ki, const double* Kx,
                    const int64_t* Mp, const int64_t* Mi, const double* Mx, const double* coords,
                    int nLevels, int64_t leafSize, double shift) {
        setProblems({ProblemIn{n, Kp, Ki, Kx, Mp, Mi, Mx, coords}}, nLevels, leafSize, shift);
    }
    void runGE(int prec);
    void spectralPrepare();
    // kind 0: keep min(value, n_i) modes per node (value <= 0 keeps all);
    // kind 1: keep modes with w <= value; one value per problem.
    // pinned (non-null, one count per node): keep exactly min(pinned[p], n_p)
    // modes of every node from the dense reduction (kind and values unused)
    void spectralSelect(int kind, const std::vector<double>& valuePerSeg,
                        const std::vector<int64_t>* pinned = nullptr);
    void spectralSelect(int kind, double value) {
        spectralSelect(kind, std::vector<double>(static_cast<size_t>(nseg), value));
    }
    void project(int prec);
    void reducedSolve(int64_t want, bool vectors);
    void backtransform();
    void subspaceIter(int steps, int64_t q, int64_t nev, std::vector<double>& thetaOut,
                      bool fromBasis);
    void sync();        // the main stream; statistics of finished launches
    void syncAll();     // every stream of the context; all statistics
    void requireGpu() const {
        if (!gpu) throw StatusError(6, "context has no device");
    }

};

// Operator-granular entry points (operators.cpp).
// A reusable dense LDL^T factor of P A P^T on the device (perm[new] = old).
struct DenseFactor {
    int device = -1;
    int64_t n = 0;
    std::unique_ptr<DevArena> mem;
    double* LD = nullptr;        // L strict lower, D on the diagonal
    double* dq = nullptr;        // n doubles of solve scratch
    int32_t *p32 = nullptr, *ip32 = nullptr;   // perm and its inverse
    std::vector<int64_t> perm;
};
std::vector<int64_t> eigenLdltOrder(const std::vector<double>& diag);
void gemmGather(Ctx& c, int precision, int ta, int tb, int64_t m, int64_t n, int64_t k,
                double alpha, const double* A, int64_t lda, int64_t aColsStored,
                const int32_t* aRows, const int32_t* aCols, const double* B, int64_t ldb,
                int64_t bColsStored, const int32_t* bRows, const int32_t* bCols, double beta,
                double* C, int64_t ldc);
void ldltFactor(Ctx& c, DenseFactor& f, int64_t n, const int64_t* cp, const int64_t* ri,
                const double* vx, int pivoting);
void ldltFactorSolve(Ctx& c, const DenseFactor& f, int64_t nrhs, const double* B, double* X);
void ldltFactorGet(Ctx& c, const DenseFactor& f, double* LD, int64_t* perm);
// block width of the filtered iteration for m of the largest eigenpairs of a
// q x q matrix, 0 where it does not apply (chfsi.cpp)
int64_t chfsiBlock(int64_t q, int64_t m);
void denseEig(Ctx& c, int64_t n, const double* K, const double* M, int range, double lo, double hi,
              int64_t il, int64_t iu, int64_t* mOut, double* w, double* V);

// The allocation-free factor-and-solve (noalloc.cpp): scratchInit records,
// with `device` current, the device address of its file-scope scratch array.
hipError_t scratchInit(int device);
size_t noallocScratchBytes();
void ldltSolveNoalloc(Ctx& c, int64_t n, double* A, int64_t nrhs, double* B);

}  // namespace amls
