// Launch plans of the engine: descriptor batches for the batched kernels and
// the Plan that turns them into a launch sequence on a stream (shared by the
// engine's phases).
#pragma once

#include <chrono>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/amls_hip.h"
#include "engine.hpp"

namespace amls {

using dev::Copy2dDesc;
using dev::EigDesc;
using dev::ExtAddDesc;
using dev::FrontDesc;
using dev::GatherDesc;
using dev::GemmDesc;
using dev::TileRef;
using dev::TrsmDesc;
using dev::TrsmWork;
using dev::VecDesc;

inline int cdiv(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

using Clock = std::chrono::steady_clock;
inline double secs(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

inline GemmDesc gdesc() {
    GemmDesc d;
    std::memset(&d, 0, sizeof(d));
    d.alpha1 = 1.0;
    d.alpha2 = 1.0;
    return d;
}

inline Copy2dDesc cdesc(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t rows,
                 int64_t cols, int mode, const double* rs = nullptr, const double* cs = nullptr,
                 double alpha = 1.0) {
    Copy2dDesc d;
    std::memset(&d, 0, sizeof(d));
    d.src = src;
    d.dst = dst;
    d.lds = lds;
    d.ldd = ldd;
    d.rows = static_cast<int32_t>(rows);
    d.cols = static_cast<int32_t>(cols);
    d.mode = mode;
    d.rs = rs;
    d.cs = cs;
    d.alpha = alpha;
    return d;
}

struct Copy2dBatch {
    std::vector<Copy2dDesc> descs;
    std::vector<int4> work;
    void add(const Copy2dDesc& d) {
        if (d.rows <= 0 || d.cols <= 0) return;
        const int id = static_cast<int>(descs.size());
        descs.push_back(d);
        for (int i = 0; i < cdiv(d.rows, 64); ++i)
            for (int j = 0; j < cdiv(d.cols, 64); ++j) work.push_back(make_int4(id, i, j, 0));
    }
};

struct ExtBatch {
    std::vector<ExtAddDesc> descs;
    std::vector<int4> work;
    void add(const ExtAddDesc& d) {
        if (d.rows <= 0 || d.cols <= 0) return;
        const int id = static_cast<int>(descs.size());
        descs.push_back(d);
        for (int i = 0; i < cdiv(d.rows, 64); ++i)
            for (int j = 0; j < cdiv(d.cols, 64); ++j) {
                if (d.symmetric && j > i) continue;
                work.push_back(make_int4(id, i, j, 0));
            }
    }
};

struct TrsmBatch {
    std::vector<TrsmDesc> descs;
    void add(const double* L, int64_t ldl, double* B, int64_t ldb, int64_t n, int64_t nrhs,
             int trans) {
        if (n <= 0 || nrhs <= 0) return;
        // the diagonal-block kernel addresses 64 columns of B with 32-bit
        // byte offsets
        if (ldb > 4000000) throw StatusError(AMLS_E_ARG, "leading dimension above 4e6 elements");
        TrsmDesc d;
        std::memset(&d, 0, sizeof(d));
        d.L = L;
        d.B = B;
        d.ldl = ldl;
        d.ldb = ldb;
        d.n = static_cast<int32_t>(n);
        d.nrhs = static_cast<int32_t>(nrhs);
        d.trans = trans;
        descs.push_back(d);
    }
};

struct GatherBatch {
    std::vector<GatherDesc> descs;
    std::vector<int4> work;
    void add(const GatherDesc& d) {
        if (d.rows <= 0 || d.cols <= 0) return;
        const int id = static_cast<int>(descs.size());
        descs.push_back(d);
        for (int i = 0; i < cdiv(d.rows, 64); ++i)
            for (int j = 0; j < cdiv(d.cols, 64); ++j) work.push_back(make_int4(id, i, j, 0));
    }
};

// A sequence of launches whose descriptor arrays live in one blob.
// algorithmic work of a GEMM batch: 2 m n k flops per product and
// (m k + k n + 2 m n) words, counting only the written triangle for `lower`
inline void gemmWork(const GemmBatch& b, bool fp32, double& flops, double& bytes) {
    flops = bytes = 0.0;
    const double w = fp32 ? 4.0 : 8.0;
    for (const auto& d : b.descs) {
        const double m = d.m, n = d.n;
        double mn = m * n;
        if (d.lower && m >= n) mn = m * n - n * (n - 1) / 2;
        const double k1 = d.A ? d.k : 0, k2 = d.A2 ? d.k2 : 0;
        // a triangular op(A) (m = k) meets only half of k on average
        const double k1eff = d.triA ? 0.5 * (k1 + 1) : k1;
        flops += 2.0 * mn * (k1eff + k2);
        bytes += ((d.triA ? 0.5 : 1.0) * m * k1 + m * k2 + (k1 + k2) * n + 2.0 * mn) * w;
    }
}

struct Plan {
    Blob* blob;
    KernelStats* st;
    std::vector<std::function<void(hipStream_t)>> steps;
    // traced runs record the host's planning ("build": from the plan's
    // creation or last run), the descriptor upload and the launch enqueue
    std::chrono::steady_clock::time_point born = std::chrono::steady_clock::now();
    Plan(Blob* b, KernelStats* k) : blob(b), st(k) {}

    void timed(const std::string& name, double flops, double bytes,
               std::function<void(hipStream_t)> f) {
        KernelStats* k = st;
        if (!k) {
            steps.push_back(std::move(f));
            return;
        }
        steps.push_back([=](hipStream_t s) { k->run(name, flops, bytes, s, [&] { f(s); }); });
    }

    void gemm(const GemmBatch& b, bool fp32) {
        if (b.empty()) return;
        // split-k products accumulate into C: clear it first when beta = 0,
        // scale it by beta first otherwise (beta != 1 only for full outputs)
        Copy2dBatch clr;
        for (const auto& d : b.descs)
            if (d.kchunk > 0 && d.beta == 0.0) {
                Copy2dDesc c = cdesc(nullptr, 0, d.C, d.ldc, d.m, d.n, 3, nullptr, nullptr, 0.0);
                c.pad = d.lower;
                clr.add(c);
            } else if (d.kchunk > 0 && d.beta != 1.0) {
                clr.add(cdesc(d.C, d.ldc, d.C, d.ldc, d.m, d.n, 0, nullptr, nullptr, d.beta));
            }
        copy(clr);
        const size_t od = blob->add(b.descs), ot = blob->add(b.tiles);
        const int nt = static_cast<int>(b.tiles.size());
        Blob* bl = blob;
        double fl, by;
        gemmWork(b, fp32, fl, by);
        std::string name = fp32 ? "gemm_f32" : "gemm_f64";
        if (st && st->byShape) {
            int mm = 0, nn = 0, kk = 0, tri = 0, low = 0, split = 0;
            for (const auto& d : b.descs) {
                mm = std::max(mm, d.m);
                nn = std::max(nn, d.n);
                kk = std::max(kk, std::max(d.k, d.k2));
                tri |= d.triA != 0;
                low |= d.lower != 0;
                split |= d.kchunk > 0;
            }
            name += " m" + std::to_string(mm) + " n" + std::to_string(nn) + " k" + std::to_string(kk) +
                    " x" + std::to_string(b.descs.size()) + (tri ? " tri" : "") + (low ? " low" : "") +
                    (split ? " split" : "") + " t" + std::to_string(b.tiles.size());
        }
        timed(name, fl, by, [=](hipStream_t s) {
            launch_gemm(bl->at<GemmDesc>(od), bl->at<TileRef>(ot), nt, fp32, s);
        });
    }
    void copy(const Copy2dBatch& b) {
        if (b.work.empty()) return;
        const size_t od = blob->add(b.descs), ow = blob->add(b.work);
        const int nw = static_cast<int>(b.work.size());
        Blob* bl = blob;
        double by = 0.0;   // each element read (unless a fill) and written once
        for (const auto& d : b.descs) by += double(d.rows) * d.cols * (d.src ? 16.0 : 8.0);
        timed("copy2d", 0.0, by, [=](hipStream_t s) {
            launch_copy2d(bl->at<Copy2dDesc>(od), bl->at<int4>(ow), nw, s);
        });
    }
    void ext(const ExtBatch& b) {
        if (b.work.empty()) return;
        const size_t od = blob->add(b.descs), ow = blob->add(b.work);
        const int nw = static_cast<int>(b.work.size());
        Blob* bl = blob;
        double fl = 0.0, by = 0.0;   // source read, target read and written
        for (const auto& d : b.descs) {
            const double e = double(d.rows) * d.cols * (d.symmetric ? 0.5 : 1.0);
            fl += e;
            by += 24.0 * e;
        }
        timed("extadd", fl, by, [=](hipStream_t s) {
            launch_extadd(bl->at<ExtAddDesc>(od), bl->at<int4>(ow), nw, s);
        });
    }
    // Blocked right-looking unit-triangular solve of every descriptor: per
    // 64-row block (in solve order) one launch of diagonal-block solves, then
    // one GEMM batch applying the solved block to the rows still to come.
    // The blocked solves of b, one 64-row block of every descriptor per step:
    // trsmSetup adds the descriptors, trsmStep(s) step s's diagonal-block
    // solves and the update of the rows after (forward) or before (backward)
    // them.  Step s of a forward solve reads only block column s of L.
    struct TrsmPlan {
        size_t od = 0;
        int maxnb = 0;
        double fl = 0.0, by = 0.0;
    };
    TrsmPlan trsmSetup(const TrsmBatch& b) {
        TrsmPlan t;
        if (b.descs.empty()) return t;
        t.od = blob->add(b.descs);
        for (const auto& d : b.descs) {
            t.maxnb = std::max(t.maxnb, cdiv(d.n, 64));
            const double n = d.n, r = d.nrhs;
            t.fl += n * (n - 1) * r;
            t.by += (n * (n + 1) / 2 + 2 * n * r) * 8.0;
        }
        return t;
    }
    void trsm(const TrsmBatch& b) {
        const TrsmPlan t = trsmSetup(b);
        for (int s = 0; s < t.maxnb; ++s) trsmStep(b, t, s);
    }
    void trsmStep(const TrsmBatch& b, const TrsmPlan& t, int s) {
        if (s >= t.maxnb) return;
        const size_t od = t.od;
        const double fl = t.fl, by = t.by;
        Blob* bl = blob;
        const int nd = static_cast<int>(b.descs.size());
        {
            std::vector<TrsmWork> work[2];   // by direction: forward, transposed
            GemmBatch gb;
            for (int id = 0; id < nd; ++id) {
                const TrsmDesc& d = b.descs[id];
                const int nb = cdiv(d.n, 64);
                if (s >= nb) continue;
                const int rb = d.trans ? nb - 1 - s : s;
                const int r0 = rb * 64, mv = std::min(64, d.n - r0);
                // a lower-triangular RHS is zero above each column block's
                // diagonal: only column blocks <= rb are nonzero in row block rb
                const int ncb = d.rhsLower && !d.trans ? std::min(cdiv(d.nrhs, 64), rb + 1)
                                                       : cdiv(d.nrhs, 64);
                for (int c = 0; c < cdiv(ncb, 4); ++c) work[d.trans ? 1 : 0].push_back({id, c, rb, ncb});
                const int ncols = std::min(d.nrhs, ncb * 64);
                GemmDesc g = gdesc();
                if (!d.trans && r0 + mv < d.n) {
                    // B(r0+mv:, :) -= L(r0+mv:, r0:r0+mv) X(r0:r0+mv, :)
                    g.A = d.L + (r0 + mv) + static_cast<int64_t>(r0) * d.ldl;
                    g.lda = d.ldl;
                    g.B = d.B + r0;
                    g.ldb = d.ldb;
                    g.C = d.B + (r0 + mv);
                    g.ldc = d.ldb;
                    g.m = d.n - (r0 + mv);
                } else if (d.trans && r0 > 0) {
                    // B(:r0, :) -= L(r0:r0+mv, :r0)^T X(r0:r0+mv, :)
                    g.A = d.L + r0;
                    g.lda = d.ldl;
                    g.ta = 1;
                    g.B = d.B + r0;
                    g.ldb = d.ldb;
                    g.C = d.B;
                    g.ldc = d.ldb;
                    g.m = r0;
                } else {
                    continue;
                }
                g.n = ncols;
                g.k = mv;
                g.alpha1 = -1.0;
                g.beta = 1.0;
                gb.add(g);
            }
            const size_t ow0 = blob->add(work[0]), ow1 = blob->add(work[1]);
            const int nw0 = static_cast<int>(work[0].size()), nw1 = static_cast<int>(work[1].size());
            const bool first = s == 0;
            timed("trsm", first ? fl : 0.0, first ? by : 0.0, [=](hipStream_t st) {
                launch_trsm_diag(bl->at<TrsmDesc>(od), bl->at<TrsmWork>(ow0), nw0, false, st);
                launch_trsm_diag(bl->at<TrsmDesc>(od), bl->at<TrsmWork>(ow1), nw1, true, st);
            });
            gemm(gb, false);
        }
    }
    void gather(const GatherBatch& b) {
        if (b.work.empty()) return;
        const size_t od = blob->add(b.descs), ow = blob->add(b.work);
        const int nw = static_cast<int>(b.work.size());
        Blob* bl = blob;
        double by = 0.0;   // each gathered element read and written once
        for (const auto& d : b.descs) by += double(d.rows) * d.cols * 16.0;
        timed("gather", 0.0, by, [=](hipStream_t s) {
            launch_gather(bl->at<GatherDesc>(od), bl->at<int4>(ow), nw, s);
        });
    }
    void vec(const std::vector<VecDesc>& v) {
        if (v.empty()) return;
        const size_t od = blob->add(v);
        const int nv = static_cast<int>(v.size());
        Blob* bl = blob;
        timed("vecop", 0.0, 0.0, [=](hipStream_t s) { launch_vecop(bl->at<VecDesc>(od), nv, s); });
    }
    // other launches (fills, column norms, permutations), timed as "misc"
    void fn(std::function<void(hipStream_t)> f) { timed("misc", 0.0, 0.0, std::move(f)); }
    // steps that time their own launches (or launch nothing)
    void hostStep(std::function<void(hipStream_t)> f) { steps.push_back(std::move(f)); }
    // side's steps run on s2 from this point of the plan on, concurrently
    // with the steps added here until join(); side shares this plan's blob
    void fork(Plan& side, hipStream_t s2, hipEvent_t ef, hipEvent_t ej) {
        auto st = std::make_shared<std::vector<std::function<void(hipStream_t)>>>(std::move(side.steps));
        side.steps.clear();
        steps.push_back([=](hipStream_t s) {
            AMLS_HIP_CHECK(hipEventRecord(ef, s));
            AMLS_HIP_CHECK(hipStreamWaitEvent(s2, ef, 0));
            for (auto& f : *st) f(s2);
            AMLS_HIP_CHECK(hipEventRecord(ej, s2));
        });
    }
    void join(hipEvent_t ej) {
        steps.push_back([=](hipStream_t s) { AMLS_HIP_CHECK(hipStreamWaitEvent(s, ej, 0)); });
    }
    void run(hipStream_t s) {
        const auto t1 = std::chrono::steady_clock::now();
        blob->upload(s);
        const auto t2 = std::chrono::steady_clock::now();
        for (auto& f : steps) f(s);
        AMLS_HIP_CHECK(hipGetLastError());
        const auto t3 = std::chrono::steady_clock::now();
        if (st && st->trace) {
            st->hostSpan("build", born, t1);
            st->hostSpan("upload", t1, t2);
            st->hostSpan("enqueue", t2, t3);
        }
        born = t3;
    }
};

// Tridiagonalisation of the matrices of `descs` (sv / sp scratch at the
// per-matrix offsets soff), its last steps in one register-resident launch.
void addSytrd(Plan& plan, const std::vector<EigDesc>& descs, const std::vector<int64_t>& soff,
              double* sv, double* sp, SytrdCache* cache = nullptr);
// Bisection, inverse iteration and (backTransform) back-application of the
// reflectors for descriptors with lam / Y / work / k set.
void addSelect(Plan& plan, const std::vector<EigDesc>& descs, bool vectors,
               bool backTransform = true);
// Y = A X for column-major X, Y (n x ncols) through the row-major scratch t1, t2.
void addSpmm(Plan& plan, const SparseOp& A, const double* X, double* Y, int64_t ncols, double* t1,
             double* t2);
// In-place LDL^T of the lower triangles of the square matrices As[m] (order
// and leading dimension ns[m]) in 64-column panels; Ws[m]: ns[m] x 64
// scratch.  A non-positive or non-finite pivot of matrix m sets *err to
// min(*err, nodes[m]).
// algorithmic flops / bytes of LDL^T panel s0 of a leading block of order n
// inside a matrix of order f (-1: f = n), added to flops / bytes
void ldltPanelWork(int64_t n, int64_t s0, double& flops, double& bytes, int64_t f = -1);
// With a pipeline: the forward solves of inv (their L being these factors,
// or null) run on s2, row block j as soon as panel j is factored,
// concurrently with the later panels.  With s3 as well (look-ahead), each
// panel's trailing update is split: the next panel's 64 columns on the
// plan's stream, the rest on s3 while the next panel is factored; the W
// scratch then holds two panels (2 x n x 64, alternating).  The plan joins
// both after the last panel.
struct InvPipeline {
    const TrsmBatch* inv = nullptr;
    hipStream_t s2 = nullptr;
    hipEvent_t ef = nullptr, ej = nullptr;
    hipStream_t s3 = nullptr;
    hipEvent_t ef3 = nullptr, eb = nullptr;
};
void addLdltBatch(Plan& plan, const std::vector<double*>& As, const std::vector<int64_t>& ns,
                  const std::vector<double*>& Ws, int* err, const std::vector<int>& nodes,
                  const InvPipeline* pipe = nullptr);
// LDL^T of the n x n lower triangle of dA in place and, for nrhs > 0,
// dX := A^-1 dX; dW: n x 64 scratch, dq: n doubles, *derr: INT_MAX unless a
// pivot is non-positive or non-finite (then 0).
void addLdltSolve(Plan& plan, double* dA, int64_t n, double* dX, int64_t nrhs, double* dW,
                  double* dq, int* derr);
// The two halves of addLdltSolve: the factorisation, and X := A^-1 X with a
// factor it left (L strict lower, D on the diagonal of dLD).
void addLdltFactor(Plan& plan, double* dA, int64_t n, double* dW, int* derr);
void addLdltSolveWith(Plan& plan, const double* dLD, int64_t n, double* dX, int64_t nrhs,
                      double* dq);

}  // namespace amls
