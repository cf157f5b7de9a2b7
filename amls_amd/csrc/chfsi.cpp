// Top eigenpairs of the final Rayleigh-Ritz problem of the subspace
// iteration by Chebyshev-filtered subspace iteration.
//
// The final Rayleigh-Ritz step (the last step of AMLS-SIMb, the subspace-
// iteration paper's Rayleigh-Ritz on span(Y)) needs the nev largest
// eigenpairs of C = D^-1/2 L^-1 (Y^T M Y) L^-T D^-1/2, an explicit dense
// q x q matrix (q = p = 1,499 at the 26,460-dof hexahedral workload).  A full
// Householder reduction of C is q dependent launches; this solver needs a
// few dozen GEMMs with C instead:
//   - a block X of s columns (s = min(192, max(2 nev, nev + 48)) rounded up
//     to 16) starts from a reproducible uniform block (columns normalised)
//     and one power step, orthonormalised;
//     the Rayleigh quotient of its last column estimates the s-th eigenvalue
//     and sets the cut nu_c of the filter;
//   - each iteration applies the Chebyshev polynomial T_d of
//     x = 2 nu / nu_c - 1, which is at most 1 on [0, nu_c] and grows fast
//     above it, through the three-term recurrence with Cs = (4 / nu_c) C - 2 I
//     (X_1 = Cs X_0 / 2, X_j = Cs X_(j-1) - X_(j-2)), then normalises the
//     columns and orthonormalises the block by Cholesky QR (G = X^T X =
//     L D L^T, X <- X L^-T D^-1/2; twice in the first filtered iteration,
//     whose block is the worst conditioned).  Cholesky
//     QR keeps the columns in order, as orthogonal iteration does, so after
//     the first iterations each column leans towards its own eigenvector and
//     the filter's growth over the block stays within what one Cholesky QR
//     resolves; the degree therefore starts at the largest of 1..3 with
//     T_d(x_1) <= 1e8 (x_1 from the first column's Rayleigh quotient) and
//     rises by one per iteration up to 5;
//   - Rayleigh-Ritz on the block (H = X^T C X, s x s, the library's dense
//     symmetric eigensolver) gives the Ritz pairs, and the residuals
//     || C v - nu v || of the nev wanted ones are checked against
//     1e-11 nu_1 on the host; more iterations follow while they are not met,
//     up to 14 in all.
// Any failure (a non-positive Cholesky pivot, residuals not met, a value
// that is not finite) returns false and the caller takes the dense path, so
// the result is always one of the two.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "plan.hpp"

namespace amls {

namespace {

double chebT(int d, double x) {   // T_d(x), x >= 1
    double t0 = 1.0, t1 = x;
    if (d == 0) return 1.0;
    for (int k = 2; k <= d; ++k) {
        const double t2 = 2.0 * x * t1 - t0;
        t0 = t1;
        t1 = t2;
    }
    return t1;
}

GemmDesc prod(const double* A, int64_t lda, int ta, const double* B, int64_t ldb, int tb, double* Cp,
              int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta, int lower = 0) {
    GemmDesc g = gdesc();
    g.A = A;
    g.lda = lda;
    g.ta = ta;
    g.B = B;
    g.ldb = ldb;
    g.tb = tb;
    g.C = Cp;
    g.ldc = ldc;
    g.m = static_cast<int32_t>(m);
    g.n = static_cast<int32_t>(n);
    g.k = static_cast<int32_t>(k);
    g.alpha1 = alpha;
    g.beta = beta;
    g.lower = lower;
    return g;
}

// split > 1: the product is split over k into that many pieces (few output
// tiles against a long k)
void addGemm(Plan& plan, const GemmDesc& g, int split = 0) {
    GemmBatch b;
    b.forceSplit = split;
    b.add(g);
    plan.gemm(b, false);
}

}  // namespace

// Block width for nev wanted pairs of a q x q matrix, 0 where the filtered
// iteration does not apply (the block must leave 48 guard columns, fit the
// single-workgroup reduction of the Rayleigh-Ritz problem, and be small
// against q).
int64_t chfsiBlock(int64_t q, int64_t m) {
    int64_t s = std::max<int64_t>(2 * m, m + 48);
    s = std::min<int64_t>(192, (s + 15) / 16 * 16);
    if (s < m + 48 || 4 * s > q) return 0;
    return s;
}

bool Ctx::chfsiTop(double* C, int64_t q, int64_t m, double* lamOut, double* Vout) {
    lastRrIters = 0;
    lastRrResidual = 0.0;
    const int64_t s = chfsiBlock(q, m);
    if (s == 0) return false;
    auto& A = fresh(chfMem);
    for (int i = 0; i < 2; ++i) A.plan(q * s * 8);       // X blocks
    A.plan(q * q * 8);                                   // Cs
    for (int i = 0; i < 3; ++i) A.plan(s * s * 8);       // G, L^-1, H
    A.plan(2 * s * 64 * 8);                              // LDL^T panel scratch (two panels)
    for (int i = 0; i < 3; ++i) A.plan(s * 8);           // D^-1/2, column norms, Rayleigh quotients
    A.plan(2 * 8);                                       // row segment [0, q)
    for (int i = 0; i < 3; ++i) A.plan(s * 8);           // d, e, tau
    A.plan(2 * s * 8);
    A.plan(3 * s * 8);
    A.plan(s * m * 8);                                   // Ritz coefficients
    A.plan(6 * s * m * 8);                               // inverse-iteration scratch
    A.plan(q * m * 8);                                   // C V
    A.plan(m * 8);                                       // residuals
    A.plan(256);
    A.commit();
    double* Xp[2] = {A.take<double>(q * s), A.take<double>(q * s)};
    double* Cs = A.take<double>(q * q);
    double* G = A.take<double>(s * s);
    double* Li = A.take<double>(s * s);
    double* H = A.take<double>(s * s);
    double* Wp = A.take<double>(2 * s * 64);
    double* dq = A.take<double>(s);
    double* nrm = A.take<double>(s);
    double* rq = A.take<double>(s);
    int64_t* seg = A.take<int64_t>(2);
    double* ed_d = A.take<double>(s);
    double* ed_e = A.take<double>(s);
    double* ed_tau = A.take<double>(s);
    double* sv = A.take<double>(2 * s);
    double* sp = A.take<double>(3 * s);
    double* Ys = A.take<double>(s * m);
    double* wk = A.take<double>(6 * s * m);
    double* R = A.take<double>(q * m);
    double* res = A.take<double>(m);
    int* err = A.take<int>(1);
    const int64_t segH[2] = {0, q};
    AMLS_HIP_CHECK(hipMemcpyAsync(seg, segH, sizeof(segH), hipMemcpyHostToDevice, stream));
    const int si = static_cast<int>(s);
    const int64_t qq = q;
    int cur = 0;   // index of the block holding X

    auto normalize = [&](Plan& plan, double* X) {
        plan.fn([=](hipStream_t st) { launch_colnorm_seg(seg, 1, X, X, si, qq, nrm, st, false); });
    };
    // X <- X L^-T D^-1/2 with X^T X = L D L^T (T: scratch block); L^-1 by
    // the forward solve on the identity, which follows the factorisation
    // panel by panel on stream2 when that stream is free
    auto cholqr = [&](Plan& plan, double* X, double* T) {
        addGemm(plan, prod(X, q, 1, X, q, 0, G, s, s, s, q, 1.0, 0.0, 1), 8);
        Copy2dBatch cz;
        cz.add(cdesc(nullptr, 0, Li, s, s, s, 3, nullptr, nullptr, 0.0));
        plan.copy(cz);
        std::vector<VecDesc> v(1), vq(1);
        std::memset(v.data(), 0, sizeof(VecDesc));
        std::memset(vq.data(), 0, sizeof(VecDesc));
        v[0].dst = Li;
        v[0].ld = s;
        v[0].n = si;
        v[0].op = 2;
        plan.vec(v);
        TrsmBatch tb;
        tb.add(G, s, Li, s, s, s, 0);
        tb.descs.back().rhsLower = 1;
        const bool piped = pipelineOk();
        const InvPipeline pipe{&tb, stream2, evFork, evJoin, stream3, evFork3, evBatch};
        addLdltBatch(plan, {G}, {s}, {Wp}, err, {-3}, piped ? &pipe : nullptr);
        if (!piped) plan.trsm(tb);
        vq[0].src = G;
        vq[0].dst = dq;
        vq[0].ld = s;
        vq[0].n = si;
        vq[0].op = 0;
        plan.vec(vq);
        addGemm(plan, prod(X, q, 0, Li, s, 1, T, q, q, s, s, 1.0, 0.0));
        Copy2dBatch sc;
        sc.add(cdesc(T, q, X, q, q, s, 0, nullptr, dq));
        plan.copy(sc);
    };
    auto readErr = [&]() {
        int e = 0;
        AMLS_HIP_CHECK(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, stream));
        sync();
        return e;
    };

    // ---- start block, one power step, Rayleigh quotients of the columns
    std::vector<double> rqH(s);
    {
        Plan plan(&fresh(chfBlob), &kstats);
        double* X0 = Xp[0];
        double* X1 = Xp[1];
        // C from its lower triangle: the products below then apply the
        // symmetric matrix the dense reduction would read (the two triangles
        // of the formed C differ by its rounding)
        Copy2dBatch cs;
        cs.add(cdesc(C, q, C, q, q, q, 2));
        plan.copy(cs);
        plan.fn([=](hipStream_t st) {
            launch_set_i32(err, INT_MAX, st);
            launch_fill_hash(X0, qq * si, 0x2545F4914F6CDD1Dull, st);
        });
        // the random block only normalised (its columns are nearly
        // orthogonal already: cond about 2 for s << q), its power step
        // orthonormalised by one Cholesky QR (cond about 2 nu_1 / nu_s)
        normalize(plan, X0);
        addGemm(plan, prod(C, q, 0, X0, q, 0, X1, q, q, s, q, 1.0, 0.0), 3);
        normalize(plan, X1);
        cholqr(plan, X1, X0);
        addGemm(plan, prod(C, q, 0, X1, q, 0, X0, q, q, s, q, 1.0, 0.0), 3);
        plan.fn([=](hipStream_t st) { launch_coldot(qq, X1, X0, si, qq, rq, st); });
        plan.run(stream);
        AMLS_HIP_CHECK(hipMemcpyAsync(rqH.data(), rq, s * 8, hipMemcpyDeviceToHost, stream));
        cur = 1;
    }
    if (readErr() != INT_MAX) return false;
    const double cut = rqH[s - 1];
    double top = 0.0;
    for (double v : rqH) {
        if (!std::isfinite(v)) return false;
        top = std::max(top, v);
    }
    if (!(cut > 0.0) || !(top > cut)) return false;
    const double x1 = 2.0 * top / cut - 1.0;
    int deg = 3;
    while (deg > 1 && chebT(deg, x1) > 1e8) --deg;

    bool first = true;
    int iters = 0;
    std::vector<double> resH(m), lamH(m);
    for (int round = 0; round < 3; ++round) {
        Plan plan(&fresh(chfBlob), &kstats);
        if (first) {
            // Cs = (4 / nu_c) C - 2 I
            Copy2dBatch cc;
            cc.add(cdesc(C, q, Cs, q, q, q, 0, nullptr, nullptr, 4.0 / cut));
            plan.copy(cc);
            std::vector<VecDesc> v(1);
            std::memset(v.data(), 0, sizeof(VecDesc));
            v[0].dst = Cs;
            v[0].ld = q;
            v[0].n = static_cast<int32_t>(q);
            v[0].op = 8;
            v[0].a = -2.0;
            plan.vec(v);
        }
        const int nit = round == 0 ? 5 : 4;
        for (int it = 0; it < nit; ++it) {
            double* prev = Xp[cur];
            double* nxt = Xp[1 - cur];
            addGemm(plan, prod(Cs, q, 0, prev, q, 0, nxt, q, q, s, q, 0.5, 0.0), 3);
            // X_j = Cs X_(j-1) - X_(j-2), written over X_(j-2)
            for (int j = 2; j <= deg; ++j) {
                addGemm(plan, prod(Cs, q, 0, nxt, q, 0, prev, q, q, s, q, 1.0, -1.0), 3);
                std::swap(prev, nxt);
            }
            normalize(plan, nxt);
            cholqr(plan, nxt, prev);
            if (first) cholqr(plan, nxt, prev);
            first = false;
            cur = nxt == Xp[0] ? 0 : 1;
            deg = std::min(deg + 1, 5);
            ++iters;
        }
        // Rayleigh-Ritz on the block, Ritz vectors, residuals
        double* X = Xp[cur];
        double* T = Xp[1 - cur];
        addGemm(plan, prod(C, q, 0, X, q, 0, T, q, q, s, q, 1.0, 0.0), 3);
        addGemm(plan, prod(X, q, 1, T, q, 0, H, s, s, s, q, 1.0, 0.0, 1), 8);
        EigDesc ed;
        std::memset(&ed, 0, sizeof(ed));
        ed.A = H;
        ed.lda = s;
        ed.n = si;
        ed.d = ed_d;
        ed.e = ed_e;
        ed.tau = ed_tau;
        ed.lam = lamOut;
        ed.Y = Ys;
        ed.work = wk;
        ed.k = ed.kcap = static_cast<int32_t>(m);
        addSytrd(plan, {ed}, {0}, sv, sp);
        addSelect(plan, {ed}, true);
        addGemm(plan, prod(X, q, 0, Ys, s, 0, Vout, q, q, m, s, 1.0, 0.0));
        addGemm(plan, prod(C, q, 0, Vout, q, 0, R, q, q, m, q, 1.0, 0.0), 4);
        const int mi = static_cast<int>(m);
        plan.fn([=](hipStream_t st) { launch_ritz_resid(R, Vout, lamOut, qq, mi, res, st); });
        plan.run(stream);
        AMLS_HIP_CHECK(hipMemcpyAsync(resH.data(), res, m * 8, hipMemcpyDeviceToHost, stream));
        AMLS_HIP_CHECK(hipMemcpyAsync(lamH.data(), lamOut, m * 8, hipMemcpyDeviceToHost, stream));
        if (readErr() != INT_MAX) return false;
        double rmax = 0.0;
        bool finite = std::isfinite(lamH[0]) && lamH[0] > 0.0;
        for (double r : resH) {
            finite = finite && std::isfinite(r);
            rmax = std::max(rmax, r);
        }
        if (!finite) return false;
        lastRrIters = iters;
        lastRrResidual = rmax / lamH[0];
        if (rmax <= 1e-11 * lamH[0]) return true;
    }
    return false;
}

}  // namespace amls
