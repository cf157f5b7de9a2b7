// Local shift-invert Lanczos (SIL) on the device: the set-up of each SIL
// node's symmetric operator and the resumable Lanczos runs.
//
// EigenpairSolver (src/amls/eigenpair_solver.hpp:38-82) sends a local
// pencil (K-hat, M-hat) of order n >= 200 whose K-hat has fewer than
// 0.15 n^2 non-zeros to sil::solveIndex / solveInterval
// (src/amls/shift_invert_lanczos.hpp:126-203), an M-inner-product Lanczos
// on C = (K-hat - sigma M-hat)^-1 M-hat.  Here, with A = K-hat - sigma M-hat
// = L_A D_A L_A^T and M-hat = R R^T (R = L_M D_M^1/2), the same process runs
// as standard Lanczos on S = R^T A^-1 R = G^T G, G = D_A^-1/2 L_A^-1 R, in
// u = R^T x coordinates: the start vector is u_0 = R^T q_0 (so |u_0|^2 =
// q_0^T M q_0, the reference's normalisation) and Ritz vectors map back as
// x = R^-T u = L_M^-T D_M^-1/2 u.
//
// K-hat itself is rebuilt from its LDL^T factor left by GE (L D L^T); the
// node's non-zero count and largest |diagonal|, which decide the switch and
// sigma, were taken from the front before it was factored.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "plan.hpp"

namespace amls {

namespace {
// std::srand(seed) + Eigen's Vec::Random(n), as glibc's rand() and Eigen
// (<= 3.4) produce them: r_i = r_{i-31} + r_{i-3} (mod 2^32) after a
// Park-Miller seeding of 31 words and 310 discarded outputs, rand() = r_i >> 1,
// coefficient -1 + 2 rand() / RAND_MAX (shift_invert_lanczos.hpp:49-50).
std::vector<double> silStartVector(int64_t n, uint32_t seed = 12345u) {
    std::vector<uint32_t> r(static_cast<size_t>(344 + n));
    r[0] = seed ? seed : 1u;
    for (int i = 1; i < 31; ++i) {
        const int64_t hi = static_cast<int64_t>(r[i - 1]) / 127773;
        const int64_t lo = static_cast<int64_t>(r[i - 1]) % 127773;
        int64_t w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        r[i] = static_cast<uint32_t>(w);
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (size_t i = 34; i < r.size(); ++i) r[i] = r[i - 31] + r[i - 3];
    std::vector<double> v(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        const double x = static_cast<double>(r[344 + i] >> 1);
        v[i] = -1.0 + (2.0 * x) / 2147483647.0;
    }
    return v;
}
}  // namespace

void Ctx::silPrepare() {
    const int64_t N = S.nNodes;
    sil.clear();
    silOf.assign(N, -1);
    lastPath.assign(N, 0);
    lastSteps.assign(N, 0);
    if (silMode < 0) return;
    int64_t nmax = 0;
    for (int64_t p = 0; p < N; ++p) {
        // the density switch; nodes above the Lanczos kernels' order limit
        // stay on the dense path
        if (!silEligible(p)) continue;
        const int64_t n = S.size[p];
        const double dmax = statHost[2 * p + 1];
        SilNode s{};
        s.p = p;
        s.n = static_cast<int>(n);
        s.cap = static_cast<int>(std::min<int64_t>(n, 1536));
        s.sigma = -std::max(1e-10, 1e-9 * (dmax > 0.0 ? dmax : 1.0));
        silOf[p] = static_cast<int>(sil.size());
        sil.push_back(s);
        nmax = std::max(nmax, n);
    }
    if (sil.empty()) return;
    auto& A = fresh(silMem);
    // one plan() per take() below (each rounds to 256 bytes)
    for (const SilNode& s : sil) {
        const int64_t n = s.n;
        for (int q = 0; q < 3; ++q) A.plan(n * n * 8);    // S (A first), Mf, X
        A.plan(n * s.cap * 8);                            // U
        for (int q = 0; q < 2; ++q) A.plan(s.cap * 8);    // alpha, beta
        for (int q = 0; q < 4; ++q) A.plan(n * 8);        // u0, dmhi, dai, dhalf
        A.plan(n * 64 * 8);                               // panel scratch
        A.plan(16);                                       // state
        A.plan(n * 8);                                    // W
        A.plan(s.cap * 8);                                // C
        A.plan(64 * 8);                                   // part
    }
    A.plan(nmax * 8);                   // q0
    A.plan(256);                        // error word
    A.commit();
    for (SilNode& s : sil) {
        const int64_t n = s.n;
        s.S = A.take<double>(n * n);
        s.Mf = A.take<double>(n * n);
        s.X = A.take<double>(n * n);
        s.U = A.take<double>(n * s.cap);
        s.alpha = A.take<double>(s.cap);
        s.beta = A.take<double>(s.cap);
        s.u0 = A.take<double>(n);
        s.dmhi = A.take<double>(n);
        s.dai = A.take<double>(n);
        s.dhalf = A.take<double>(n);
        s.Wp = A.take<double>(n * 64);
        s.state = A.take<int32_t>(4);
        s.W = A.take<double>(n);
        s.C = A.take<double>(s.cap);
        s.part = A.take<double>(64);
        s.mdone = 0;
        s.broke = false;
    }
    double* q0 = A.take<double>(nmax);
    int* err = A.take<int>(1);
    const std::vector<double> q0h = silStartVector(nmax);
    AMLS_HIP_CHECK(hipMemcpyAsync(q0, q0h.data(), nmax * 8, hipMemcpyHostToDevice, stream));
    const int big = INT_MAX;
    AMLS_HIP_CHECK(hipMemcpyAsync(err, &big, sizeof(int), hipMemcpyHostToDevice, stream));
    for (SilNode& s : sil)
        AMLS_HIP_CHECK(hipMemsetAsync(s.state, 0, 4 * sizeof(int32_t), stream));

    fresh(silBlob);
    Plan plan(silBlob.get(), &kstats);
    // A = K-hat - sigma M-hat = (L D^1/2)(L D^1/2)^T + |sigma| M-hat, into S
    std::vector<VecDesc> v1;
    Copy2dBatch c1, c2;
    GemmBatch g1;
    std::vector<double*> fac;
    std::vector<int64_t> fn;
    std::vector<double*> fw;
    std::vector<int> fnode;
    for (SilNode& s : sil) {
        const int64_t p = s.p, n = s.n, f = S.f(p);
        double* F = dF + fo[p];
        double* G = dG + fo[p];
        VecDesc v;
        std::memset(&v, 0, sizeof(v));
        v.src = F;
        v.ld = f;
        v.dst = s.dhalf;
        v.n = s.n;
        v.op = 6;
        v1.push_back(v);
        Copy2dDesc cx = cdesc(F, f, s.X, n, n, n, 4, nullptr, s.dhalf);   // X = L D^1/2
        c1.add(cx);
        c1.add(cdesc(G, f, s.S, n, n, n, 2, nullptr, nullptr, -s.sigma));  // |sigma| M-hat
        c1.add(cdesc(G, f, s.Mf, n, n, n, 2));                             // M-hat
        GemmDesc g = gdesc();
        g.A = s.X;
        g.lda = n;
        g.triA = 1;
        g.B = s.X;
        g.ldb = n;
        g.tb = 1;
        g.k = s.n;
        g.C = s.S;
        g.ldc = n;
        g.m = g.n = s.n;
        g.beta = 1.0;
        g.lower = 1;
        g1.add(g);
        fac.push_back(s.S);
        fac.push_back(s.Mf);
        fn.push_back(n);
        fn.push_back(n);
        fw.push_back(s.Wp);
        fw.push_back(s.X);   // X is free again once A is formed (its n x 64 head)
        fnode.push_back(static_cast<int>(p));
        fnode.push_back(static_cast<int>(p));
    }
    plan.vec(v1);
    plan.copy(c1);
    plan.gemm(g1, false);
    // the two factorisations: A = L_A D_A L_A^T and M-hat = L_M D_M L_M^T
    addLdltBatch(plan, fac, fn, fw, err, fnode);
    // R = L_M D_M^1/2 into X; u0 = R^T q0; X = D_A^-1/2 L_A^-1 R (= G);
    // S = G^T G (G lower triangular), completed to full storage
    std::vector<VecDesc> v2;
    GemmBatch gu, gs;
    TrsmBatch tr;
    Copy2dBatch sc, cs;
    for (SilNode& s : sil) {
        const int64_t n = s.n;
        VecDesc v;
        std::memset(&v, 0, sizeof(v));
        v.n = s.n;
        v.src = s.Mf;
        v.ld = n;
        v.dst = s.dhalf;     // D_M^1/2 (the K-hat factor's D^1/2 is no longer needed)
        v.op = 6;
        v2.push_back(v);
        v.dst = s.dmhi;      // D_M^-1/2
        v.op = 0;
        v2.push_back(v);
        v.src = s.S;
        v.dst = s.dai;       // D_A^-1/2
        v2.push_back(v);
        c2.add(cdesc(s.Mf, n, s.X, n, n, n, 4, nullptr, s.dhalf));
        GemmDesc g = gdesc();                 // u0 = R^T q0
        g.A = s.X;
        g.lda = n;
        g.ta = 1;
        g.triA = 2;
        g.B = q0;
        g.ldb = nmax;
        g.k = s.n;
        g.C = s.u0;
        g.ldc = n;
        g.m = s.n;
        g.n = 1;
        gu.add(g);
        tr.add(s.S, n, s.X, n, n, n, 0);      // X = L_A^-1 R
        tr.descs.back().rhsLower = 1;
        sc.add(cdesc(s.X, n, s.X, n, n, n, 0, s.dai));
        GemmDesc h = gdesc();                 // S = X^T X (lower)
        h.A = s.X;
        h.lda = n;
        h.ta = 1;
        h.triA = 2;
        h.B = s.X;
        h.ldb = n;
        h.k = s.n;
        h.C = s.S;
        h.ldc = n;
        h.m = h.n = s.n;
        h.lower = 1;
        gs.add(h);
        cs.add(cdesc(s.S, n, s.S, n, n, n, 2));
    }
    plan.vec(v2);
    plan.copy(c2);
    plan.gemm(gu, false);
    plan.trsm(tr);
    plan.copy(sc);
    plan.gemm(gs, false);
    plan.copy(cs);
    plan.run(stream);
    int e = 0;
    AMLS_HIP_CHECK(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, stream));
    sync();
    if (e != INT_MAX)
        throw StatusError(AMLS_E_NOT_SPD, "shift-invert Lanczos: K-hat - sigma M-hat or M-hat of node " +
                                              std::to_string(S.treeId[e]) + " (post-order " +
                                              std::to_string(e) + ") is not SPD");
}

// Runs the Lanczos of SIL nodes idx[i] until ncv[i] steps are done (or a
// breakdown); the host copies of the step counts are updated.
void Ctx::silAdvance(const std::vector<int>& idx, const std::vector<int>& ncv) {
    std::vector<dev::LanczosDesc> ds;
    int nmax = 1, steps = 0, jtop = 0;
    for (size_t i = 0; i < idx.size(); ++i) {
        SilNode& s = sil[idx[i]];
        const int target = std::min(ncv[i], s.cap);
        if (s.broke || s.mdone >= target) continue;
        dev::LanczosDesc d;
        std::memset(&d, 0, sizeof(d));
        d.S = s.S;
        d.U = s.U;
        d.alpha = s.alpha;
        d.beta = s.beta;
        d.u0 = s.u0;
        d.state = s.state;
        d.n = s.n;
        d.ncv = target;
        d.W = s.W;
        d.C = s.C;
        d.part = s.part;
        ds.push_back(d);
        nmax = std::max(nmax, s.n);
        steps = std::max(steps, target - s.mdone);
        jtop = std::max(jtop, target - 1);
    }
    if (ds.empty()) return;
    Blob blob;
    const size_t od = blob.add(ds);
    blob.upload(stream);
    const int cnt = static_cast<int>(ds.size());
    double flops = 0.0, bytes = 0.0;
    for (size_t i = 0; i < idx.size(); ++i) {
        const SilNode& s = sil[idx[i]];
        const int target = std::min(ncv[i], s.cap);
        if (s.broke || s.mdone >= target) continue;
        for (int j = s.mdone; j < target; ++j) {
            flops += 2.0 * s.n * s.n + 8.0 * s.n * (j + 1);
            bytes += 8.0 * (double(s.n) * s.n + 4.0 * s.n * (j + 1));
        }
    }
    // one step of every active node per round; a node that has reached its
    // target or broken down drops out inside the kernels
    std::vector<int> mstart;
    for (size_t i = 0; i < idx.size(); ++i) {
        const SilNode& s = sil[idx[i]];
        if (!(s.broke || s.mdone >= std::min(ncv[i], s.cap))) mstart.push_back(s.mdone);
    }
    // round k: node i is at step mdone_i + k at most
    const int mmax = *std::max_element(mstart.begin(), mstart.end());
    kstats.run("lanczos", flops, bytes, stream, [&] {
        const dev::LanczosDesc* d = blob.at<dev::LanczosDesc>(od);
        launch_lanczos_init(d, cnt, stream);
        for (int k = 0; k < steps; ++k)
            launch_lanczos_step(d, cnt, nmax, std::min(jtop, mmax + k), stream);
    }, std::max(steps, 1));
    std::vector<int32_t> st(4 * sil.size());
    for (size_t i = 0; i < sil.size(); ++i)
        AMLS_HIP_CHECK(hipMemcpyAsync(st.data() + 4 * i, sil[i].state, 2 * sizeof(int32_t),
                                      hipMemcpyDeviceToHost, stream));
    sync();
    for (size_t i = 0; i < sil.size(); ++i) {
        sil[i].mdone = st[4 * i];
        sil[i].broke = st[4 * i + 1] != 0;
    }
}

// EigenpairSolver's choice for every SIL node in one truncation pass, and
// the number of pairs kept (k holds the dense counts on entry):
//  - FixedCount(k0 > 0), the index hint [1, k0] (solveLocalIndex,
//    eigenpair_solver.hpp:51-63): SIL when k0 < 4n/5, with
//    ncv = min(n, max(4 k0 + 16, 48)); its k0 lowest Ritz pairs unless the
//    run broke down before k0 steps (then dense);
//  - Threshold(hi), the interval (-1e300, hi] (solveLocalInterval, :38-47,
//    solveInterval, shift_invert_lanczos.hpp:126-168): ncv = min(n, 96),
//    doubled up to min(n, 1536) while no Ritz value or every Ritz value
//    lies below hi (the residual test reads the never-stored beta_{m-1}, so
//    it always passes), keeping every Ritz value <= hi; none kept: dense;
//  - FixedCount(0) (no hint, solveLocalAll): dense.
// mUse[p] > 0: node p keeps the Ritz pairs of the first mUse[p] Lanczos steps.
void Ctx::silSelect(int kind, const std::vector<double>& valueSeg, std::vector<int>& mUse) {
    const int64_t N = S.nNodes;
    mUse.assign(N, 0);
    std::fill(lastPath.begin(), lastPath.end(), 0);
    std::fill(lastSteps.begin(), lastSteps.end(), 0);
    if (sil.empty()) return;
    const int ns = static_cast<int>(sil.size());
    if (kind == 0) {
        std::vector<int> idx, ncv;
        std::vector<int64_t> kk;
        for (int i = 0; i < ns; ++i) {
            const SilNode& s = sil[i];
            const double v = valueSeg[nodeSeg[s.p]];
            if (v <= 0) continue;
            const int64_t kq = static_cast<int64_t>(v);
            if (!(kq < (int64_t(s.n) * 4) / 5)) continue;
            idx.push_back(i);
            ncv.push_back(static_cast<int>(std::min<int64_t>(s.n, std::max<int64_t>(4 * kq + 16, 48))));
            kk.push_back(kq);
        }
        silAdvance(idx, ncv);
        for (size_t t = 0; t < idx.size(); ++t) {
            const SilNode& s = sil[idx[t]];
            const int m = std::min(s.mdone, ncv[t]);
            if (std::min<int64_t>(kk[t], m) == kk[t]) {
                mUse[s.p] = m;
                k[s.p] = kk[t];
                lastPath[s.p] = 1;
                lastSteps[s.p] = m;
            } else {
                lastPath[s.p] = 2;
                lastSteps[s.p] = m;
            }
        }
        return;
    }
    std::vector<int> idx, ncv, cnt(ns, 0), mOf(ns, 0);
    for (int i = 0; i < ns; ++i) {
        idx.push_back(i);
        ncv.push_back(std::min(sil[i].n, 96));
    }
    while (!idx.empty()) {
        silAdvance(idx, ncv);
        // Sturm counts of nu >= 1 / (hi - sigma) on T_m: the Ritz values
        // lambda = 1 / nu + sigma <= hi
        std::vector<EigDesc> ed;
        std::vector<double> thr;
        int nmaxc = 1;
        for (size_t t = 0; t < idx.size(); ++t) {
            const SilNode& s = sil[idx[t]];
            const int m = std::min(s.mdone, ncv[t]);
            EigDesc e;
            std::memset(&e, 0, sizeof(e));
            e.d = s.alpha;
            e.e = s.beta;
            e.n = m;
            e.kcap = m;
            ed.push_back(e);
            const double hi = valueSeg[nodeSeg[s.p]];
            thr.push_back(hi - s.sigma > 0.0 ? 1.0 / (hi - s.sigma) : INFINITY);
            nmaxc = std::max(nmaxc, m);
        }
        Blob cb;
        std::vector<int32_t> c(idx.size(), 0);
        const size_t od = cb.add(ed), ot = cb.add(thr), oc = cb.add(c);
        cb.upload(stream);
        launch_eig_count(cb.at<EigDesc>(od), static_cast<int>(ed.size()), cb.at<double>(ot),
                         const_cast<int32_t*>(cb.at<int32_t>(oc)), nmaxc, stream);
        AMLS_HIP_CHECK(hipMemcpyAsync(c.data(), cb.at<int32_t>(oc), c.size() * sizeof(int32_t),
                                      hipMemcpyDeviceToHost, stream));
        sync();
        std::vector<int> nidx, nncv;
        for (size_t t = 0; t < idx.size(); ++t) {
            const SilNode& s = sil[idx[t]];
            const int m = std::min(s.mdone, ncv[t]);
            cnt[idx[t]] = c[t];
            mOf[idx[t]] = m;
            if (c[t] < m && c[t] > 0) continue;          // an interval edge was seen
            if (ncv[t] >= s.cap || m < ncv[t]) continue;  // capacity reached, or breakdown
            nidx.push_back(idx[t]);
            nncv.push_back(std::min(s.cap, 2 * ncv[t]));
        }
        idx.swap(nidx);
        ncv.swap(nncv);
    }
    for (int i = 0; i < ns; ++i) {
        const SilNode& s = sil[i];
        lastSteps[s.p] = mOf[i];
        if (cnt[i] > 0) {
            mUse[s.p] = mOf[i];
            k[s.p] = cnt[i];
            lastPath[s.p] = 1;
        } else {
            lastPath[s.p] = 2;
        }
    }
}

}  // namespace amls
