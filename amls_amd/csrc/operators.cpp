// Operator-granular entry points behind the reference's plugin interfaces,
// on host buffers (the reference's Eigen matrices):
//  - ColCompOps' gathered GEMMs (src/linalg/colcomp.hpp:124-423): a GEMM on
//    row- and column-gathered operands, gathered on the device;
//  - ISparseDirectSolver::factorize / solveWithFactor
//    (src/linalg/sparse_direct.hpp:13-64, .cpp:10-63): a factor kept on the
//    device, with Eigen LDLT's diagonal pivoting for the dense solver;
//  - IDenseEigensolver::solveAll / solveInterval / solveIndex
//    (src/linalg/dense_eigensolver.hpp:10-31, .cpp:11-133): K x = lambda M x
//    with M SPD, ascending lambda, M-orthonormal vectors.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>

#include "plan.hpp"

namespace amls {

namespace {

void up(Ctx& c, void* d, const void* h, size_t bytes) {
    if (bytes) AMLS_HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c.stream));
}

void addGather2d(Plan& plan, const double* src, int64_t lds, double* dst, int64_t ldd, int64_t rows,
                 int64_t cols, const int32_t* ri, const int32_t* ci) {
    if (rows <= 0 || cols <= 0) return;
    dev::Gather2dDesc d;
    std::memset(&d, 0, sizeof(d));
    d.src = src;
    d.dst = dst;
    d.ri = ri;
    d.ci = ci;
    d.lds = lds;
    d.ldd = ldd;
    d.rows = static_cast<int32_t>(rows);
    d.cols = static_cast<int32_t>(cols);
    std::vector<dev::Gather2dDesc> ds{d};
    std::vector<int4> work;
    for (int i = 0; i < cdiv(rows, 64); ++i)
        for (int j = 0; j < cdiv(cols, 64); ++j) work.push_back(make_int4(0, i, j, 0));
    const size_t od = plan.blob->add(ds), ow = plan.blob->add(work);
    const int nw = static_cast<int>(work.size());
    Blob* bl = plan.blob;
    plan.fn([=](hipStream_t s) { launch_gather2d(bl->at<dev::Gather2dDesc>(od), bl->at<int4>(ow), nw, s); });
}

void checkIndex(const int32_t* idx, int64_t count, int64_t bound, const char* what) {
    if (!idx) {
        if (count > bound)
            throw StatusError(AMLS_E_ARG, std::string(what) + ": operand smaller than the product");
        return;
    }
    for (int64_t i = 0; i < count; ++i)
        if (idx[i] < 0 || idx[i] >= bound)
            throw StatusError(AMLS_E_ARG, std::string(what) + ": index out of range");
}

}  // namespace

// C = alpha op(Ag) op(Bg) + beta C with Ag = A(aRows, aCols), Bg = B(bRows,
// bCols) (identity where a list is null); the stored A is lda x aCols.
void gemmGather(Ctx& c, int precision, int ta, int tb, int64_t m, int64_t n, int64_t k,
                double alpha, const double* A, int64_t lda, int64_t aColsStored,
                const int32_t* aRows, const int32_t* aCols, const double* B, int64_t ldb,
                int64_t bColsStored, const int32_t* bRows, const int32_t* bCols, double beta,
                double* C, int64_t ldc) {
    if (precision != 0 && precision != 1)
        throw StatusError(AMLS_E_ARG, "precision must be 0 (FP64) or 1 (FP32 MFMA)");
    if (m < 0 || n < 0 || k < 0 || aColsStored < 0 || bColsStored < 0)
        throw StatusError(AMLS_E_ARG, "negative size");
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX || lda > INT32_MAX || ldb > INT32_MAX)
        throw StatusError(AMLS_E_ARG, "gemm_gather: sizes must fit in 32 bits");
    const int64_t ar = ta ? k : m, ac = ta ? m : k;   // the gathered operands' shapes
    const int64_t br = tb ? n : k, bc = tb ? k : n;
    if (ldc < std::max<int64_t>(1, m)) throw StatusError(AMLS_E_ARG, "leading dimension too small");
    if (m == 0 || n == 0) return;
    if (!C || (k > 0 && (!A || !B))) throw StatusError(AMLS_E_ARG, "null matrix argument");
    if (k > 0) {
        checkIndex(aRows, ar, lda, "gemm_gather A rows");
        checkIndex(aCols, ac, aColsStored, "gemm_gather A columns");
        checkIndex(bRows, br, ldb, "gemm_gather B rows");
        checkIndex(bCols, bc, bColsStored, "gemm_gather B columns");
    }
    c.requireGpu();
    const bool gA = aRows || aCols, gB = bRows || bCols;
    DevArena mem;
    mem.plan(lda * aColsStored * 8);
    mem.plan(ldb * bColsStored * 8);
    mem.plan(ldc * n * 8);
    mem.plan(gA ? ar * ac * 8 : 0);
    mem.plan(gB ? br * bc * 8 : 0);
    for (int64_t q : {ar, ac, br, bc}) mem.plan(q * 4);
    mem.commit();
    double* dA = mem.take<double>(lda * aColsStored);
    double* dB = mem.take<double>(ldb * bColsStored);
    double* dC = mem.take<double>(ldc * n);
    double* dAg = gA ? mem.take<double>(ar * ac) : nullptr;
    double* dBg = gB ? mem.take<double>(br * bc) : nullptr;
    int32_t* ix[4];
    const int32_t* hx[4] = {aRows, aCols, bRows, bCols};
    const int64_t nx[4] = {ar, ac, br, bc};
    for (int q = 0; q < 4; ++q) {
        ix[q] = mem.take<int32_t>(nx[q]);
        if (hx[q]) up(c, ix[q], hx[q], nx[q] * 4);
    }
    if (k > 0) {
        up(c, dA, A, lda * aColsStored * 8);
        up(c, dB, B, ldb * bColsStored * 8);
    }
    up(c, dC, C, ldc * n * 8);
    Blob blob;
    Plan plan(&blob, &c.kstats);
    if (k > 0 && gA)
        addGather2d(plan, dA, lda, dAg, ar, ar, ac, aRows ? ix[0] : nullptr, aCols ? ix[1] : nullptr);
    if (k > 0 && gB)
        addGather2d(plan, dB, ldb, dBg, br, br, bc, bRows ? ix[2] : nullptr, bCols ? ix[3] : nullptr);
    if (k > 0) {
        GemmBatch gb;
        GemmDesc d = gdesc();
        d.A = gA ? dAg : dA;
        d.lda = gA ? std::max<int64_t>(ar, 1) : lda;
        d.ta = ta;
        d.B = gB ? dBg : dB;
        d.ldb = gB ? std::max<int64_t>(br, 1) : ldb;
        d.tb = tb;
        d.C = dC;
        d.ldc = ldc;
        d.m = static_cast<int32_t>(m);
        d.n = static_cast<int32_t>(n);
        d.k = static_cast<int32_t>(k);
        d.alpha1 = alpha;
        d.beta = beta;
        gb.add(d);
        plan.gemm(gb, precision != 0);
    } else {
        Copy2dBatch cb;   // no product: C = beta C
        cb.add(cdesc(dC, ldc, dC, ldc, m, n, 0, nullptr, nullptr, beta));
        plan.copy(cb);
    }
    plan.run(c.stream);
    AMLS_HIP_CHECK(hipMemcpyAsync(C, dC, ldc * n * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

// ---------------------------------------------------------------------------
// Reusable LDL^T factor
// ---------------------------------------------------------------------------
// Eigen's LDLT (the reference's DenseLDLTSolver) pivots on the largest
// remaining diagonal entry; its factorisation is left-looking, so the entries
// it compares are the input's own diagonal (each is updated only when its
// column is reached), and the pivot sequence is a selection on |A_ii|: at step
// k the first largest of positions k..n-1, swapped into k.  The factor is then
// the unpivoted LDL^T of P A P^T.  perm[new] = old.
std::vector<int64_t> eigenLdltOrder(const std::vector<double>& diag) {
    const int64_t n = static_cast<int64_t>(diag.size());
    std::vector<int64_t> pos(n);
    std::iota(pos.begin(), pos.end(), 0);
    std::vector<double> a(n);
    for (int64_t i = 0; i < n; ++i) a[i] = std::fabs(diag[i]);
    for (int64_t k = 0; k < n; ++k) {
        int64_t best = k;
        for (int64_t i = k + 1; i < n; ++i)
            if (a[i] > a[best]) best = i;
        std::swap(a[k], a[best]);
        std::swap(pos[k], pos[best]);
    }
    return pos;
}

void ldltFactor(Ctx& c, DenseFactor& f, int64_t n, const int64_t* cp, const int64_t* ri,
                const double* vx, int pivoting) {
    if (n <= 0) throw StatusError(AMLS_E_ARG, "ldlt_factor: empty matrix");
    if (n > 46340) throw StatusError(AMLS_E_ARG, "ldlt_factor: n^2 must fit in 31 bits");
    if (!cp || !ri || !vx) throw StatusError(AMLS_E_ARG, "null matrix argument");
    if (pivoting != 0 && pivoting != 1) throw StatusError(AMLS_E_ARG, "pivoting must be 0 or 1");
    if (cp[0] != 0) throw StatusError(AMLS_E_ARG, "column pointers must start at 0");
    for (int64_t j = 0; j < n; ++j)
        if (cp[j + 1] < cp[j]) throw StatusError(AMLS_E_ARG, "column pointers decrease");
    c.requireGpu();
    // the lower triangle (Eigen's LDLT<.., Lower> and SimplicialLDLT read only
    // that), duplicates summed, completed to full symmetric storage
    std::vector<double> A(static_cast<size_t>(n * n), 0.0);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t e = cp[j]; e < cp[j + 1]; ++e) {
            const int64_t i = ri[e];
            if (i < 0 || i >= n) throw StatusError(AMLS_E_ARG, "row index out of range");
            if (i >= j) A[i + j * n] += vx[e];
        }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = j + 1; i < n; ++i) A[j + i * n] = A[i + j * n];
    std::vector<double> dg(n);
    for (int64_t i = 0; i < n; ++i) dg[i] = A[i + i * n];
    f.perm = pivoting ? eigenLdltOrder(dg) : std::vector<int64_t>();
    if (!pivoting) {
        f.perm.resize(n);
        std::iota(f.perm.begin(), f.perm.end(), 0);
    }
    f.n = n;
    f.device = c.device;
    f.mem.reset(new DevArena());
    f.mem->plan(n * n * 8 * 2);
    f.mem->plan(n * 64 * 8);
    f.mem->plan(n * 8);
    f.mem->plan(n * 4 * 2);
    f.mem->plan(256);
    f.mem->commit();
    double* dIn = f.mem->take<double>(n * n);
    f.LD = f.mem->take<double>(n * n);
    double* dW = f.mem->take<double>(n * 64);
    f.dq = f.mem->take<double>(n);
    f.p32 = f.mem->take<int32_t>(n);
    f.ip32 = f.mem->take<int32_t>(n);
    int* derr = f.mem->take<int>(1);
    std::vector<int32_t> p32(n), ip32(n);
    for (int64_t i = 0; i < n; ++i) {
        p32[i] = static_cast<int32_t>(f.perm[i]);
        ip32[f.perm[i]] = static_cast<int32_t>(i);
    }
    up(c, dIn, A.data(), n * n * 8);
    up(c, f.p32, p32.data(), n * 4);
    up(c, f.ip32, ip32.data(), n * 4);
    Blob blob;
    Plan plan(&blob, &c.kstats);
    addGather2d(plan, dIn, n, f.LD, n, n, n, f.p32, f.p32);   // P A P^T
    addLdltFactor(plan, f.LD, n, dW, derr);
    plan.run(c.stream);
    int e = 0;
    AMLS_HIP_CHECK(hipMemcpyAsync(&e, derr, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (e != INT_MAX)
        throw StatusError(AMLS_E_NOT_SPD, "LDL^T: non-positive or non-finite pivot");
}

// X = A^-1 B = P^T L^-T D^-1 L^-1 P B with the factor of ldltFactor
void ldltFactorSolve(Ctx& c, const DenseFactor& f, int64_t nrhs, const double* B, double* X) {
    if (nrhs < 0) throw StatusError(AMLS_E_ARG, "negative nrhs");
    if (nrhs > INT32_MAX) throw StatusError(AMLS_E_ARG, "nrhs must fit in 32 bits");
    if (nrhs == 0) return;
    if (!B || !X) throw StatusError(AMLS_E_ARG, "null matrix argument");
    c.requireGpu();
    if (f.device != c.device) throw StatusError(AMLS_E_ARG, "factor belongs to another device");
    const int64_t n = f.n;
    DevArena mem;
    mem.plan(n * nrhs * 8 * 2);
    mem.commit();
    double* dB = mem.take<double>(n * nrhs);
    double* dX = mem.take<double>(n * nrhs);
    up(c, dB, B, n * nrhs * 8);
    Blob blob;
    Plan plan(&blob, &c.kstats);
    addGather2d(plan, dB, n, dX, n, n, nrhs, f.p32, nullptr);    // P B
    addLdltSolveWith(plan, f.LD, n, dX, nrhs, f.dq);
    addGather2d(plan, dX, n, dB, n, n, nrhs, f.ip32, nullptr);   // P^T (.)
    plan.run(c.stream);
    AMLS_HIP_CHECK(hipMemcpyAsync(X, dB, n * nrhs * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

void ldltFactorGet(Ctx& c, const DenseFactor& f, double* LD, int64_t* perm) {
    c.requireGpu();
    if (LD) AMLS_HIP_CHECK(hipMemcpyAsync(LD, f.LD, f.n * f.n * 8, hipMemcpyDeviceToHost, c.stream));
    if (perm) std::copy(f.perm.begin(), f.perm.end(), perm);
    c.sync();
}

// ---------------------------------------------------------------------------
// Dense generalized eigensolver
// ---------------------------------------------------------------------------
// K x = lambda M x through the Cholesky-type factor of M: M = L D L^T,
// S = -D^-1/2 L^-1 K L^-T D^-1/2, whose largest eigenvalues nu = -lambda are
// the lowest lambda; x = L^-T D^-1/2 y is M-orthonormal.  range 0: all
// (solveAll), 1: lambda in (lo, hi] (solveInterval, RANGE = 'V'), 2: the
// il-th to iu-th lowest, 1-based (solveIndex, RANGE = 'I', iu clamped to n).
void denseEig(Ctx& c, int64_t n, const double* K, const double* M, int range, double lo, double hi,
              int64_t il, int64_t iu, int64_t* mOut, double* w, double* V) {
    if (range < 0 || range > 2) throw StatusError(AMLS_E_ARG, "range must be 0, 1 or 2");
    if (n < 0) throw StatusError(AMLS_E_ARG, "negative size");
    if (n > 46340) throw StatusError(AMLS_E_ARG, "dense_eig: n^2 must fit in 31 bits");
    if (!mOut) throw StatusError(AMLS_E_ARG, "null count argument");
    *mOut = 0;
    if (range == 2 && n > 0 && il < 1) throw StatusError(AMLS_E_ARG, "il must be >= 1");
    // the reference's early returns (dense_eigensolver.cpp:49-53, 94-98)
    if (n == 0 || (range == 1 && !(hi > lo)) || (range == 2 && iu < il)) return;
    if (!K || !M || !w) throw StatusError(AMLS_E_ARG, "null matrix argument");
    c.requireGpu();
    const int64_t nn = n * n;
    DevArena mem;
    for (int q = 0; q < 5; ++q) mem.plan(nn * 8);          // Kin, Min (L), Li, T, S
    mem.plan(n * 64 * 8);
    for (int q = 0; q < 4; ++q) mem.plan(n * 8);          // dq, d, e, tau
    mem.plan(2 * n * 8);
    mem.plan(3 * n * 8);
    mem.plan(n * 8);                                      // lam
    mem.plan(nn * 8);                                     // Y
    mem.plan(6 * nn * 8);                                 // inverse-iteration scratch
    mem.plan(nn * 8);                                     // X
    mem.plan(256);
    mem.plan(256);
    mem.commit();
    double* dK = mem.take<double>(nn);
    double* dM = mem.take<double>(nn);
    double* Li = mem.take<double>(nn);
    double* T = mem.take<double>(nn);
    double* S = mem.take<double>(nn);
    double* Wp = mem.take<double>(n * 64);
    double* dq = mem.take<double>(n);
    double* d = mem.take<double>(n);
    double* e = mem.take<double>(n);
    double* tau = mem.take<double>(n);
    double* sv = mem.take<double>(2 * n);
    double* sp = mem.take<double>(3 * n);
    double* lam = mem.take<double>(n);
    double* Y = mem.take<double>(nn);
    double* wk = mem.take<double>(6 * nn);
    double* X = mem.take<double>(nn);
    int* derr = mem.take<int>(1);
    int32_t* cnt = mem.take<int32_t>(2);
    up(c, dK, K, nn * 8);
    up(c, dM, M, nn * 8);
    {
        Blob blob;
        Plan plan(&blob, &c.kstats);
        addLdltFactor(plan, dM, n, Wp, derr);                   // M = L D L^T in place
        Copy2dBatch c0;
        c0.add(cdesc(nullptr, 0, Li, n, n, n, 3, nullptr, nullptr, 0.0));
        c0.add(cdesc(dK, n, dK, n, n, n, 2));                   // K from its lower triangle
        plan.copy(c0);
        std::vector<VecDesc> v0(2);
        std::memset(v0.data(), 0, 2 * sizeof(VecDesc));
        v0[0].dst = Li;
        v0[0].ld = n;
        v0[0].n = static_cast<int32_t>(n);
        v0[0].op = 2;
        v0[1].src = dM;
        v0[1].dst = dq;
        v0[1].ld = n;
        v0[1].n = static_cast<int32_t>(n);
        v0[1].op = 0;                                           // D^-1/2
        plan.vec(v0);
        TrsmBatch ti;                                           // Li = L^-1
        ti.add(dM, n, Li, n, n, n, 0);
        ti.descs.back().rhsLower = 1;
        plan.trsm(ti);
        GemmBatch g1, g2;
        GemmDesc a = gdesc();                                   // T = L^-1 K
        a.A = Li;
        a.lda = n;
        a.triA = 1;
        a.B = dK;
        a.ldb = n;
        a.k = static_cast<int32_t>(n);
        a.C = T;
        a.ldc = n;
        a.m = a.n = static_cast<int32_t>(n);
        g1.add(a);
        GemmDesc b = gdesc();                                   // S = L^-1 T^T (lower)
        b.A = Li;
        b.lda = n;
        b.triA = 1;
        b.B = T;
        b.ldb = n;
        b.tb = 1;
        b.k = static_cast<int32_t>(n);
        b.C = S;
        b.ldc = n;
        b.m = b.n = static_cast<int32_t>(n);
        b.lower = 1;
        g2.add(b);
        plan.gemm(g1, false);
        plan.gemm(g2, false);
        Copy2dBatch c1;                                         // S = -D^-1/2 S D^-1/2, full
        c1.add(cdesc(S, n, S, n, n, n, 2, dq, dq, -1.0));
        plan.copy(c1);
        EigDesc ed;
        std::memset(&ed, 0, sizeof(ed));
        ed.A = S;
        ed.lda = n;
        ed.n = static_cast<int32_t>(n);
        ed.d = d;
        ed.e = e;
        ed.tau = tau;
        ed.kcap = static_cast<int32_t>(n);
        addSytrd(plan, {ed}, {0}, sv, sp);
        plan.run(c.stream);
        int err = 0;
        AMLS_HIP_CHECK(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        if (err != INT_MAX)
            throw StatusError(AMLS_E_NOT_SPD, "dense_eig: M is not positive definite");
    }
    // which eigenvalues of S (descending nu): j0 .. j1-1
    int64_t j0 = 0, j1 = n;
    if (range == 1) {
        EigDesc ed;
        std::memset(&ed, 0, sizeof(ed));
        ed.d = d;
        ed.e = e;
        ed.n = static_cast<int32_t>(n);
        ed.kcap = static_cast<int32_t>(n);
        std::vector<EigDesc> eds{ed, ed};
        // #{lambda <= x} = #{nu >= -x}
        std::vector<double> thr{-lo, -hi};
        Blob cb;
        const size_t od = cb.add(eds), ot = cb.add(thr);
        cb.upload(c.stream);
        launch_eig_count(cb.at<EigDesc>(od), 2, cb.at<double>(ot), cnt, static_cast<int>(n), c.stream);
        int32_t h[2] = {0, 0};
        AMLS_HIP_CHECK(hipMemcpyAsync(h, cnt, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        j0 = h[0];
        j1 = h[1];
    } else if (range == 2) {
        j0 = il - 1;
        j1 = std::min<int64_t>(iu, n);
    }
    if (j1 <= j0) return;
    {
        Blob blob;
        Plan plan(&blob, &c.kstats);
        EigDesc ed;
        std::memset(&ed, 0, sizeof(ed));
        ed.A = S;
        ed.lda = n;
        ed.n = static_cast<int32_t>(n);
        ed.d = d;
        ed.e = e;
        ed.tau = tau;
        ed.lam = lam;
        ed.Y = Y;
        ed.work = wk;
        ed.k = ed.kcap = static_cast<int32_t>(j1);
        addSelect(plan, {ed}, V != nullptr);
        if (V) {
            // Cholesky QR of the returned slice Yo: inverse iteration
            // orthogonalises only inside clusters within 1e-7 |T|, so vectors
            // of eigenvalues a gap g apart overlap by about eps |T| / g.
            // G = Yo^T Yo = L_G D_G L_G^T, Yo <- Yo L_G^-T D_G^-1/2 restores
            // orthonormality to rounding and moves each residual by about
            // eps |T| (the mixing coefficients are those overlaps).  Scratch:
            // G in T, L_G^-1 in the (consumed) K input, Q in S, D_G^-1/2 in sv.
            const int32_t kk = static_cast<int32_t>(j1 - j0);
            double* Yo = Y + j0 * n;
            GemmBatch gg, gq;
            GemmDesc g = gdesc();
            g.A = Yo;
            g.lda = n;
            g.ta = 1;
            g.B = Yo;
            g.ldb = n;
            g.k = static_cast<int32_t>(n);
            g.C = T;
            g.ldc = kk;
            g.m = g.n = kk;
            g.lower = 1;
            gg.add(g);
            plan.gemm(gg, false);
            addLdltFactor(plan, T, kk, Wp, derr);
            Copy2dBatch cz;
            cz.add(cdesc(nullptr, 0, dK, kk, kk, kk, 3, nullptr, nullptr, 0.0));
            plan.copy(cz);
            std::vector<VecDesc> vq(2);
            std::memset(vq.data(), 0, 2 * sizeof(VecDesc));
            vq[0].dst = dK;
            vq[0].ld = kk;
            vq[0].n = kk;
            vq[0].op = 2;                                       // unit diagonal
            vq[1].src = T;
            vq[1].dst = sv;
            vq[1].ld = kk;
            vq[1].n = kk;
            vq[1].op = 0;                                       // D_G^-1/2
            plan.vec(vq);
            TrsmBatch tq;
            tq.add(T, kk, dK, kk, kk, kk, 0);
            tq.descs.back().rhsLower = 1;
            plan.trsm(tq);
            GemmDesc q = gdesc();                               // Q = Yo L_G^-T
            q.A = Yo;
            q.lda = n;
            q.B = dK;
            q.ldb = kk;
            q.tb = 1;
            q.k = kk;
            q.C = S;
            q.ldc = n;
            q.m = static_cast<int32_t>(n);
            q.n = kk;
            gq.add(q);
            plan.gemm(gq, false);
            Copy2dBatch cq;
            cq.add(cdesc(S, n, Yo, n, n, kk, 0, nullptr, sv));  // Yo = Q D_G^-1/2
            plan.copy(cq);
            Copy2dBatch c2;                                     // y <- D^-1/2 y
            c2.add(cdesc(Yo, n, Yo, n, n, kk, 0, dq));
            plan.copy(c2);
            GemmBatch g3;                                       // X = L^-T y
            GemmDesc x = gdesc();
            x.A = Li;
            x.lda = n;
            x.ta = 1;
            x.triA = 2;
            x.B = Yo;
            x.ldb = n;
            x.k = static_cast<int32_t>(n);
            x.C = X;
            x.ldc = n;
            x.m = static_cast<int32_t>(n);
            x.n = kk;
            g3.add(x);
            plan.gemm(g3, false);
        }
        plan.run(c.stream);
    }
    std::vector<double> nu(j1);
    AMLS_HIP_CHECK(hipMemcpyAsync(nu.data(), lam, j1 * 8, hipMemcpyDeviceToHost, c.stream));
    if (V)
        AMLS_HIP_CHECK(hipMemcpyAsync(V, X, (j1 - j0) * n * 8, hipMemcpyDeviceToHost, c.stream));
    int err = INT_MAX;
    if (V) AMLS_HIP_CHECK(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (err != INT_MAX)
        throw StatusError(AMLS_E_NOCONV, "dense_eig: eigenvectors numerically dependent");
    for (int64_t j = j0; j < j1; ++j) w[j - j0] = -nu[j];
    *mOut = j1 - j0;
}

}  // namespace amls
