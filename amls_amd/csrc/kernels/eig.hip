// Batched dense symmetric eigensolver for selected eigenpairs:
//   Householder tridiagonalisation (one launch per column step, all matrices
//   of a batch advancing together), Sturm-count bisection for eigenvalues
//   selected by index or by value, inverse iteration with modified
//   Gram-Schmidt inside clusters of close eigenvalues, and back-application of
//   the Householder reflectors.
//
// This is the GPU replacement of the LAPACK dsygvx / dsygvd calls behind
// LapackDenseEigensolver (src/linalg/dense_eigensolver.cpp:11-133): the same
// reduction-to-tridiagonal / bisection / inverse-iteration family
// (RANGE='I' by index, RANGE='V' by value, abstol = 0), written from the
// algorithms' textbook definitions.
#include "common.hpp"
#include "launch.hpp"

#include <cfloat>

namespace amls {
namespace dev {

constexpr int RS = 64;   // rows per workgroup in a tridiagonalisation step

__device__ __forceinline__ double block_sum(double v, double* red) {
    // 256-thread block reduction; red has >= 4 slots
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int t = threadIdx.x;
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// Step i of the lower Householder reduction, one 64 x 64 tile of the
// trailing lower triangle [i+1, n)^2 per workgroup (absolute tile grid).
// Only the lower triangle is read or written.  Fused per tile: apply the
// rank-2 update of step i-1 and accumulate p_i = A v_i (row sums and, by
// symmetry, column sums) with FP64 atomics.  Per-matrix scratch:
//   sv (2n): v_{i-1} / v_i by step parity,
//   sp (3n): p_{i-1} is read, p_i accumulated, p_{i+1} cleared (step mod 3).
// Every workgroup recomputes the scalars it needs (p.v, the Householder
// vector of column i) from the shared vectors, so one launch per step
// suffices.  Work item (x = matrix, y = tile row, z = tile col, w = 1 for the
// matrix's leader, which writes d, e, tau, v_i and the previous reflector).
constexpr int SY_KR = 8;   // strided rows per thread held in registers (n - i <= 2048)

// Step i for tile (wk.y, wk.z) of matrix d (wk.y < 0: the leader's last
// step i = n - 1 only); sv_m / sp_m: the matrix's 2n / 3n scratch.
__device__ __forceinline__ void sytrd_step_body(const EigDesc& d, int i, int4 wk,
                                                double* __restrict__ sv_m,
                                                double* __restrict__ sp_m) {
    __shared__ double tile[64 * 65];
    __shared__ double vr[64], wr[64], xr[64], vc[64], wc[64], xc[64];
    __shared__ double red[8];
    const int n = d.n;
    const int64_t lda = d.lda;
    double* A = d.A;
    const int t = threadIdx.x;
    const double* svr = sv_m + (int64_t)((i + 1) & 1) * n;   // v_{i-1}
    double* svw = sv_m + (int64_t)(i & 1) * n;               // v_i
    const double* spr = sp_m + (int64_t)((i + 2) % 3) * n;   // p_{i-1}
    double* spw = sp_m + (int64_t)(i % 3) * n;               // p_i
    double* spz = sp_m + (int64_t)((i + 1) % 3) * n;         // p_{i+1}
    const bool hasPrev = i > 0;
    const double* Ai = A + (int64_t)i * lda;   // column i (of A after update i-2)

    // ---- every load that does not depend on this step's scalars is issued
    // first, so the step costs one round trip to memory before its two
    // block reductions: the strided rows r = i + t + 256 j of p_{i-1},
    // v_{i-1} and column i, this workgroup's tile, and the tile's row / column
    // entries of the same three vectors.
    const double tauP = hasPrev ? d.tau[i - 1] : 0.0;
    const double vI = hasPrev ? svr[i] : 0.0, pI = hasPrev ? spr[i] : 0.0, aII = Ai[i];
    const bool has1 = i + 1 < n;
    const double v1 = (hasPrev && has1) ? svr[i + 1] : 0.0, p1 = (hasPrev && has1) ? spr[i + 1] : 0.0;
    const double a1 = has1 ? Ai[i + 1] : 0.0;
    double pr[SY_KR], vv[SY_KR], ac[SY_KR];
#pragma unroll
    for (int j = 0; j < SY_KR; ++j) {
        const int r = i + t + 256 * j;
        const bool in = r < n;
        pr[j] = (hasPrev && in) ? spr[r] : 0.0;
        vv[j] = (hasPrev && in) ? svr[r] : 0.0;
        ac[j] = in ? Ai[r] : 0.0;
    }
    const int r0 = wk.y * 64, c0 = wk.z * 64;
    const bool haveTile = wk.y >= 0 && n - i - 1 > 0;
    double ta[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = t + 256 * k, rr = e & 63, cc = e >> 6;
        const int r = r0 + rr, c = c0 + cc;
        ta[k] = (haveTile && r < n && c > i && r >= c) ? A[r + (int64_t)c * lda] : 0.0;
    }
    // tile vectors: thread t < 64 takes row r0 + t, 64 <= t < 128 column c0 + t - 64
    const int vidx = t < 64 ? r0 + t : c0 + t - 64;
    const bool vok = haveTile && t < 128 && vidx > i && vidx < n;
    const double vv_t = (vok && hasPrev) ? svr[vidx] : 0.0;
    const double pv_t = (vok && hasPrev) ? spr[vidx] : 0.0;
    const double av_t = vok ? Ai[vidx] : 0.0;

    // ---- w_{i-1} = tau p + alpha v with alpha = -1/2 tau^2 (p.v)
    double alphaP = 0.0;
    if (hasPrev) {
        double part = 0.0;
#pragma unroll
        for (int j = 0; j < SY_KR; ++j) part += pr[j] * vv[j];
        for (int r = i + t + 256 * SY_KR; r < n; r += 256) part += spr[r] * svr[r];
        alphaP = -0.5 * tauP * tauP * block_sum(part, red);
    }
    const double wI = tauP * pI + alphaP * vI;
    // column i below the diagonal after update i-1: x_r = a_r - (v_r w_i + w_r v_i)
    auto xof = [&](double a, double v, double p) { return a - (v * wI + (tauP * p + alphaP * v) * vI); };
    const int m = n - i - 1;
    if (m <= 0) {
        if (wk.w && t == 0) d.d[i] = aII - 2.0 * vI * wI;
        return;
    }
    double part = 0.0;
#pragma unroll
    for (int j = 0; j < SY_KR; ++j) {
        const int r = i + t + 256 * j;
        if (r >= i + 2 && r < n) {
            const double x = xof(ac[j], vv[j], pr[j]);
            part += x * x;
        }
    }
    for (int r = i + t + 256 * SY_KR; r < n; r += 256) {
        const double x = xof(Ai[r], hasPrev ? svr[r] : 0.0, hasPrev ? spr[r] : 0.0);
        part += x * x;
    }
    const double xnorm2 = block_sum(part, red);
    const double alpha = xof(a1, v1, p1);
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (xnorm2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    if (wk.w) {
        // leader: d, e, tau, v_i, clear p_{i+1}, store the reflector of step i-1
        if (t == 0) {
            d.d[i] = aII - 2.0 * vI * wI;
            d.e[i] = beta;
            d.tau[i] = tau;
            svw[i] = 0.0;
        }
#pragma unroll
        for (int j = 0; j < SY_KR; ++j) {
            const int r = i + t + 256 * j;
            if (r >= i + 1 && r < n) {
                svw[r] = r == i + 1 ? 1.0 : xof(ac[j], vv[j], pr[j]) * scal;
                spz[r] = 0.0;
                if (hasPrev) A[r + (int64_t)(i - 1) * lda] = vv[j];
            }
        }
        for (int r = i + t + 256 * SY_KR; r < n; r += 256) {
            const double v = hasPrev ? svr[r] : 0.0;
            svw[r] = xof(Ai[r], v, hasPrev ? spr[r] : 0.0) * scal;
            spz[r] = 0.0;
            if (hasPrev) A[r + (int64_t)(i - 1) * lda] = v;
        }
    }
    if (wk.y < 0) return;
    if (t < 128) {
        const double w = tauP * pv_t + alphaP * vv_t;
        const double x = vok ? (vidx == i + 1 ? 1.0 : xof(av_t, vv_t, pv_t) * scal) : 0.0;
        if (t < 64) {
            vr[t] = vv_t;
            wr[t] = vok ? w : 0.0;
            xr[t] = x;
        } else {
            vc[t - 64] = vv_t;
            wc[t - 64] = vok ? w : 0.0;
            xc[t - 64] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int e = t + 256 * k, rr = e & 63, cc = e >> 6;
        const int r = r0 + rr, c = c0 + cc;
        double a = ta[k];
        if (r < n && c > i && r >= c && hasPrev) {
            a -= vr[rr] * wc[cc] + wr[rr] * vc[cc];
            A[r + (int64_t)c * lda] = a;
        }
        tile[rr * 65 + cc] = a;
    }
    __syncthreads();
    if (t < 64) {
        const int r = r0 + t;
        double s = 0.0;
        for (int cc = 0; cc < 64; ++cc) s += tile[t * 65 + cc] * xc[cc];
        if (r > i && r < n) atomicAdd(spw + r, s);
    } else if (t < 128) {
        const int cc = t - 64;
        const int c = c0 + cc;
        double s = 0.0;
        for (int rr = 0; rr < 64; ++rr)
            if (r0 + rr != c) s += tile[rr * 65 + cc] * xr[rr];
        if (c > i && c < n) atomicAdd(spw + c, s);
    }
}

// Batched step: work item (matrix, tile row, tile col, leader flag).
__global__ void __launch_bounds__(256) sytrd_tile_kernel(const EigDesc* __restrict__ descs,
                                                         const int4* __restrict__ work, int i,
                                                         double* __restrict__ sv_all,
                                                         double* __restrict__ sp_all,
                                                         const int64_t* __restrict__ soff) {
    const int4 wk = work[blockIdx.x];
    const EigDesc d = descs[wk.x];
    sytrd_step_body(d, i, wk, sv_all + 2 * soff[wk.x], sp_all + 3 * soff[wk.x]);
}

// One matrix: the descriptor comes by value and the tile from the block
// index (tiles R >= C >= t0 in row order, t0 = (i + 1) / 64), so a step
// starts its data loads without first fetching work items and descriptors.
// last != 0: the single-workgroup final step i = n - 1.
__global__ void __launch_bounds__(256) sytrd_tile1_kernel(EigDesc d, int i, int t0, int last,
                                                          double* __restrict__ sv,
                                                          double* __restrict__ sp) {
    int4 wk;
    wk.x = 0;
    if (last) {
        wk.y = wk.z = -1;
        wk.w = 1;
    } else {
        int idx = blockIdx.x, R = t0;
        while (idx > R - t0) {
            idx -= R - t0 + 1;
            ++R;
        }
        wk.y = R;
        wk.z = t0 + idx;
        wk.w = (R == t0 && wk.z == t0) ? 1 : 0;
    }
    sytrd_step_body(d, i, wk, sv, sp);
}

// Sturm count: number of eigenvalues of T strictly below x (pivmin guarded).
__device__ int sturm_count(const double* __restrict__ dd, const double* __restrict__ ee, int n,
                           double x, double pivmin) {
    int cnt = 0;
    double q = dd[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
    for (int j = 1; j < n; ++j) {
        const double e = ee[j - 1];
        q = dd[j] - x - e * e / q;
        if (fabs(q) < pivmin) q = -pivmin;
        if (q < 0.0) ++cnt;
    }
    return cnt;
}

__device__ void gersh(const double* dd, const double* ee, int n, double& lo, double& hi,
                      double& pivmin, double& tnorm) {
    lo = 1e308;
    hi = -1e308;
    double emax = 0.0;
    for (int j = 0; j < n; ++j) {
        const double el = j > 0 ? fabs(ee[j - 1]) : 0.0;
        const double er = j + 1 < n ? fabs(ee[j]) : 0.0;
        lo = fmin(lo, dd[j] - el - er);
        hi = fmax(hi, dd[j] + el + er);
        if (j + 1 < n) emax = fmax(emax, ee[j] * ee[j]);
    }
    tnorm = fmax(fabs(lo), fabs(hi));
    const double pad = 2.0 * DBL_EPSILON * tnorm + 1e-300;
    lo -= pad;
    hi += pad;
    pivmin = DBL_MIN * fmax(1.0, emax);
}

// counts[mat] = number of eigenvalues >= thr[mat] (capped at cap[mat])
// Gershgorin bounds and pivmin of a tridiagonal (dd: n, ee: n - 1 entries,
// any memory space), lanes of one wave striding the rows.
__device__ void gersh_wave(const double* dd, const double* ee, int n, int lane, double& lo,
                           double& hi, double& pivmin, double& tn) {
    lo = 1e308;
    hi = -1e308;
    double emax = 0.0;
    for (int r = lane; r < n; r += 64) {
        const double el = r > 0 ? fabs(ee[r - 1]) : 0.0;
        const double er = r + 1 < n ? fabs(ee[r]) : 0.0;
        lo = fmin(lo, dd[r] - el - er);
        hi = fmax(hi, dd[r] + el + er);
        if (r + 1 < n) emax = fmax(emax, ee[r] * ee[r]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
        emax = fmax(emax, __shfl_xor(emax, o, 64));
    }
    tn = fmax(fabs(lo), fabs(hi));
    lo -= 2.0 * DBL_EPSILON * tn + 1e-300;
    hi += 2.0 * DBL_EPSILON * tn + 1e-300;
    pivmin = DBL_MIN * fmax(1.0, emax);
}

// Copies matrix d's diagonal and off-diagonal into LDS (sde: 2 nmax doubles)
// when useLds, and returns where to read them from.
__device__ void stage_tridiag(const EigDesc& d, double* sde, int useLds, const double*& dd,
                              const double*& ee) {
    dd = d.d;
    ee = d.e;
    if (useLds) {
        for (int r = threadIdx.x; r < d.n; r += blockDim.x) {
            sde[r] = d.d[r];
            sde[d.n + r] = r + 1 < d.n ? d.e[r] : 0.0;
        }
        __syncthreads();
        dd = sde;
        ee = sde + d.n;
    }
}

// counts[mat] = number of eigenvalues >= thr[mat] (capped at kcap); one
// 64-lane workgroup per matrix, its tridiagonal staged in LDS
__global__ void __launch_bounds__(64) eig_count_kernel(const EigDesc* __restrict__ descs,
                                                       const double* __restrict__ thr,
                                                       int32_t* __restrict__ counts, int useLds) {
    extern __shared__ double sde[];
    const int m = blockIdx.x;
    const EigDesc d = descs[m];
    if (d.n <= 0) {
        if (threadIdx.x == 0) counts[m] = 0;
        return;
    }
    const double *dd, *ee;
    stage_tridiag(d, sde, useLds, dd, ee);
    double lo, hi, pivmin, tn;
    gersh_wave(dd, ee, d.n, threadIdx.x, lo, hi, pivmin, tn);
    if (threadIdx.x == 0) {
        const int below = sturm_count(dd, ee, d.n, thr[m], pivmin);
        counts[m] = min(d.n - below, d.kcap);
    }
}

// lam[j] = (j+1)-th largest eigenvalue, j < k; one wave per work item
// (x = matrix, y = j; the host pads each matrix's items to whole workgroups
// of four with y = -1, so a workgroup stages one tridiagonal in LDS).
// Multisection: each round the 64 lanes evaluate the Sturm count at 64
// interior points of the current bracket, which shrinks the bracket 65-fold
// per round (instead of 2-fold for plain bisection).
__global__ void __launch_bounds__(256) bisect_kernel(const EigDesc* __restrict__ descs,
                                                     const int2* __restrict__ work, int nwork,
                                                     int useLds) {
    extern __shared__ double sde[];
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const EigDesc d = descs[work[blockIdx.x * 4].x];
    const double *dd, *ee;
    stage_tridiag(d, sde, useLds, dd, ee);
    if (wid >= nwork) return;
    const int2 wk = work[wid];
    if (wk.y < 0) return;
    const int n = d.n;
    const int j = wk.y;
    double lo, hi, pivmin, tn;
    gersh_wave(dd, ee, n, lane, lo, hi, pivmin, tn);
    const int a = n - 1 - j;    // ascending index
    for (int round = 0; round < 16; ++round) {
        const double tol = 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin;
        if (hi - lo <= tol) break;
        const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
        const int c = sturm_count(dd, ee, n, x, pivmin);
        // counts grow with x: lanes with c <= a form a prefix
        const unsigned long long mask = __ballot(c <= a);
        const int np = __popcll(mask);
        const double nlo = np > 0 ? lo + (hi - lo) * (double)np / 65.0 : lo;
        const double nhi = np < 64 ? lo + (hi - lo) * (double)(np + 1) / 65.0 : hi;
        if (!(nhi < hi || nlo > lo)) break;   // no progress at this resolution
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) d.lam[j] = 0.5 * (lo + hi);
}

// Inverse iteration.  Thread per eigenvalue; a thread whose eigenvalue starts
// a cluster (gap to the previous one above ctol*|T|) computes the whole
// cluster in order, orthogonalising each vector against the earlier ones.
// Scratch of matrix m is interleaved by eigenvalue, element (r, j) of array
// q at work[(q n + r) k + j], so a wave's accesses are contiguous.
__global__ void __launch_bounds__(64) stein_kernel(const EigDesc* __restrict__ descs,
                                                   const int2* __restrict__ work, int nwork,
                                                   const int64_t* __restrict__ woff,
                                                   int useLds) {
    // every 64-item block belongs to one matrix (the host pads per matrix);
    // its tridiagonal is staged in LDS when it fits (useLds)
    extern __shared__ double sde[];
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int2 wk0 = work[blockIdx.x * blockDim.x];
    const EigDesc d = descs[wk0.x];
    const int n = d.n;
    const double* dd = d.d;
    const double* ee = d.e;
    if (useLds) {
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            sde[r] = d.d[r];
            sde[n + r] = r + 1 < n ? d.e[r] : 0.0;
        }
        __syncthreads();
        dd = sde;
        ee = sde + n;
    }
    if (id >= nwork) return;
    const int2 wk = work[id];
    if (wk.y < 0) return;
    const int K = d.k;
    const int j0 = wk.y;
    double lo, hi, pivmin, tn;
    gersh(dd, ee, n, lo, hi, pivmin, tn);
    const double ctol = 1e-7 * tn;
    if (j0 > 0 && fabs(d.lam[j0 - 1] - d.lam[j0]) <= ctol) return;   // not a leader
    int j1 = j0 + 1;
    while (j1 < K && fabs(d.lam[j1 - 1] - d.lam[j1]) <= ctol) ++j1;
    (void)woff;
    double* base = d.work + j0;
    const int64_t sK = K;
    double* ud = base;
    double* u1 = base + (int64_t)n * sK;
    double* u2 = base + 2 * (int64_t)n * sK;
    double* ml = base + 3 * (int64_t)n * sK;
    double* pv = base + 4 * (int64_t)n * sK;
    double* x = base + 5 * (int64_t)n * sK;
#define S_(arr, r) arr[(int64_t)(r) * sK]
    const double tiny = DBL_EPSILON * fmax(tn, DBL_MIN);
    for (int j = j0; j < j1; ++j) {
        const double lam = d.lam[j] + (j - j0) * 10.0 * DBL_EPSILON * tn;
        // LU with partial pivoting of T - lam I, kept in registers along the
        // recurrence and stored for the solves
        double udk = dd[0] - lam;
        double u1k = n > 1 ? ee[0] : 0.0;
        double u2k = 0.0;
        for (int k = 0; k + 1 < n; ++k) {
            const double sub = ee[k];
            const double dg = dd[k + 1] - lam;
            const double sp = k + 2 < n ? ee[k + 1] : 0.0;
            double nud, nu1;
            if (fabs(udk) >= fabs(sub)) {
                double p = udk;
                if (fabs(p) < tiny) p = copysign(tiny, p == 0.0 ? 1.0 : p);
                const double l = sub / p;
                S_(ud, k) = p;
                S_(u1, k) = u1k;
                S_(u2, k) = u2k;
                S_(ml, k) = l;
                S_(pv, k) = 0.0;
                nud = dg - l * u1k;
                nu1 = sp - l * u2k;
            } else {
                const double l = udk / sub;
                S_(ud, k) = sub;
                S_(u1, k) = dg;
                S_(u2, k) = sp;
                S_(ml, k) = l;
                S_(pv, k) = 1.0;
                nud = u1k - l * dg;
                nu1 = u2k - l * sp;
            }
            udk = nud;
            u1k = nu1;
            u2k = 0.0;
        }
        if (fabs(udk) < tiny) udk = copysign(tiny, udk == 0.0 ? 1.0 : udk);
        S_(ud, n - 1) = udk;
        S_(u1, n - 1) = 0.0;
        S_(u2, n - 1) = 0.0;
        // inverse iteration: two solves for an isolated eigenvalue, three inside
        // a cluster; the start vector is generated inside the first forward
        // pass, and the normalisation of each solve is applied when the next
        // pass reads the vector
        uint32_t s = 2463534242u ^ (uint32_t)(j * 2654435761u) ^ (uint32_t)(wk.x * 40503u);
        auto start = [&]() {
            s ^= s << 13; s ^= s >> 17; s ^= s << 5;
            return 1.0 + (double)(s & 0xffff) * (1.0 / 65536.0);
        };
        const int iters = j > j0 ? 3 : 2;
        double sc = 1.0;
        for (int it = 0; it < iters; ++it) {
            // forward: row interchanges and multipliers
            double xk = it == 0 ? start() : S_(x, 0) * sc;
            for (int k = 0; k + 1 < n; ++k) {
                double xn = it == 0 ? start() : S_(x, k + 1) * sc;
                if (S_(pv, k) != 0.0) {
                    const double tmp = xk;
                    xk = xn;
                    xn = tmp;
                }
                S_(x, k) = xk;
                xk = xn - S_(ml, k) * xk;
            }
            // back substitution with U (bandwidth 3), accumulating the norm
            double nrm = 0.0, amax = 0.0, sgn = 1.0;
            auto acc = [&](double v) {
                nrm += v * v;
                if (fabs(v) > amax) { amax = fabs(v); sgn = v < 0.0 ? -1.0 : 1.0; }
            };
            double x1 = xk / S_(ud, n - 1);
            S_(x, n - 1) = x1;
            acc(x1);
            double x2 = 0.0;
            if (n > 1) {
                const double v = (S_(x, n - 2) - S_(u1, n - 2) * x1) / S_(ud, n - 2);
                S_(x, n - 2) = v;
                acc(v);
                x2 = x1;
                x1 = v;
            }
            for (int k = n - 3; k >= 0; --k) {
                const double v = (S_(x, k) - S_(u1, k) * x1 - S_(u2, k) * x2) / S_(ud, k);
                S_(x, k) = v;
                acc(v);
                x2 = x1;
                x1 = v;
            }
            // orthogonalise against earlier members of the cluster (modified
            // Gram-Schmidt; the last subtraction pass recomputes the norm)
            for (int q = j0; q < j; ++q) {
                const double* y = d.Y + (int64_t)q * n;
                double dot = 0.0;
                for (int r = 0; r < n; ++r) dot += S_(x, r) * y[r];
                const bool lastq = q + 1 == j;
                if (lastq) { nrm = 0.0; amax = 0.0; sgn = 1.0; }
                for (int r = 0; r < n; ++r) {
                    const double v = S_(x, r) - dot * y[r];
                    S_(x, r) = v;
                    if (lastq) acc(v);
                }
            }
            sc = sgn / sqrt(nrm);
        }
        double* y = d.Y + (int64_t)j * n;
        for (int r = 0; r < n; ++r) y[r] = S_(x, r) * sc;
    }
#undef S_
}

// Y := Q Y for the reflectors stored by the tridiagonalisation.  One
// 256-thread workgroup per (matrix, block of nc <= 8 columns); the block of Y
// lives in LDS for the whole sweep over the reflectors.  Thread t owns the
// rows r = t + 256 m of Y for every reflector, so the only cross-thread step
// per reflector is the reduction of the dot products (one barrier, partial
// sums double-buffered between consecutive reflectors), and each reflector's vector is
// fetched into registers three reflectors ahead of its use.
// Work item (x = matrix, y = first column, z = column count).
constexpr int OR_KR = 8;   // rows per thread in registers (n <= 2048)

__global__ void __launch_bounds__(256) ormtr_kernel(const EigDesc* __restrict__ descs,
                                                    const int4* __restrict__ work) {
    extern __shared__ double sY[];       // n x nc, column-major (ld n)
    __shared__ double part[2][4][8];
    const int4 wk = work[blockIdx.x];
    const EigDesc d = descs[wk.x];
    const int n = d.n;
    const int c0 = wk.y, nc = wk.z;
    double* Y = d.Y + (int64_t)c0 * n;
    const double* A = d.A;
    const int64_t lda = d.lda;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    for (int e = t; e < n * nc; e += 256) sY[e] = Y[e];
    // v_i(r): 0 for r <= i, 1 at r = i + 1, stored below
    auto load = [&](double (&v)[OR_KR], double& tau, int i) {
        tau = i >= 0 ? d.tau[i] : 0.0;
#pragma unroll
        for (int m = 0; m < OR_KR; ++m) {
            const int r = t + 256 * m;
            v[m] = (i < 0 || r <= i || r >= n) ? 0.0 : (r == i + 1 ? 1.0 : A[r + (int64_t)i * lda]);
        }
    };
    int par = 0;   // partial-sum buffer, alternating over the applied reflectors
    auto step = [&](const double (&v)[OR_KR], double tau, int i) {
        if (i < 0 || tau == 0.0) return;   // uniform over the workgroup
        double s[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) s[c] = 0.0;
#pragma unroll
        for (int m = 0; m < OR_KR; ++m) {
            const int r = t + 256 * m;
            if (r > i && r < n) {
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (c < nc) s[c] += v[m] * sY[r + c * n];
            }
        }
        for (int r = t + 256 * OR_KR; r < n; r += 256) {
            const double vr = r <= i ? 0.0 : (r == i + 1 ? 1.0 : A[r + (int64_t)i * lda]);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c < nc) s[c] += vr * sY[r + c * n];
        }
        double (*pb)[8] = part[par];
        par ^= 1;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (c < nc) {
                double x = s[c];
                for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
                if (lane == 0) pb[w][c] = x;
            }
        }
        __syncthreads();
        double coef[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            coef[c] = c < nc ? tau * (pb[0][c] + pb[1][c] + pb[2][c] + pb[3][c]) : 0.0;
#pragma unroll
        for (int m = 0; m < OR_KR; ++m) {
            const int r = t + 256 * m;
            if (r > i && r < n) {
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (c < nc) sY[r + c * n] -= coef[c] * v[m];
            }
        }
        for (int r = t + 256 * OR_KR; r < n; r += 256) {
            const double vr = r <= i ? 0.0 : (r == i + 1 ? 1.0 : A[r + (int64_t)i * lda]);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c < nc) sY[r + c * n] -= coef[c] * vr;
        }
    };
    double va[OR_KR], vb[OR_KR], vc[OR_KR], vd[OR_KR];
    double ta, tb, tc, td;
    load(va, ta, n - 2);
    load(vb, tb, n - 3);
    load(vc, tc, n - 4);
    __syncthreads();
    for (int i = n - 2; i >= 0; i -= 4) {
        load(vd, td, i - 3);
        step(va, ta, i);
        load(va, ta, i - 4);
        step(vb, tb, i - 1);
        load(vb, tb, i - 5);
        step(vc, tc, i - 2);
        load(vc, tc, i - 6);
        step(vd, td, i - 3);
    }
    __syncthreads();
    for (int e = t; e < n * nc; e += 256) Y[e] = sY[e];
}

}  // namespace dev

void launch_sytrd_step1(const dev::EigDesc& d, int i, double* sv, double* sp, hipStream_t s) {
    const int n = d.n;
    if (i >= n) return;
    if (i == n - 1) {
        hipLaunchKernelGGL(dev::sytrd_tile1_kernel, dim3(1), dim3(256), 0, s, d, i, 0, 1, sv, sp);
        return;
    }
    const int t0 = (i + 1) / 64, t1 = (n - 1) / 64;
    const int nt = (t1 - t0 + 1) * (t1 - t0 + 2) / 2;
    hipLaunchKernelGGL(dev::sytrd_tile1_kernel, dim3(nt), dim3(256), 0, s, d, i, t0, 0, sv, sp);
}

void launch_sytrd_step(const dev::EigDesc* d, const int4* work, int nwork, int i, double* sv,
                       double* sp, const int64_t* soff, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::sytrd_tile_kernel, dim3(nwork), dim3(256), 0, s, d, work, i, sv, sp,
                       soff);
}

void launch_eig_count(const dev::EigDesc* d, int nmat, const double* thr, int32_t* counts,
                      int nmax, hipStream_t s) {
    if (nmat <= 0) return;
    const size_t lds = (size_t)2 * nmax * sizeof(double);
    const int useLds = lds <= 60000 ? 1 : 0;
    hipLaunchKernelGGL(dev::eig_count_kernel, dim3(nmat), dim3(64), useLds ? lds : 0, s, d, thr,
                       counts, useLds);
}

void launch_bisect(const dev::EigDesc* d, const int2* work, int nwork, int nmax, hipStream_t s) {
    if (nwork <= 0) return;
    const size_t lds = (size_t)2 * nmax * sizeof(double);
    const int useLds = lds <= 60000 ? 1 : 0;
    hipLaunchKernelGGL(dev::bisect_kernel, dim3((nwork + 3) / 4), dim3(256), useLds ? lds : 0, s, d,
                       work, nwork, useLds);
}

void launch_stein(const dev::EigDesc* d, const int2* work, int nwork, const int64_t* woff,
                  int nmax, hipStream_t s) {
    if (nwork <= 0) return;
    const size_t lds = (size_t)2 * nmax * sizeof(double);
    const int useLds = lds <= 60000 ? 1 : 0;
    hipLaunchKernelGGL(dev::stein_kernel, dim3((nwork + 63) / 64), dim3(64), useLds ? lds : 0, s, d,
                       work, nwork, woff, useLds);
}

void launch_ormtr(const dev::EigDesc* d, const int4* work, int nwork, int ldsDoubles, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::ormtr_kernel, dim3(nwork), dim3(256), (size_t)ldsDoubles * 8, s, d, work);
}

}  // namespace amls
