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

// Step i of the lower Householder reduction.  Matrices are stored full
// (both triangles kept current on the trailing block).  Per-matrix scratch
// (2n each, double-buffered by step parity): v of the previous step and
// p = A v.  Launch i reads the (i-1) buffers and writes the i buffers, so no
// workgroup overwrites a value another one of the same launch still reads.
// Work item (x = matrix, y = row slice of RS rows starting at i+1).
__global__ void __launch_bounds__(256) sytrd_step_kernel(const EigDesc* __restrict__ descs,
                                                         const int2* __restrict__ work, int i,
                                                         double* __restrict__ sv_all,
                                                         double* __restrict__ sp_all,
                                                         const int64_t* __restrict__ soff) {
    extern __shared__ double lds[];
    __shared__ double red[8];
    __shared__ double pacc[4 * RS];
    const int2 wk = work[blockIdx.x];
    const EigDesc d = descs[wk.x];
    const int n = d.n;
    const int64_t lda = d.lda;
    double* A = d.A;
    const double* svr = sv_all + soff[wk.x] + (int64_t)((i + 1) & 1) * n;   // step i-1
    const double* spr = sp_all + soff[wk.x] + (int64_t)((i + 1) & 1) * n;
    double* svw = sv_all + soff[wk.x] + (int64_t)(i & 1) * n;               // step i
    double* spw = sp_all + soff[wk.x] + (int64_t)(i & 1) * n;
    const int t = threadIdx.x;
    double* v = lds;            // indexed by global row/col
    double* w = lds + n;
    double* vprev = lds + 2 * n;
    // ---- w_{i-1} = tau p - 1/2 tau^2 (p.v) v, recomputed redundantly per block
    const bool hasPrev = i > 0;
    const double tauPrev = hasPrev ? d.tau[i - 1] : 0.0;
    if (hasPrev) {
        double part = 0.0;
        for (int r = i + t; r < n; r += 256) part += spr[r] * svr[r];
        const double pv = block_sum(part, red);
        const double alpha = -0.5 * tauPrev * tauPrev * pv;
        for (int r = i + t; r < n; r += 256) {
            vprev[r] = svr[r];
            w[r] = tauPrev * spr[r] + alpha * svr[r];
        }
    }
    __syncthreads();
    const int m = n - i - 1;
    if (m <= 0) {
        if (wk.y == 0 && t == 0) {
            double a = A[i + (int64_t)i * lda];
            if (hasPrev) a -= 2.0 * vprev[i] * w[i];
            d.d[i] = a;
        }
        return;
    }
    // ---- column i after the previous update; Householder of A(i+1:n, i)
    double part = 0.0;
    for (int r = i + 1 + t; r < n; r += 256) {
        double a = A[r + (int64_t)i * lda];
        if (hasPrev) a -= vprev[r] * w[i] + w[r] * vprev[i];
        v[r] = a;
        if (r > i + 1) part += a * a;
    }
    const double xnorm2 = block_sum(part, red);
    double aii = A[i + (int64_t)i * lda];
    if (hasPrev) aii -= 2.0 * vprev[i] * w[i];
    const double alpha = v[i + 1];
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (xnorm2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    __syncthreads();
    for (int r = i + 1 + t; r < n; r += 256) v[r] = (r == i + 1) ? 1.0 : v[r] * scal;
    __syncthreads();
    if (wk.y == 0 && t == 0) {
        d.d[i] = aii;
        d.e[i] = beta;
        d.tau[i] = tau;
    }
    // ---- fused: apply the previous update to my rows and form p = A v;
    // the previous reflector is stored below the subdiagonal of column i-1
    const int r0 = i + 1 + wk.y * RS;
    const int rl = t & 63, cg = t >> 6;
    const int r = r0 + rl;
    double acc = 0.0;
    if (r < n) {
        if (hasPrev && cg == 0) A[r + (int64_t)(i - 1) * lda] = vprev[r];
        const double vr = hasPrev ? vprev[r] : 0.0, wr = hasPrev ? w[r] : 0.0;
        for (int c = i + 1 + cg; c < n; c += 4) {
            double* pa = A + r + (int64_t)c * lda;
            double a = *pa;
            if (hasPrev) {
                a -= vr * w[c] + wr * vprev[c];
                *pa = a;
            }
            acc += a * v[c];
        }
    }
    pacc[cg * RS + rl] = acc;
    __syncthreads();
    if (t < RS && r0 + t < n) {
        const double pr = pacc[t] + pacc[RS + t] + pacc[2 * RS + t] + pacc[3 * RS + t];
        spw[r0 + t] = pr;
        svw[r0 + t] = v[r0 + t];
    }
    // slice 0 also records v(i) = 0 so the next step's window [i+1, n) is clean
    if (wk.y == 0 && t == 0) {
        svw[i] = 0.0;
        spw[i] = 0.0;
    }
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
__global__ void eig_count_kernel(const EigDesc* __restrict__ descs, int nmat,
                                 const double* __restrict__ thr, int32_t* __restrict__ counts) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmat) return;
    const EigDesc d = descs[m];
    if (d.n <= 0) { counts[m] = 0; return; }
    double lo, hi, pivmin, tn;
    gersh(d.d, d.e, d.n, lo, hi, pivmin, tn);
    const int below = sturm_count(d.d, d.e, d.n, thr[m], pivmin);
    int k = d.n - below;
    if (k > d.kcap) k = d.kcap;
    counts[m] = k;
}

// lam[j] = (j+1)-th largest eigenvalue, j < k; work item (x = matrix, y = j)
__global__ void bisect_kernel(const EigDesc* __restrict__ descs, const int2* __restrict__ work,
                              int nwork) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nwork) return;
    const int2 wk = work[id];
    const EigDesc d = descs[wk.x];
    const int n = d.n;
    const int j = wk.y;
    double lo, hi, pivmin, tn;
    gersh(d.d, d.e, n, lo, hi, pivmin, tn);
    const int a = n - 1 - j;    // ascending index
    for (int it = 0; it < 256; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double tol = 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin;
        if (hi - lo <= tol || mid <= lo || mid >= hi) break;
        if (sturm_count(d.d, d.e, n, mid, pivmin) > a) hi = mid;
        else lo = mid;
    }
    d.lam[j] = 0.5 * (lo + hi);
}

// Inverse iteration.  Thread per eigenvalue; a thread whose eigenvalue starts
// a cluster (gap to the previous one above ctol*|T|) computes the whole
// cluster in order, orthogonalising each vector against the earlier ones.
__global__ void stein_kernel(const EigDesc* __restrict__ descs, const int2* __restrict__ work,
                             int nwork, const int64_t* __restrict__ woff) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nwork) return;
    const int2 wk = work[id];
    const EigDesc d = descs[wk.x];
    const int n = d.n;
    const int j0 = wk.y;
    double lo, hi, pivmin, tn;
    gersh(d.d, d.e, n, lo, hi, pivmin, tn);
    const double ctol = 1e-7 * tn;
    if (j0 > 0 && fabs(d.lam[j0 - 1] - d.lam[j0]) <= ctol) return;   // not a leader
    int j1 = j0 + 1;
    while (j1 < d.k && fabs(d.lam[j1 - 1] - d.lam[j1]) <= ctol) ++j1;
    double* ud = d.work + woff[id];
    double* u1 = ud + n;
    double* u2 = u1 + n;
    double* ml = u2 + n;
    double* pv = ml + n;
    const double tiny = DBL_EPSILON * fmax(tn, DBL_MIN);
    for (int j = j0; j < j1; ++j) {
        // small perturbation inside a cluster keeps the factorisations distinct
        const double lam = d.lam[j] + (j - j0) * 10.0 * DBL_EPSILON * tn;
        // ---- LU with partial pivoting of T - lam I
        ud[0] = d.d[0] - lam;
        u1[0] = n > 1 ? d.e[0] : 0.0;
        u2[0] = 0.0;
        for (int k = 0; k + 1 < n; ++k) {
            const double sub = d.e[k];
            const double dg = d.d[k + 1] - lam;
            const double sp = k + 2 < n ? d.e[k + 1] : 0.0;
            if (fabs(ud[k]) >= fabs(sub)) {
                double p = ud[k];
                if (fabs(p) < tiny) p = copysign(tiny, p == 0.0 ? 1.0 : p);
                ud[k] = p;
                const double l = sub / p;
                ml[k] = l;
                pv[k] = 0.0;
                ud[k + 1] = dg - l * u1[k];
                u1[k + 1] = sp - l * u2[k];
                u2[k + 1] = 0.0;
            } else {
                const double l = ud[k] / sub;
                const double o1 = u1[k], o2 = u2[k];
                ml[k] = l;
                pv[k] = 1.0;
                ud[k] = sub;
                u1[k] = dg;
                u2[k] = sp;
                ud[k + 1] = o1 - l * dg;
                u1[k + 1] = o2 - l * sp;
                u2[k + 1] = 0.0;
            }
        }
        if (fabs(ud[n - 1]) < tiny) ud[n - 1] = copysign(tiny, ud[n - 1] == 0.0 ? 1.0 : ud[n - 1]);
        double* x = d.Y + (int64_t)j * n;
        // deterministic start vector
        uint32_t s = 2463534242u ^ (uint32_t)(j * 2654435761u) ^ (uint32_t)(wk.x * 40503u);
        for (int r = 0; r < n; ++r) {
            s ^= s << 13; s ^= s >> 17; s ^= s << 5;
            x[r] = 1.0 + (double)(s & 0xffff) * (1.0 / 65536.0);
        }
        for (int it = 0; it < 4; ++it) {
            // forward: apply row interchanges / multipliers
            for (int k = 0; k + 1 < n; ++k) {
                if (pv[k] != 0.0) {
                    const double tmp = x[k];
                    x[k] = x[k + 1];
                    x[k + 1] = tmp;
                }
                x[k + 1] -= ml[k] * x[k];
            }
            // back substitution with U (bandwidth 3)
            x[n - 1] = x[n - 1] / ud[n - 1];
            if (n > 1) x[n - 2] = (x[n - 2] - u1[n - 2] * x[n - 1]) / ud[n - 2];
            for (int k = n - 3; k >= 0; --k)
                x[k] = (x[k] - u1[k] * x[k + 1] - u2[k] * x[k + 2]) / ud[k];
            // orthogonalise against earlier members of the cluster
            for (int q = j0; q < j; ++q) {
                const double* y = d.Y + (int64_t)q * n;
                double dot = 0.0;
                for (int r = 0; r < n; ++r) dot += x[r] * y[r];
                for (int r = 0; r < n; ++r) x[r] -= dot * y[r];
            }
            double nrm = 0.0, amax = 0.0;
            for (int r = 0; r < n; ++r) nrm += x[r] * x[r];
            nrm = sqrt(nrm);
            double sgn = 1.0;
            for (int r = 0; r < n; ++r)
                if (fabs(x[r]) > amax) { amax = fabs(x[r]); sgn = x[r] < 0.0 ? -1.0 : 1.0; }
            const double sc = sgn / nrm;
            for (int r = 0; r < n; ++r) x[r] *= sc;
        }
    }
}

// y := Q y for the reflectors stored by the tridiagonalisation; one wave per
// (matrix, column).
__global__ void __launch_bounds__(256) ormtr_kernel(const EigDesc* __restrict__ descs,
                                                    const int2* __restrict__ work, int nwork) {
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid >= nwork) return;
    const int2 wk = work[wid];
    const EigDesc d = descs[wk.x];
    const int n = d.n;
    double* y = d.Y + (int64_t)wk.y * n;
    const double* A = d.A;
    const int64_t lda = d.lda;
    for (int i = n - 2; i >= 0; --i) {
        const double tau = d.tau[i];
        if (tau == 0.0) continue;
        double s = 0.0;
        for (int r = i + 1 + lane; r < n; r += 64) {
            const double vr = r == i + 1 ? 1.0 : A[r + (int64_t)i * lda];
            s += vr * y[r];
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        s *= tau;
        for (int r = i + 1 + lane; r < n; r += 64) {
            const double vr = r == i + 1 ? 1.0 : A[r + (int64_t)i * lda];
            y[r] -= s * vr;
        }
    }
}

}  // namespace dev

void launch_sytrd_step(const dev::EigDesc* d, const int2* work, int nwork, int i, double* sv,
                       double* sp, const int64_t* soff, int nmax, hipStream_t s) {
    if (nwork <= 0) return;
    const size_t lds = (size_t)3 * nmax * sizeof(double);
    hipLaunchKernelGGL(dev::sytrd_step_kernel, dim3(nwork), dim3(256), lds, s, d, work, i, sv, sp,
                       soff);
}

void launch_eig_count(const dev::EigDesc* d, int nmat, const double* thr, int32_t* counts,
                      hipStream_t s) {
    if (nmat <= 0) return;
    hipLaunchKernelGGL(dev::eig_count_kernel, dim3((nmat + 63) / 64), dim3(64), 0, s, d, nmat, thr,
                       counts);
}

void launch_bisect(const dev::EigDesc* d, const int2* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::bisect_kernel, dim3((nwork + 63) / 64), dim3(64), 0, s, d, work, nwork);
}

void launch_stein(const dev::EigDesc* d, const int2* work, int nwork, const int64_t* woff,
                  hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::stein_kernel, dim3((nwork + 63) / 64), dim3(64), 0, s, d, work, nwork,
                       woff);
}

void launch_ormtr(const dev::EigDesc* d, const int2* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::ormtr_kernel, dim3((nwork + 3) / 4), dim3(256), 0, s, d, work, nwork);
}

}  // namespace amls
