// Shift-invert Lanczos steps for the local eigenproblems that the
// reference's EigenpairSolver sends to sil::ShiftInvertLanczos
// (src/amls/eigenpair_solver.hpp:38-82, src/amls/shift_invert_lanczos.hpp
// :44-85).
//
// The reference runs M-inner-product Lanczos on C = (K - sigma M)^-1 M in
// the original coordinates x.  With M = R R^T, u = R^T x turns it into the
// standard Lanczos process on the symmetric S = R^T (K - sigma M)^-1 R: the
// same alpha, beta and Ritz values, and x = R^-T u.  The host forms S by
// batched factorisations and GEMMs; this kernel runs the sequential part,
// one workgroup per local problem:
//   w = S u_j;  alpha_j = u_j^T w;
//   two passes of full reorthogonalisation against u_0 .. u_j (classical
//   Gram-Schmidt: coefficients first, then one update; the reference's
//   modified form computes the same numbers in exact arithmetic);
//   beta_j = ||w||; stop at beta_j < 1e-13 (breakdown) or after step ncv-1;
//   u_{j+1} = w / beta_j.
// u_{j+1} and beta_j are stored after the last requested step as well, so a
// later call with a larger ncv continues the same run; every reduction has a
// fixed order, so continuing gives the numbers a fresh, longer run gives.
#include "common.hpp"
#include "launch.hpp"

namespace amls {
namespace dev {

constexpr int LZ_THREADS = 1024;
constexpr int LZ_WAVES = LZ_THREADS / 64;
constexpr int LZ_NMAX = 3200;
constexpr int LZ_CMAX = 1536;

__device__ __forceinline__ double lz_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum over the workgroup, the same value (same order) on every thread
__device__ __forceinline__ double lz_block_sum(double v, double* red) {
    v = lz_wave_sum(v);
    __syncthreads();   // earlier readers of red are done
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < LZ_WAVES; ++i) s += red[i];
    return s;
}

__global__ void __launch_bounds__(LZ_THREADS) lanczos_kernel(const LanczosDesc* __restrict__ descs) {
    extern __shared__ double lds[];
    const LanczosDesc d = descs[blockIdx.x];
    const int n = d.n;
    int m = d.state[0];
    if (d.state[1] != 0 || m >= d.ncv || n <= 0) return;   // uniform over the block
    double* red = lds;            // LZ_WAVES partial sums
    double* su = lds + 32;        // u_j
    double* sw = su + n;          // w
    double* sc = sw + n;          // reorthogonalisation coefficients (<= ncv)
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int cap = min(n, LZ_CMAX);
    if (m == 0) {
        double s = 0.0;
        for (int r = t; r < n; r += LZ_THREADS) {
            const double v = d.u0[r];
            su[r] = v;
            s += v * v;
        }
        // the reference's q0 / sqrt(max(q0^T M q0, 1e-300)), q0^T M q0 = |u0|^2
        const double nrm = sqrt(fmax(lz_block_sum(s, red), 1e-300));
        for (int r = t; r < n; r += LZ_THREADS) {
            const double v = su[r] / nrm;
            su[r] = v;
            d.U[r] = v;
        }
    } else {
        for (int r = t; r < n; r += LZ_THREADS) su[r] = d.U[r + (int64_t)m * n];
    }
    __syncthreads();
    for (int j = m; j < d.ncv; ++j) {
        // w = S u_j: thread per row (column-major S, consecutive rows on
        // consecutive lanes), u_j broadcast from LDS
        for (int r = t; r < n; r += LZ_THREADS) {
            const double* row = d.S + r;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            int c = 0;
            for (; c + 4 <= n; c += 4) {
                a0 = fma(row[(int64_t)c * n], su[c], a0);
                a1 = fma(row[(int64_t)(c + 1) * n], su[c + 1], a1);
                a2 = fma(row[(int64_t)(c + 2) * n], su[c + 2], a2);
                a3 = fma(row[(int64_t)(c + 3) * n], su[c + 3], a3);
            }
            for (; c < n; ++c) a0 = fma(row[(int64_t)c * n], su[c], a0);
            sw[r] = (a0 + a1) + (a2 + a3);
        }
        __syncthreads();
        double s = 0.0;
        for (int r = t; r < n; r += LZ_THREADS) s += su[r] * sw[r];
        const double aj = lz_block_sum(s, red);
        for (int pass = 0; pass < 2; ++pass) {
            // sc[l] = u_l^T w, one wave per l (coalesced column reads)
            for (int l = wv; l <= j; l += LZ_WAVES) {
                const double* ul = d.U + (int64_t)l * n;
                double acc = 0.0;
                for (int r = lane; r < n; r += 64) acc = fma(ul[r], sw[r], acc);
                acc = lz_wave_sum(acc);
                if (lane == 0) sc[l] = acc;
            }
            __syncthreads();
            // w -= U sc, thread per row
            for (int r = t; r < n; r += LZ_THREADS) {
                double a0 = 0.0, a1 = 0.0;
                int l = 0;
                for (; l + 2 <= j + 1; l += 2) {
                    a0 = fma(d.U[r + (int64_t)l * n], sc[l], a0);
                    a1 = fma(d.U[r + (int64_t)(l + 1) * n], sc[l + 1], a1);
                }
                if (l <= j) a0 = fma(d.U[r + (int64_t)l * n], sc[l], a0);
                sw[r] -= a0 + a1;
            }
            __syncthreads();
        }
        s = 0.0;
        for (int r = t; r < n; r += LZ_THREADS) s += sw[r] * sw[r];
        const double bj = sqrt(fmax(lz_block_sum(s, red), 0.0));
        if (t == 0) d.alpha[j] = aj;
        m = j + 1;
        if (bj < 1e-13) {
            if (t == 0) d.state[1] = 1;
            break;
        }
        if (t == 0) d.beta[j] = bj;
        if (j + 1 < cap) {
            for (int r = t; r < n; r += LZ_THREADS) {
                const double v = sw[r] / bj;
                su[r] = v;
                d.U[r + (int64_t)(j + 1) * n] = v;
            }
        }
        __syncthreads();
    }
    if (t == 0) d.state[0] = m;
}

}  // namespace dev

int lanczos_max_order() { return dev::LZ_NMAX; }

void launch_lanczos(const dev::LanczosDesc* d, int count, int nmax, int cmax, hipStream_t s) {
    if (count <= 0) return;
    const size_t lds = (size_t)(32 + 2 * nmax + cmax) * sizeof(double);
    hipLaunchKernelGGL(dev::lanczos_kernel, dim3(count), dim3(dev::LZ_THREADS), lds, s, d);
}

}  // namespace amls
