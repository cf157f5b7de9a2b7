// Shift-invert Lanczos steps for the local eigenproblems that the
// reference's EigenpairSolver sends to sil::ShiftInvertLanczos
// (src/amls/eigenpair_solver.hpp:38-82, src/amls/shift_invert_lanczos.hpp
// :44-85).
//
// The reference runs M-inner-product Lanczos on C = (K - sigma M)^-1 M in
// the original coordinates x.  With M = R R^T, u = R^T x turns it into the
// standard Lanczos process on the symmetric S = R^T (K - sigma M)^-1 R: the
// same alpha, beta and Ritz values, and x = R^-T u.  The host forms S by
// batched factorisations and GEMMs; these kernels run the sequential part.
// One Lanczos step of every active node is six launches, each spread over
// the whole device (a node's basis is n x up to 1,536 doubles, and the
// step reads it twice per reorthogonalisation pass, which one workgroup
// per node could only stream at a single CU's bandwidth):
//   lz_matvec:  w = S u_j, four rows per wave (S is stored symmetric, so
//               row r is the contiguous column r);
//   lz_coef:    c_l = u_l^T w, one wave per basis vector (alpha_j = c_j of
//               the first pass);
//   lz_update:  w -= U c, a thread per row, c staged in LDS; after the
//               second pass each block's part of |w|^2;
//   (lz_coef and lz_update twice: two passes of classical Gram-Schmidt;
//   the reference's modified form computes the same numbers in exact
//   arithmetic)
//   lz_finish:  beta_j = |w| from the block parts in order; stop at
//               beta_j < 1e-13 (breakdown) or after step ncv-1;
//               u_{j+1} = w / beta_j.
// u_{j+1} and beta_j are stored after the last requested step as well, so a
// later call with a larger ncv continues the same run; every reduction has a
// fixed order, so continuing gives the numbers a fresh, longer run gives.
#include "common.hpp"
#include "launch.hpp"

namespace amls {
namespace dev {

constexpr int LZ_NMAX = 3200;
constexpr int LZ_CMAX = 1536;
constexpr int LZ_ROWS = 16;    // rows per lz_matvec workgroup (four per wave)
constexpr int LZ_UPD = 256;    // rows per lz_update workgroup

__device__ __forceinline__ double lz_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// the step j of node d, or -1 when the node has stopped (uniform per block)
__device__ __forceinline__ int lz_step(const LanczosDesc& d) {
    const int j = d.state[0];
    return (d.state[1] != 0 || j >= d.ncv || d.n <= 0) ? -1 : j;
}

// u_0 = u0 / sqrt(max(u0^T u0, 1e-300)) (the reference's q0 normalised in
// the M inner product, u0 = R^T q0) for nodes that have not started
__global__ void __launch_bounds__(256) lz_init_kernel(const LanczosDesc* __restrict__ descs) {
    __shared__ double red[4];
    const LanczosDesc d = descs[blockIdx.x];
    if (lz_step(d) != 0) return;
    const int t = threadIdx.x;
    double s = 0.0;
    for (int r = t; r < d.n; r += 256) s = fma(d.u0[r], d.u0[r], s);
    s = lz_wave_sum(s);
    if ((t & 63) == 0) red[t >> 6] = s;
    __syncthreads();
    const double nrm = sqrt(fmax(((red[0] + red[1]) + red[2]) + red[3], 1e-300));
    for (int r = t; r < d.n; r += 256) d.U[r] = d.u0[r] / nrm;
}

__global__ void __launch_bounds__(256) lz_matvec_kernel(const LanczosDesc* __restrict__ descs) {
    const LanczosDesc d = descs[blockIdx.y];
    const int j = lz_step(d);
    const int r0 = blockIdx.x * LZ_ROWS + 4 * (threadIdx.x >> 6);
    if (j < 0 || r0 >= d.n) return;
    const int lane = threadIdx.x & 63, n = d.n;
    const double* u = d.U + (int64_t)j * n;
    const double* c0 = d.S + (int64_t)r0 * n;
    const int nr = min(4, n - r0);
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = lane; c < n; c += 64) {
        const double uc = u[c];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (i < nr) a[i] = fma(c0[(int64_t)i * n + c], uc, a[i]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double v = lz_wave_sum(a[i]);
        if (lane == 0 && i < nr) d.W[r0 + i] = v;
    }
}

__global__ void __launch_bounds__(256) lz_coef_kernel(const LanczosDesc* __restrict__ descs,
                                                      int pass) {
    const LanczosDesc d = descs[blockIdx.y];
    const int j = lz_step(d);
    const int l = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j < 0 || l > j) return;
    const int lane = threadIdx.x & 63, n = d.n;
    const double* ul = d.U + (int64_t)l * n;
    double acc = 0.0;
    for (int r = lane; r < n; r += 64) acc = fma(ul[r], d.W[r], acc);
    acc = lz_wave_sum(acc);
    if (lane == 0) {
        d.C[l] = acc;
        if (pass == 0 && l == j) d.alpha[j] = acc;
    }
}

__global__ void __launch_bounds__(LZ_UPD) lz_update_kernel(const LanczosDesc* __restrict__ descs,
                                                           int last) {
    __shared__ double sc[LZ_CMAX];
    __shared__ double red[LZ_UPD / 64];
    const LanczosDesc d = descs[blockIdx.y];
    const int j = lz_step(d);
    const int n = d.n;
    if (j < 0 || (int)blockIdx.x * LZ_UPD >= n) return;
    const int t = threadIdx.x;
    for (int l = t; l <= j; l += LZ_UPD) sc[l] = d.C[l];
    __syncthreads();
    const int r = blockIdx.x * LZ_UPD + t;
    double w = 0.0;
    if (r < n) {
        double a0 = 0.0, a1 = 0.0;
        int l = 0;
        for (; l + 2 <= j + 1; l += 2) {
            a0 = fma(d.U[r + (int64_t)l * n], sc[l], a0);
            a1 = fma(d.U[r + (int64_t)(l + 1) * n], sc[l + 1], a1);
        }
        if (l <= j) a0 = fma(d.U[r + (int64_t)l * n], sc[l], a0);
        w = d.W[r] - (a0 + a1);
        d.W[r] = w;
    }
    if (last) {
        double s = lz_wave_sum(w * w);
        if ((t & 63) == 0) red[t >> 6] = s;
        __syncthreads();
        if (t == 0) d.part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
    }
}

__global__ void __launch_bounds__(256) lz_finish_kernel(const LanczosDesc* __restrict__ descs) {
    __shared__ double sb;
    const LanczosDesc d = descs[blockIdx.x];
    const int j = lz_step(d);
    if (j < 0) return;
    const int n = d.n, t = threadIdx.x;
    if (t == 0) {
        double s = 0.0;
        for (int b = 0; b * LZ_UPD < n; ++b) s += d.part[b];
        sb = sqrt(fmax(s, 0.0));
    }
    __syncthreads();
    const double bj = sb;
    const int cap = min(n, LZ_CMAX);
    if (bj < 1e-13) {
        __syncthreads();   // every thread has read the state before it changes
        if (t == 0) {
            d.state[1] = 1;
            d.state[0] = j + 1;
        }
        return;
    }
    if (j + 1 < cap)
        for (int r = t; r < n; r += 256) d.U[r + (int64_t)(j + 1) * n] = d.W[r] / bj;
    __syncthreads();
    if (t == 0) {
        d.beta[j] = bj;
        d.state[0] = j + 1;
    }
}

}  // namespace dev

int lanczos_max_order() { return dev::LZ_NMAX; }
int lanczos_max_ncv() { return dev::LZ_CMAX; }

void launch_lanczos_init(const dev::LanczosDesc* d, int count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(dev::lz_init_kernel, dim3(count), dim3(256), 0, s, d);
}

// one step of every active node: nmax = largest order, jmax = largest step
// index any node may be at
void launch_lanczos_step(const dev::LanczosDesc* d, int count, int nmax, int jmax, hipStream_t s) {
    if (count <= 0 || nmax <= 0) return;
    const unsigned nb = (unsigned)count;
    hipLaunchKernelGGL(dev::lz_matvec_kernel, dim3((nmax + dev::LZ_ROWS - 1) / dev::LZ_ROWS, nb),
                       dim3(256), 0, s, d);
    const unsigned gc = (unsigned)((jmax + 1 + 3) / 4), gu = (unsigned)((nmax + dev::LZ_UPD - 1) / dev::LZ_UPD);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(dev::lz_coef_kernel, dim3(gc, nb), dim3(256), 0, s, d, pass);
        hipLaunchKernelGGL(dev::lz_update_kernel, dim3(gu, nb), dim3(dev::LZ_UPD), 0, s, d, pass);
    }
    hipLaunchKernelGGL(dev::lz_finish_kernel, dim3(nb), dim3(256), 0, s, d);
}

}  // namespace amls
