// Small kernels of the Chebyshev-filtered subspace iteration: a reproducible
// uniform start block and the Ritz residual norms.
#include "common.hpp"
#include "launch.hpp"

namespace amls {
namespace dev {

// x[i] = uniform in [-1, 1) from a 64-bit mix (splitmix64 finaliser) of
// (seed, i): the same block for the same seed on every device and launch.
__global__ void __launch_bounds__(256) fill_hash_kernel(double* __restrict__ x, int64_t n,
                                                        uint64_t seed) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
}

// res[j] = || R(:, j) - lam[j] V(:, j) ||_2, R = C V (q x m, ld q); one
// workgroup per column.
__global__ void __launch_bounds__(256) ritz_resid_kernel(const double* __restrict__ R,
                                                         const double* __restrict__ V,
                                                         const double* __restrict__ lam, int64_t q,
                                                         double* __restrict__ res) {
    __shared__ double red[4];
    const int j = blockIdx.x;
    const double l = lam[j];
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < q; r += 256) {
        const double v = R[r + (int64_t)j * q] - l * V[r + (int64_t)j * q];
        s += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) res[j] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

}  // namespace dev

void launch_fill_hash(double* x, int64_t n, uint64_t seed, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(dev::fill_hash_kernel, dim3(g), dim3(256), 0, s, x, n, seed);
}

void launch_ritz_resid(const double* R, const double* V, const double* lam, int64_t q, int m,
                       double* res, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(dev::ritz_resid_kernel, dim3(m), dim3(256), 0, s, R, V, lam, q, res);
}

}  // namespace amls
