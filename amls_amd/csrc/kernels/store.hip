// Accumulation kernel of the contribution store: out += src over a
// rows x cols column-major block (each element one IEEE addition, so a sum
// of records in append order equals the host's element-wise sum bit for bit).
#include "common.hpp"
#include "launch.hpp"

namespace amls {
namespace dev {

// a 256-row slice of one column per block; grid (row blocks, columns)
__global__ void __launch_bounds__(256) add2d_kernel(double* __restrict__ out, int64_t ldo,
                                                    const double* __restrict__ src, int64_t lds,
                                                    int64_t rows, int64_t cols) {
    const int64_t c = blockIdx.y;
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * 256)
        out[r + c * ldo] += src[r + c * lds];
}

}  // namespace dev

void launch_add2d(double* out, int64_t ldo, const double* src, int64_t lds, int64_t rows,
                  int64_t cols, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return;
    // cols <= 65,535 per launch (grid y); the store's pieces stay far below
    for (int64_t c0 = 0; c0 < cols; c0 += 65535) {
        const int64_t nc = cols - c0 < 65535 ? cols - c0 : 65535;
        int64_t bx = (rows + 255) / 256;
        if (bx > 64) bx = 64;
        hipLaunchKernelGGL(dev::add2d_kernel, dim3((unsigned)bx, (unsigned)nc), dim3(256), 0, s,
                           out + c0 * ldo, ldo, src + c0 * lds, lds, rows, nc);
    }
}

}  // namespace amls
