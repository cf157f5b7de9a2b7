// CSR products and column reductions (HBM-bound).
//
// Replaces the M-normalisation SpMVs of AmlsSolver::solve
// (src/amls/solver.cpp:307-310) and provides the SpMM used by the subspace
// iteration and the residual checks.
#include "common.hpp"
#include "launch.hpp"

namespace amls {
namespace dev {

// One wave per row; lanes stride the row's nonzeros and the columns of X are
// processed in chunks of 8 so each nonzero is read once per chunk.
__global__ void __launch_bounds__(256) spmm_kernel(int64_t n, const int64_t* __restrict__ rowptr,
                                                   const int64_t* __restrict__ colidx,
                                                   const double* __restrict__ vals,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ y, int ncols, int64_t ld) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int64_t b = rowptr[row], e = rowptr[row + 1];
    for (int c0 = 0; c0 < ncols; c0 += 8) {
        double acc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = 0.0;
        for (int64_t k = b + lane; k < e; k += 64) {
            const double a = vals[k];
            const int64_t col = colidx[k];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (c0 + q < ncols) acc[q] += a * x[col + (int64_t)(c0 + q) * ld];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double v = acc[q];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0 && c0 + q < ncols) y[row + (int64_t)(c0 + q) * ld] = v;
        }
    }
}

constexpr unsigned NXCD = 8;   // MI355X: 8 XCDs, workgroups dealt round-robin

// Row-major operands: Y(r, :) = sum_e val[e] X(col[e], :), one wave per
// (row, 64-column chunk), lanes over contiguous columns so every nonzero is
// one coalesced 512-byte read of X.  Chunks are the slow grid dimension, so
// the whole GPU streams through one 64-column slab of X at a time.  The grid
// (a multiple of 8 in x) is remapped so that each XCD takes one contiguous
// eighth of the rows: the X rows its rows touch stay in that XCD's L2 instead
// of being fetched by all eight.
__global__ void __launch_bounds__(256) spmm_rm_kernel(int64_t n, const int64_t* __restrict__ rowptr,
                                                      const int64_t* __restrict__ colidx,
                                                      const double* __restrict__ vals,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y, int ncols) {
    const unsigned per = gridDim.x / NXCD;
    const unsigned blk = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
    const int64_t row = (int64_t)blk * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int c = blockIdx.y * 64 + lane;
    const bool on = c < ncols;
    const int cc = on ? c : 0;   // in-bounds column for the loads of idle lanes
    const int64_t b = rowptr[row], e = rowptr[row + 1];
    // eight nonzeros per round: their indices, then their X rows, are all in
    // flight together instead of one dependent round trip per nonzero
    double acc0 = 0.0, acc1 = 0.0;
    int64_t k = b;
    for (; k + 8 <= e; k += 8) {
        int64_t col[8];
        double a[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            col[u] = colidx[k + u];
            a[u] = vals[k + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[col[u] * ncols + cc];
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            acc0 += a[u] * xv[u];
            acc1 += a[u + 1] * xv[u + 1];
        }
    }
    for (; k < e; ++k) acc0 += vals[k] * x[colidx[k] * ncols + cc];
    if (on) y[row * ncols + c] = acc0 + acc1;
}

// 3 x 3 block-sparse rows (elasticity: one block per pair of coupled mesh
// nodes): Y(3b + i, :) = sum over blocks (b, c) of A_bc(i, :) X(3c .. 3c+2, :).
// One wave per block row.  The wave stages its row's blocks (values and
// column indices, at most SPM_MAXB blocks) in LDS once and then sweeps every
// column of X in chunks of 128 (two columns per lane), so the matrix is read
// from memory once instead of once per chunk; the block values are LDS
// broadcasts.  Each block costs six coalesced 512-byte reads of X for 18
// FMAs per lane, four blocks in flight at a time.  Rows with more blocks
// than SPM_MAXB read the matrix from global memory per chunk.  Same
// XCD-contiguous row mapping as spmm_rm_kernel.
constexpr int SPM_MAXB = 128;
constexpr int SPM_CPL = 2;   // columns per lane

template <bool Lds>
__device__ __forceinline__ void bsr3_row(const double* __restrict__ av, const int* __restrict__ ac,
                                         int nbk, const double* __restrict__ x,
                                         double* __restrict__ y, int64_t br, int ncols, int lane) {
    for (int c0 = 0; c0 < ncols; c0 += 64 * SPM_CPL) {
        int cc[SPM_CPL];
        bool on[SPM_CPL];
#pragma unroll
        for (int q = 0; q < SPM_CPL; ++q) {
            const int c = c0 + lane + 64 * q;
            on[q] = c < ncols;
            cc[q] = on[q] ? c : 0;   // in-bounds column for the loads of idle lanes
        }
        double a0[SPM_CPL], a1[SPM_CPL], a2[SPM_CPL];
#pragma unroll
        for (int q = 0; q < SPM_CPL; ++q) a0[q] = a1[q] = a2[q] = 0.0;
        int k = 0;
        for (; k + 4 <= nbk; k += 4) {
            double xv[4][3][SPM_CPL];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t xr = ac[k + u];
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int q = 0; q < SPM_CPL; ++q) xv[u][j][q] = x[(xr + j) * ncols + cc[q]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double* a = av + 9 * (k + u);
#pragma unroll
                for (int q = 0; q < SPM_CPL; ++q) {
                    a0[q] += a[0] * xv[u][0][q] + a[1] * xv[u][1][q] + a[2] * xv[u][2][q];
                    a1[q] += a[3] * xv[u][0][q] + a[4] * xv[u][1][q] + a[5] * xv[u][2][q];
                    a2[q] += a[6] * xv[u][0][q] + a[7] * xv[u][1][q] + a[8] * xv[u][2][q];
                }
            }
        }
        for (; k < nbk; ++k) {
            const int64_t xr = ac[k];
            const double* a = av + 9 * k;
#pragma unroll
            for (int q = 0; q < SPM_CPL; ++q) {
                const double x0 = x[xr * ncols + cc[q]], x1 = x[(xr + 1) * ncols + cc[q]],
                             x2 = x[(xr + 2) * ncols + cc[q]];
                a0[q] += a[0] * x0 + a[1] * x1 + a[2] * x2;
                a1[q] += a[3] * x0 + a[4] * x1 + a[5] * x2;
                a2[q] += a[6] * x0 + a[7] * x1 + a[8] * x2;
            }
        }
#pragma unroll
        for (int q = 0; q < SPM_CPL; ++q)
            if (on[q]) {
                const int c = c0 + lane + 64 * q;
                y[(3 * br) * ncols + c] = a0[q];
                y[(3 * br + 1) * ncols + c] = a1[q];
                y[(3 * br + 2) * ncols + c] = a2[q];
            }
    }
}

__global__ void __launch_bounds__(256) spmm_bsr3_kernel(int64_t nb, const int64_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ col,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y, int ncols) {
    __shared__ double sv[4][SPM_MAXB * 9];
    __shared__ int sc[4][SPM_MAXB];
    const unsigned per = gridDim.x / NXCD;
    const unsigned blk = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
    const int w = threadIdx.x >> 6;
    const int64_t br = (int64_t)blk * 4 + w;
    const int lane = threadIdx.x & 63;
    if (br >= nb) return;
    const int64_t b0 = ptr[br], b1 = ptr[br + 1];
    const int nbk = (int)(b1 - b0);
    if (nbk > SPM_MAXB) {
        // X row offsets of the blocks are recomputed from col per chunk
        for (int c0 = 0; c0 < ncols; c0 += 64) {
            const int c = c0 + lane;
            const bool on = c < ncols;
            const int cc = on ? c : 0;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int64_t k = b0; k < b1; ++k) {
                const int64_t xr = 3 * (int64_t)col[k];
                const double x0 = x[xr * ncols + cc], x1 = x[(xr + 1) * ncols + cc],
                             x2 = x[(xr + 2) * ncols + cc];
                const double* a = val + 9 * k;
                a0 += a[0] * x0 + a[1] * x1 + a[2] * x2;
                a1 += a[3] * x0 + a[4] * x1 + a[5] * x2;
                a2 += a[6] * x0 + a[7] * x1 + a[8] * x2;
            }
            if (on) {
                y[(3 * br) * ncols + c] = a0;
                y[(3 * br + 1) * ncols + c] = a1;
                y[(3 * br + 2) * ncols + c] = a2;
            }
        }
        return;
    }
    // the wave's own LDS slice: written and read by this wave only
    for (int e = lane; e < nbk * 9; e += 64) sv[w][e] = val[9 * b0 + e];
    for (int e = lane; e < nbk; e += 64) sc[w][e] = 3 * col[b0 + e];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bsr3_row<true>(sv[w], sc[w], nbk, x, y, br, ncols, lane);
}

// One block per column.
__global__ void __launch_bounds__(256) coldot_kernel(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, int64_t ld,
                                                     double* __restrict__ out) {
    __shared__ double red[4];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += 256) s += x[r + c * ld] * y[r + c * ld];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[c] = red[0] + red[1] + red[2] + red[3];
}

// Per problem segment: out[s + c nseg] = sum over rows [seg[s], seg[s+1]) of
// x(r, c) y(r, c).  One block per (segment, column).
// segMajor = 0: out[s + c nseg]; 1: out[s ncols + c]
__global__ void __launch_bounds__(256) coldot_seg_kernel(const int64_t* __restrict__ seg, int nseg,
                                                         const double* __restrict__ x,
                                                         const double* __restrict__ y, int64_t ld,
                                                         double* __restrict__ out, int segMajor) {
    __shared__ double red[4];
    const int sg = blockIdx.x, c = blockIdx.y;
    const int64_t r0 = seg[sg], r1 = seg[sg + 1];
    double s = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) s += x[r + c * ld] * y[r + c * ld];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        out[segMajor ? (int64_t)sg * gridDim.y + c : sg + (int64_t)c * nseg] =
            red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) colscale_seg_kernel(const int64_t* __restrict__ seg,
                                                           int nseg, double* __restrict__ x,
                                                           int64_t ld,
                                                           const double* __restrict__ nrm) {
    const int sg = blockIdx.x, c = blockIdx.y;
    const double sc = 1.0 / sqrt(fmax(nrm[sg + (int64_t)c * nseg], 1e-300));
    for (int64_t r = seg[sg] + threadIdx.x; r < seg[sg + 1]; r += 256) x[r + c * ld] *= sc;
}

__global__ void colscale_kernel(int64_t n, double* __restrict__ x, int64_t ld,
                                const double* __restrict__ nrm) {
    const int c = blockIdx.y;
    const double s = 1.0 / sqrt(fmax(nrm[c], 1e-300));
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256)
        x[r + c * ld] *= s;
}

__global__ void permute_rows_kernel(int64_t n, const int64_t* __restrict__ perm,
                                    const double* __restrict__ src, double* __restrict__ dst,
                                    int64_t ld) {
    const int c = blockIdx.y;
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256)
        dst[perm[j] + c * ld] = src[j + c * ld];
}

}  // namespace dev

void launch_spmm(int64_t n, const int64_t* rowptr, const int64_t* colidx, const double* vals,
                 const double* x, double* y, int ncols, int64_t ld, hipStream_t s) {
    if (n <= 0 || ncols <= 0) return;
    hipLaunchKernelGGL(dev::spmm_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, n, rowptr,
                       colidx, vals, x, y, ncols, ld);
}

void launch_spmm_rm(int64_t n, const int64_t* rowptr, const int64_t* colidx, const double* vals,
                    const double* x, double* y, int ncols, hipStream_t s) {
    if (n <= 0 || ncols <= 0) return;
    const unsigned nb = (unsigned)((n + 3) / 4);
    const unsigned nbx = (nb + dev::NXCD - 1) / dev::NXCD * dev::NXCD;
    hipLaunchKernelGGL(dev::spmm_rm_kernel, dim3(nbx, (unsigned)((ncols + 63) / 64)), dim3(256), 0, s,
                       n, rowptr, colidx, vals, x, y, ncols);
}

void launch_spmm_bsr3(int64_t nb, const int64_t* ptr, const int32_t* col, const double* val,
                      const double* x, double* y, int ncols, hipStream_t s) {
    if (nb <= 0 || ncols <= 0) return;
    const unsigned g = (unsigned)((nb + 3) / 4);
    const unsigned gx = (g + dev::NXCD - 1) / dev::NXCD * dev::NXCD;
    hipLaunchKernelGGL(dev::spmm_bsr3_kernel, dim3(gx), dim3(256), 0, s, nb, ptr, col, val, x, y,
                       ncols);
}

void launch_coldot(int64_t n, const double* x, const double* y, int ncols, int64_t ld, double* out,
                   hipStream_t s) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(dev::coldot_kernel, dim3(ncols), dim3(256), 0, s, n, x, y, ld, out);
}

void launch_colnorm_seg(const int64_t* seg, int nseg, double* x, double* y, int ncols,
                        int64_t ld, double* nrm, hipStream_t s, bool scaleY) {
    if (ncols <= 0 || nseg <= 0) return;
    hipLaunchKernelGGL(dev::coldot_seg_kernel, dim3(nseg, ncols), dim3(256), 0, s, seg, nseg, x, y,
                       ld, nrm, 0);
    hipLaunchKernelGGL(dev::colscale_seg_kernel, dim3(nseg, ncols), dim3(256), 0, s, seg, nseg, x,
                       ld, nrm);
    if (scaleY)
        hipLaunchKernelGGL(dev::colscale_seg_kernel, dim3(nseg, ncols), dim3(256), 0, s, seg, nseg,
                           y, ld, nrm);
}

void launch_coldot_seg(const int64_t* seg, int nseg, const double* x, const double* y, int ncols,
                       int64_t ld, double* out, hipStream_t s) {
    if (ncols <= 0 || nseg <= 0) return;
    hipLaunchKernelGGL(dev::coldot_seg_kernel, dim3(nseg, ncols), dim3(256), 0, s, seg, nseg, x, y,
                       ld, out, 1);
}

void launch_colscale_rsqrt(int64_t n, double* x, int ncols, int64_t ld, const double* nrm,
                           hipStream_t s) {
    if (ncols <= 0 || n <= 0) return;
    unsigned bx = (unsigned)((n + 255) / 256);
    if (bx > 64) bx = 64;
    hipLaunchKernelGGL(dev::colscale_kernel, dim3(bx, ncols), dim3(256), 0, s, n, x, ld, nrm);
}

void launch_permute_rows(int64_t n, const int64_t* perm, const double* src, double* dst, int ncols,
                         int64_t ld, hipStream_t s) {
    if (ncols <= 0 || n <= 0) return;
    unsigned bx = (unsigned)((n + 255) / 256);
    if (bx > 64) bx = 64;
    hipLaunchKernelGGL(dev::permute_rows_kernel, dim3(bx, ncols), dim3(256), 0, s, n, perm, src, dst,
                       ld);
}

}  // namespace amls
