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
// nodes): Y(3b + i, :) = sum over blocks (b, c) of A_bc(i, :) X(3c .. 3c+2, :),
// row-major X and Y.  Every lane holds SPM_CPL columns of a 64 SPM_CPL
// column chunk, so each X row triple is three coalesced 512-byte reads per
// column group, and the block values are LDS broadcasts.
constexpr int SPM_CPL = 2;   // columns per lane

// The block rows are taken four at a time ("quads"):
// the quad's four column lists are merged, so each distinct column block
// costs one set of X reads for all four rows (on the elasticity meshes a
// quad touches ~100 column blocks against ~226 blocks in its rows, so X
// traffic from L2 drops by ~2.3x).  Per quad: qptr -> merged columns ucol
// (X row offset 3c in bits 0..27, which of the four rows hold a block in
// bits 28..31), qvptr -> the blocks that exist, in merged-column order and
// row order within a column.  One workgroup per (quad, group of four column
// chunks), one chunk per wave; the column groups are the slow grid
// dimension, so the whole GPU streams through one 512-column slab of X at a
// time (the quad's blocks are read once per slab).  A quad's columns and
// blocks are staged in LDS when they fit (SPQ_MAXU / SPQ_MAXB), otherwise
// read from global memory.
constexpr int SPQ_MAXU = 160;
constexpr int SPQ_MAXB = 320;

template <bool Lds>
__device__ __forceinline__ void bsr3q_quad(const uint32_t* __restrict__ uc, int nu,
                                           const double* __restrict__ av,
                                           const double* __restrict__ x, double* __restrict__ y,
                                           int64_t q, int64_t nb, int ncols, int lane, int c_first,
                                           int c_step, int64_t ldy) {
    for (int c0 = c_first; c0 < ncols; c0 += c_step) {
        int cc[SPM_CPL];
        bool on[SPM_CPL];
#pragma unroll
        for (int k = 0; k < SPM_CPL; ++k) {
            const int c = c0 + lane + 64 * k;
            on[k] = c < ncols;
            cc[k] = on[k] ? c : 0;
        }
        double acc[4][3][SPM_CPL];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < SPM_CPL; ++k) acc[r][i][k] = 0.0;
        int bi = 0;
        // two merged columns per round: their X rows are in flight together
        int u = 0;
        for (; u + 2 <= nu; u += 2) {
            double xv[2][3][SPM_CPL];
            uint32_t w[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                w[h] = uc[u + h];
                const int64_t xr = w[h] & 0x0fffffffu;
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int k = 0; k < SPM_CPL; ++k) xv[h][j][k] = x[(xr + j) * ncols + cc[k]];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t mk = w[h] >> 28;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (mk & (1u << r)) {
                        const double* a = av + 9 * bi;
                        ++bi;
#pragma unroll
                        for (int k = 0; k < SPM_CPL; ++k) {
                            acc[r][0][k] += a[0] * xv[h][0][k] + a[1] * xv[h][1][k] + a[2] * xv[h][2][k];
                            acc[r][1][k] += a[3] * xv[h][0][k] + a[4] * xv[h][1][k] + a[5] * xv[h][2][k];
                            acc[r][2][k] += a[6] * xv[h][0][k] + a[7] * xv[h][1][k] + a[8] * xv[h][2][k];
                        }
                    }
            }
        }
        for (; u < nu; ++u) {
            const uint32_t w = uc[u];
            const int64_t xr = w & 0x0fffffffu;
            double xv[3][SPM_CPL];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < SPM_CPL; ++k) xv[j][k] = x[(xr + j) * ncols + cc[k]];
            const uint32_t mk = w >> 28;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (mk & (1u << r)) {
                    const double* a = av + 9 * bi;
                    ++bi;
#pragma unroll
                    for (int k = 0; k < SPM_CPL; ++k) {
                        acc[r][0][k] += a[0] * xv[0][k] + a[1] * xv[1][k] + a[2] * xv[2][k];
                        acc[r][1][k] += a[3] * xv[0][k] + a[4] * xv[1][k] + a[5] * xv[2][k];
                        acc[r][2][k] += a[6] * xv[0][k] + a[7] * xv[1][k] + a[8] * xv[2][k];
                    }
                }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t br = 4 * q + r;
            if (br < nb) {
#pragma unroll
                for (int k = 0; k < SPM_CPL; ++k)
                    if (on[k]) {
                        const int c = c0 + lane + 64 * k;
                        // ldy = 0: row-major Y; else column-major with leading dimension ldy
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            const int64_t e = ldy ? (3 * br + i) + (int64_t)c * ldy : (3 * br + i) * ncols + c;
                            y[e] = acc[r][i][k];
                        }
                    }
            }
        }
    }
}

__global__ void __launch_bounds__(256) spmm_bsr3q_kernel(int64_t nq, int64_t nb,
                                                         const int64_t* __restrict__ qptr,
                                                         const uint32_t* __restrict__ ucol,
                                                         const int64_t* __restrict__ qvptr,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y, int ncols,
                                                         int64_t ldy) {
    __shared__ double sv[SPQ_MAXB * 9];
    __shared__ uint32_t su[SPQ_MAXU];
    const unsigned per = gridDim.x / NXCD;
    const int64_t q = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    if (q >= nq) return;
    const int64_t u0 = qptr[q], b0 = qvptr[q];
    const int nu = (int)(qptr[q + 1] - u0), nbk = (int)(qvptr[q + 1] - b0);
    // column group blockIdx.y (four chunks), one chunk per wave
    const int cs = 64 * SPM_CPL;
    if (nu > SPQ_MAXU || nbk > SPQ_MAXB) {
        bsr3q_quad<false>(ucol + u0, nu, val + 9 * b0, x, y, q, nb, ncols, lane, (blockIdx.y * 4 + w) * cs, 1 << 30, ldy);
        return;
    }
    for (int e = t; e < nbk * 9; e += 256) sv[e] = val[9 * b0 + e];
    for (int e = t; e < nu; e += 256) su[e] = ucol[u0 + e];
    __syncthreads();
    bsr3q_quad<true>(su, nu, sv, x, y, q, nb, ncols, lane, (blockIdx.y * 4 + w) * cs, 1 << 30, ldy);
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

void launch_spmm_bsr3q(int64_t nq, int64_t nb, const int64_t* qptr, const uint32_t* ucol,
                       const int64_t* qvptr, const double* val, const double* x, double* y, int ncols,
                       hipStream_t s, int64_t ldy) {
    if (nq <= 0 || ncols <= 0) return;
    const unsigned gx = (unsigned)((nq + dev::NXCD - 1) / dev::NXCD * dev::NXCD);
    const unsigned gy = (unsigned)((ncols + 4 * 64 * dev::SPM_CPL - 1) / (4 * 64 * dev::SPM_CPL));
    hipLaunchKernelGGL(dev::spmm_bsr3q_kernel, dim3(gx, gy), dim3(256), 0, s, nq, nb, qptr, ucol, qvptr,
                       val, x, y, ncols, ldy);
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
