// Front assembly, partial LDL^T of fronts, extend-add and batched copies.
//
// Replaces the per-node factorisation of the reference
// (GaussianElimination::factorNode_, src/amls/gaussian_elimination.cpp:458-508,
// with Eigen SimplicialLDLT / LDLT in src/linalg/sparse_direct.cpp:10-63,
// including the "D > 0 and finite" check) and the block scatter of its Schur /
// congruence updates into ancestor blocks (:678-781), restated as a
// multifrontal extend-add so no two workgroups write one element.
#include "common.hpp"
#include "launch.hpp"
#include "packed_solve.hpp"

#include <climits>

namespace amls {
namespace dev {

constexpr int NB = 64;

// lane `src`'s value of v, as a wave-uniform value
__device__ __forceinline__ double lane_value(double v, int src) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// One wave per front with n > s0: unblocked LDL^T of the diagonal block of
// panel s0, lane i holding row i (its lower part) in registers; L11 (strict
// lower) and D (diagonal) are written back.
//
// Right-looking: step j computes l_i = a_ij / d_j on every lane i > j and
// keeps it in register j, then subtracts a_ij l_k from a_ik for every k > j,
// with l_k broadcast to every lane through LDS (lane i writes l_i, every
// lane reads l_k at the same address).  The workgroup is one wave, whose LDS
// operations complete in issue order, so no barrier is needed; the two
// alternating columns let one step's reads overlap the next step's write.
// Registers k > i of lane i (the upper part) collect harmless junk and are
// never read or stored.
__global__ void __launch_bounds__(64) ldlt_diag_kernel(const FrontDesc* __restrict__ descs,
                                                       const int32_t* __restrict__ list,
                                                       int s0, int* __restrict__ err) {
    const FrontDesc d = descs[list[blockIdx.x]];
    const int64_t f = d.f;
    const int jb = min(NB, d.n - s0);
    double* F = d.F + s0 + (int64_t)s0 * f;
    const int i = threadIdx.x;
    __shared__ double sl[2][NB];
    double a[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) a[k] = (k <= i && i < jb) ? F[i + (int64_t)k * f] : 0.0;
    bool bad = false;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (j < jb) {
            const double dj = lane_value(a[j], j);
            bad = bad || !(dj > 0.0) || !(dj < INFINITY);
            const double aij = a[j];
            const double l = aij / dj;
            if (i > j) a[j] = l;
            double* lj = sl[j & 1];
            lj[i] = l;
#pragma unroll
            for (int k = j + 1; k < NB; ++k) a[k] -= aij * lj[k];
        }
    }
    if (bad && i == 0) atomicMin(err, d.node);
    if (i < jb) {
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (k <= i) F[i + (int64_t)k * f] = a[k];
    }
}

// Rows below the diagonal block of panel s0, one row per thread: the row's
// 64 entries in registers, w L11^T = a solved as L11 w^T = a^T against L11
// packed in LDS in solve order (packed_solve), then L = w D^-1 and W = w.
// Work item (x = front index, y = chunk of 256 rows).
__global__ void __launch_bounds__(256) ldlt_rows_kernel(const FrontDesc* __restrict__ descs,
                                                        const int2* __restrict__ work, int s0) {
    static_assert(NB == PS_N, "the panel solve is 64 wide");
    __shared__ __align__(16) double sL[PS_NP];
    __shared__ double sdinv[NB];
    const int2 wk = work[blockIdx.x];
    const FrontDesc d = descs[wk.x];
    const int64_t f = d.f;
    const int jb = min(NB, d.n - s0);
    const int t = threadIdx.x;
    for (int e = t; e < NB * NB; e += 256) {
        const int r = e & 63, c = e >> 6;
        if (r <= c) continue;
        sL[packed_pos<false>(r, c)] = r < jb ? d.F[(s0 + r) + (int64_t)(s0 + c) * f] : 0.0;
    }
    if (t < NB) sdinv[t] = t < jb ? 1.0 / d.F[(s0 + t) + (int64_t)(s0 + t) * f] : 0.0;
    __syncthreads();
    const int64_t rb = s0 + jb;
    const int64_t row0 = rb + (int64_t)wk.y * 256;   // the chunk's first row
    // The chunk's rows of F's panel columns and of W through buffer
    // resources: lane t at byte offset 8 t, column k at the scalar offset
    // 8 k f.  Rows past f and columns past jb get offsets beyond the range
    // (loads read zero, stores are dropped), so no lane branches.
    // (a front of order f holds f^2 doubles, so f stays far below the
    // 2 million at which the range would reach the cap)
    constexpr int kOut = 0x40000000;   // past every range below; two of them stay below 2^32
    const int64_t span = (63 * f + 256) * 8;
    const int nrec = (int)min(span, (int64_t)kOut - 8);
    const __amdgpu_buffer_rsrc_t rF =
        __builtin_amdgcn_make_buffer_rsrc(d.F + row0 + (int64_t)s0 * f, (short)0, nrec, 0x00020000);
    const __amdgpu_buffer_rsrc_t rW =
        __builtin_amdgcn_make_buffer_rsrc(d.W + (row0 - rb), (short)0, nrec, 0x00020000);
    const int vo = row0 + t < f ? 8 * t : kOut;
    const int cstep = 8 * (int)f;
    double a[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k)
        a[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                              rF, vo, k < jb ? k * cstep : kOut, 0));
    packed_solve<false>(a, sL, jb);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int so = k < jb ? k * cstep : kOut;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ps_u32x2, a[k]), rW, vo, so, 0);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ps_u32x2, a[k] * sdinv[k]), rF, vo, so, 0);
    }
}

__global__ void __launch_bounds__(256) extadd_kernel(const ExtAddDesc* __restrict__ descs,
                                                     const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    const ExtAddDesc d = descs[w.x];
    const int t = threadIdx.x;
    const int s = w.y * 64 + (t & 63);
    if (s >= d.rows) return;
    const int64_t ds = (d.rowmap ? d.rowmap[s] : s) - d.rowoff;
    for (int q = 0; q < 16; ++q) {
        const int c = w.z * 64 + (t >> 6) + 4 * q;
        if (c >= d.cols) break;
        if (d.symmetric) {
            if (c > s) continue;
            const int64_t dc = (d.rowmap ? d.rowmap[c] : c) - d.rowoff;
            d.dst[ds + dc * d.ldd] += d.src[s + (int64_t)c * d.lds];
        } else {
            d.dst[ds + (int64_t)(d.coloff + c) * d.ldd] += d.src[s + (int64_t)c * d.lds];
        }
    }
}

// dst[tgt[e]] (+)= scale * vals[src[e]]
__global__ void scatter_kernel(const int64_t* __restrict__ src, const int64_t* __restrict__ tgt,
                               const double* __restrict__ vals, double* __restrict__ dst,
                               int64_t count, double scale, int accumulate) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const double v = scale * vals[src[e]];
    if (accumulate) dst[tgt[e]] += v;
    else dst[tgt[e]] = v;
}

__global__ void set_i32_kernel(int* __restrict__ p, int v) {
    if (threadIdx.x == 0) *p = v;
}

__global__ void fill_kernel(double* __restrict__ p, int64_t count, double v) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * blockDim.x)
        p[e] = v;
}

// dst(i,j) = alpha * rs[i] * cs[j] * S(i,j), S = src, src^T or the symmetric
// completion of src's lower triangle.
__global__ void __launch_bounds__(256) copy2d_kernel(const Copy2dDesc* __restrict__ descs,
                                                     const int4* __restrict__ work) {
    __shared__ double tileT[64 * 65];
    const int4 w = work[blockIdx.x];
    const Copy2dDesc d = descs[w.x];
    const int t = threadIdx.x;
    if (d.mode == 1) {
        // transpose through LDS: read along src's contiguous dimension, write
        // along dst's; all sixteen loads of a thread are issued before the
        // first LDS store, and all LDS reads before the first global store
        const int i0 = w.y * 64, j0 = w.z * 64;
        const int jj = t & 63, j = j0 + jj;
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = i0 + (t >> 6) + 4 * q;   // src(j, i)
            v[q] = (i < d.rows && j < d.cols) ? d.src[j + (int64_t)i * d.lds] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) tileT[((t >> 6) + 4 * q) * 65 + jj] = v[q];
        __syncthreads();
        const int i = i0 + (t & 63);
        if (i >= d.rows) return;
        const double sr = d.alpha * (d.rs ? d.rs[i] : 1.0);
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = tileT[(t & 63) * 65 + (t >> 6) + 4 * q];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int jc = j0 + (t >> 6) + 4 * q;
            if (jc < d.cols) d.dst[i + (int64_t)jc * d.ldd] = sr * (d.cs ? d.cs[jc] : 1.0) * v[q];
        }
        return;
    }
    const int i = w.y * 64 + (t & 63);
    if (i >= d.rows) return;
    const int jb = w.z * 64 + (t >> 6);
    if (d.mode == 3) {   // fill; pad = 1 keeps to the lower triangle
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = jb + 4 * q;
            if (j < d.cols && (!d.pad || i >= j)) d.dst[i + (int64_t)j * d.ldd] = d.alpha;
        }
        return;
    }
    // all sixteen loads of the thread go out before its first store (src and
    // dst may be one block: an element is then read and written by the same
    // thread, or, for the symmetric completion, rewritten with its own value)
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int j = jb + 4 * q;
        v[q] = 0.0;
        if (j < d.cols) {
            if (d.mode == 4)
                v[q] = i > j ? d.src[i + (int64_t)j * d.lds] : (i == j ? 1.0 : 0.0);
            else
                v[q] = (d.mode == 2 && i < j) ? d.src[j + (int64_t)i * d.lds]
                                              : d.src[i + (int64_t)j * d.lds];
        }
    }
    const double sr = d.rs ? d.alpha * d.rs[i] : d.alpha;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int j = jb + 4 * q;
        if (j < d.cols) {
            double sc = sr;
            if (d.cs) sc *= d.cs[j];
            d.dst[i + (int64_t)j * d.ldd] = sc * v[q];
        }
    }
}

// dst(r, c) = src(rows[r], c)   (gather rows of a column-major block)
__global__ void gather_rows_kernel(const GatherDesc* __restrict__ descs, const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    const GatherDesc d = descs[w.x];
    const int t = threadIdx.x;
    const int r = w.y * 64 + (t & 63);
    if (r >= d.rows) return;
    const int64_t sr = d.rows_idx[r];
    const int cb = w.z * 64 + (t >> 6);
    const int64_t rs = d.scatter ? r : sr, rd = d.scatter ? sr : r;   // source / target row
    double v[16];   // loads first, then stores
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = cb + 4 * q;
        v[q] = c < d.cols ? d.src[rs + (int64_t)c * d.lds] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = cb + 4 * q;
        if (c < d.cols) d.dst[rd + (int64_t)c * d.ldd] = v[q];
    }
}

// dst(r, c) = src(ri[r], ci[c]) (identity where a list is null); work item
// (descriptor, row block, column block) of 64 x 64
__global__ void __launch_bounds__(256) gather2d_kernel(const Gather2dDesc* __restrict__ descs,
                                                       const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    const Gather2dDesc d = descs[w.x];
    const int t = threadIdx.x;
    const int r = w.y * 64 + (t & 63);
    if (r >= d.rows) return;
    const int64_t sr = d.ri ? d.ri[r] : r;
    const int cb = w.z * 64 + (t >> 6);
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = cb + 4 * q;
        const int64_t sc = c < d.cols ? (d.ci ? d.ci[c] : c) : 0;
        v[q] = c < d.cols ? d.src[sr + sc * d.lds] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = cb + 4 * q;
        if (c < d.cols) d.dst[r + (int64_t)c * d.ldd] = v[q];
    }
}

__global__ void vecop_kernel(const VecDesc* __restrict__ descs) {
    const VecDesc d = descs[blockIdx.x];
    for (int i = threadIdx.x; i < d.n; i += blockDim.x) {
        switch (d.op) {
            case 0: d.dst[i] = 1.0 / sqrt(d.src[i + (int64_t)i * d.ld]); break;
            case 1: {
                const double r = 1.0 / d.src[i];
                d.dst[i] = r;
                d.dst2[i] = sqrt(r);
                break;
            }
            case 2: d.dst[i + (int64_t)i * d.ld] = 1.0; break;
            case 3: d.dst[i] = 1.0 / sqrt(d.src[i]); break;
            case 5: d.dst[i] = 1.0 / d.src[i + (int64_t)i * d.ld]; break;
            case 6: d.dst[i] = sqrt(d.src[i + (int64_t)i * d.ld]); break;
            case 7: d.dst[i] = 1.0 / d.src[i] + d.a; break;
            case 8: d.dst[i + (int64_t)i * d.ld] += d.a; break;
            default: d.dst[i] = sqrt(1.0 / d.src[i]); break;
        }
    }
}

// Non-zeros of a front's leading n x n block (lower triangle stored; both
// triangles counted) and its largest |diagonal|.  Work item (x = front, y =
// block of 64 columns); a workgroup adds its count into stat[2 node] and
// raises stat[2 node + 1] to its maximum (both cleared beforehand): counts
// are integers far below 2^53, so the sum is exact in any order, and the
// maximum of non-negative doubles is the maximum of their bit patterns.
__global__ void __launch_bounds__(256) front_stats_kernel(const FrontStatDesc* __restrict__ descs,
                                                          const int2* __restrict__ work,
                                                          double* __restrict__ stat) {
    const int2 wk = work[blockIdx.x];
    const FrontStatDesc d = descs[wk.x];
    __shared__ double scnt[256];
    __shared__ double smax[256];
    double cnt = 0.0, mx = 0.0;
    const int t = threadIdx.x;
    const int c0 = wk.y * 64, c1 = min(d.n, c0 + 64);
    for (int c = c0; c < c1; ++c) {
        const double* col = d.F + (int64_t)c * d.f;
        for (int r = c + t; r < d.n; r += 256) {
            const double v = col[r];
            if (v != 0.0) cnt += r == c ? 1.0 : 2.0;
            if (r == c) mx = fmax(mx, fabs(v));
        }
    }
    scnt[t] = cnt;
    smax[t] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            scnt[t] += scnt[t + s];
            smax[t] = fmax(smax[t], smax[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        atomicAdd(stat + 2 * d.node, scnt[0]);
        atomicMax(reinterpret_cast<unsigned long long*>(stat + 2 * d.node + 1),
                  __builtin_bit_cast(unsigned long long, smax[0]));
    }
}

// Empty kernel: back-to-back launches of it measure the dispatch gap
// between kernels on a stream (profiling only).
__global__ void __launch_bounds__(64) noop_kernel() {}

}  // namespace dev

void launch_vecop(const dev::VecDesc* d, int count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(dev::vecop_kernel, dim3(count), dim3(256), 0, s, d);
}

void launch_ldlt_panel(const dev::FrontDesc* d, const int32_t* list, int count, int s0, int* err,
                       const int2* rows, int nrows, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(dev::ldlt_diag_kernel, dim3(count), dim3(64), 0, s, d, list, s0, err);
    if (nrows > 0)
        hipLaunchKernelGGL(dev::ldlt_rows_kernel, dim3(nrows), dim3(256), 0, s, d, rows, s0);
}

void launch_front_stats(const dev::FrontStatDesc* d, const int2* work, int nwork, double* stat,
                        hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::front_stats_kernel, dim3(nwork), dim3(256), 0, s, d, work, stat);
}

void launch_extadd(const dev::ExtAddDesc* d, const int4* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::extadd_kernel, dim3(nwork), dim3(256), 0, s, d, work);
}

void launch_scatter(const int64_t* src, const int64_t* tgt, const double* vals, double* dst,
                    int64_t count, double scale, bool accumulate, hipStream_t s) {
    if (count <= 0) return;
    const int64_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(dev::scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, tgt, vals,
                       dst, count, scale, accumulate ? 1 : 0);
}

void launch_noop(hipStream_t s) { hipLaunchKernelGGL(dev::noop_kernel, dim3(1), dim3(64), 0, s); }

void launch_set_i32(int* p, int v, hipStream_t s) {
    hipLaunchKernelGGL(dev::set_i32_kernel, dim3(1), dim3(64), 0, s, p, v);
}

void launch_fill(double* p, int64_t count, double v, hipStream_t s) {
    if (count <= 0) return;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(dev::fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, count, v);
}

void launch_copy2d(const dev::Copy2dDesc* d, const int4* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::copy2d_kernel, dim3(nwork), dim3(256), 0, s, d, work);
}

void launch_gather2d(const dev::Gather2dDesc* d, const int4* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::gather2d_kernel, dim3(nwork), dim3(256), 0, s, d, work);
}

void launch_gather(const dev::GatherDesc* d, const int4* work, int nwork, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::gather_rows_kernel, dim3(nwork), dim3(256), 0, s, d, work);
}

}  // namespace amls
