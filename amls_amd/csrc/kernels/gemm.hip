// Batched variable-size GEMM and triangular solves on CDNA4 matrix cores.
//
// Tile: 64 x 64 outputs per 256-thread workgroup (four wave64s, each a 32 x 32
// quarter made of 2 x 2 MFMA 16x16 blocks).  Operands are staged through LDS
// in 32-deep k-chunks, the next chunk prefetched into registers.  The FP64 path uses v_mfma_f64_16x16x4f64; the mixed
// path converts the FP64 operands to FP32 while staging and runs
// v_mfma_f32_16x16x4f32 with FP32 accumulation.
//
// These kernels carry the GEMM-shaped work of the reference's ColCompOps
// (src/linalg/colcomp.hpp:124-171, OpenBLAS backend :174-423) and of its
// Eigen triangular solves (src/linalg/sparse_direct.cpp:26-31, :56-61).
#include "common.hpp"
#include "launch.hpp"
#include "packed_solve.hpp"

namespace amls {
namespace dev {

template <typename T> struct Mfma;
// Operand layout (both types): lane l supplies A(l % 16, l / 16) and
// B(l / 16, l % 16) of the 16x4 / 4x16 blocks.  Result layout differs: the
// FP64 instruction interleaves rows (row = l/16 + 4*item), the FP32 one
// stacks them (row = 4*(l/16) + item); column = l % 16 for both.  Checked on
// gfx950 with random integer operands.
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int item) { return (lane >> 4) + 4 * item; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int item) { return 4 * (lane >> 4) + item; }
};

constexpr int TM = 64;      // tile rows / cols
constexpr int KC = 32;      // k-chunk staged in LDS
constexpr int LDSP = 65;    // padded LDS row (elements): odd, so the k-major stores of
                            // transposed operands spread over the banks

template <typename T>
struct TileAcc {
    typename Mfma<T>::v4 c[2][2];
    __device__ void zero() {
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int q = 0; q < 4; ++q) c[i][j][q] = T(0);
    }
};

typedef unsigned int gemm_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t gemm_rsrc(const double* p) {
    // raw buffer (stride 0, dword 3 as for every gfx9 target); the range
    // lies past every in-tile offset and below the sentinel GEMM_OOB
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), (short)0, 0x7fffff00, 0x00020000);
}
constexpr int GEMM_OOB = 0x7fffff80;   // offset that reads as zero

// Accumulate op(A)[0:mv, kb:ke] * op(B)[kb:ke, 0:nv] (scaled by `scale`)
// into the tile accumulator.  A, B point at the tile's first row / column
// (see the element formulas below); out-of-range elements read as zero.
// The next k-chunk is fetched into registers while the matrix cores work on
// the chunk in LDS, so memory latency overlaps the MFMA issue.  Each
// chunk's elements are read through a buffer resource based at the chunk,
// at per-thread byte offsets fixed for the whole loop; an element outside
// the tile gets an offset past the resource's range, which reads as zero,
// so the loop body has no branches and the accumulators stay in the
// matrix-core registers across it (the host keeps 64 lda and 64 ldb below
// 2^22 elements).
template <typename T>
__device__ void tile_mma(TileAcc<T>& acc, const double* __restrict__ A, int64_t lda,
                         int ta, int mv, const double* __restrict__ B, int64_t ldb,
                         int tb, int nv, int kb, int ke, double scale, T* sA, T* sB) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
    if (kb >= ke) return;
    // element j of this thread's share of a chunk: (row or column, k offset),
    // coalesced along the contiguous dimension of each operand.
    //   ta = 0: r = t & 63, k = (t >> 6) + 4 j, at r + k lda
    //   ta = 1: k = t & 31, r = (t >> 5) + 8 j, at k + r lda
    //   tb = 0: k = t & 31, c = (t >> 5) + 8 j, at k + c ldb
    //   tb = 1: c = t & 63, k = (t >> 6) + 4 j, at c + k ldb
    const int la = (int)lda, lb = (int)ldb;
    const int aR0 = ta ? (t >> 5) : (t & 63), aK0 = ta ? (t & 31) : (t >> 6);
    const int aRs = ta ? 8 : 0, aKs = ta ? 0 : 4;   // per-j steps of row and k
    const int bC0 = tb ? (t & 63) : (t >> 5), bK0 = tb ? (t >> 6) : (t & 31);
    const int bCs = tb ? 0 : 8, bKs = tb ? 4 : 0;
    const int aOff0 = 8 * (ta ? aK0 + aR0 * la : aR0 + aK0 * la);
    const int aOffS = 8 * (ta ? 8 * la : 4 * la);
    const int bOff0 = 8 * (tb ? bC0 + bK0 * lb : bK0 + bC0 * lb);
    const int bOffS = 8 * (tb ? 4 * lb : 8 * lb);
    const int64_t aChunk = ta ? 1 : lda, bChunk = tb ? ldb : 1;   // element step per k
    double ra[8], rb[8];
    auto fetch = [&](int k0) {
        const __amdgpu_buffer_rsrc_t rA = gemm_rsrc(A + (int64_t)k0 * aChunk);
        const __amdgpu_buffer_rsrc_t rB = gemm_rsrc(B + (int64_t)k0 * bChunk);
        const int kl = ke - k0;   // valid k offsets in this chunk: < kl
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool okA = aR0 + aRs * j < mv && aK0 + aKs * j < kl;
            const bool okB = bC0 + bCs * j < nv && bK0 + bKs * j < kl;
            ra[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                                   rA, okA ? aOff0 + aOffS * j : GEMM_OOB, 0, 0));
            rb[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                                   rB, okB ? bOff0 + bOffS * j : GEMM_OOB, 0, 0));
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            sA[(aK0 + aKs * j) * LDSP + aR0 + aRs * j] = T(ra[j] * scale);
            sB[(bK0 + bKs * j) * LDSP + bC0 + bCs * j] = T(rb[j]);
        }
    };
    fetch(kb);
    for (int k0 = kb; k0 < ke; k0 += KC) {
        stage();
        __syncthreads();
        fetch(k0 + KC < ke ? k0 + KC : k0);   // the last round re-reads (unused)
#pragma unroll
        for (int ks = 0; ks < KC / 4; ++ks) {
            const int kr = ks * 4 + (lane >> 4);
            const T a0 = sA[kr * LDSP + wr + (lane & 15)];
            const T a1 = sA[kr * LDSP + wr + 16 + (lane & 15)];
            const T b0 = sB[kr * LDSP + wc + (lane & 15)];
            const T b1 = sB[kr * LDSP + wc + 16 + (lane & 15)];
            acc.c[0][0] = Mfma<T>::mma(a0, b0, acc.c[0][0]);
            acc.c[0][1] = Mfma<T>::mma(a0, b1, acc.c[0][1]);
            acc.c[1][0] = Mfma<T>::mma(a1, b0, acc.c[1][0]);
            acc.c[1][1] = Mfma<T>::mma(a1, b1, acc.c[1][1]);
        }
        __syncthreads();
    }
}

// Visit the tile's elements held by this lane: fn(row_in_tile, col_in_tile, value).
template <typename T, typename Fn>
__device__ __forceinline__ void tile_foreach(const TileAcc<T>& acc, Fn fn) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                fn(wr + i * 16 + Mfma<T>::row(lane, q), wc + j * 16 + (lane & 15),
                   (double)acc.c[i][j][q]);
}

template <typename T>
__global__ void __launch_bounds__(256) gemm_batched_kernel(const GemmDesc* __restrict__ descs,
                                                           const TileRef* __restrict__ tiles) {
    __shared__ T sA[KC * LDSP];
    __shared__ T sB[KC * LDSP];
    const TileRef tr = tiles[blockIdx.x];
    const GemmDesc d = descs[tr.desc];
    const int r0 = tr.tm * TM, c0 = tr.tn * TM;
    const int mv = min(TM, d.m - r0), nv = min(TM, d.n - c0);
    TileAcc<T> acc;
    acc.zero();
    if (d.k > 0 && d.A) {
        const double* A = d.ta ? d.A + (int64_t)r0 * d.lda : d.A + r0;
        const double* B = d.tb ? d.B + c0 : d.B + (int64_t)c0 * d.ldb;
        // triangular op(A): row block tm only meets k < (tm+1) 64 (lower) or
        // k >= tm 64 (upper); split-k: this work item's chunk of k
        int kb = d.triA == 2 ? r0 : 0;
        int ke = d.triA == 1 ? min(d.k, r0 + TM) : d.k;
        if (d.kchunk > 0) {
            kb = tr.pad * d.kchunk;
            ke = min(d.k, kb + d.kchunk);
        }
        tile_mma<T>(acc, A, d.lda, d.ta, mv, B, d.ldb, d.tb, nv, kb, ke, d.alpha1, sA, sB);
    }
    if (d.k2 > 0 && d.A2) {
        const double* A = d.ta2 ? d.A2 + (int64_t)r0 * d.lda2 : d.A2 + r0;
        const double* B = d.tb2 ? d.B2 + c0 : d.B2 + (int64_t)c0 * d.ldb2;
        tile_mma<T>(acc, A, d.lda2, d.ta2, mv, B, d.ldb2, d.tb2, nv, 0, d.k2, d.alpha2, sA, sB);
    }
    const double beta = d.beta;
    const int lower = d.lower;
    double* C = d.C;
    const int64_t ldc = d.ldc;
    const bool split = d.kchunk > 0;
    const double* rs = d.rsC;
    tile_foreach<T>(acc, [&](int r, int c, double v) {
        const int gr = r0 + r, gc = c0 + c;
        if (r < mv && c < nv && (!lower || gr >= gc)) {
            double* p = C + gr + (int64_t)gc * ldc;
            if (rs) v *= rs[gr];
            if (split) atomicAdd(p, v);
            else *p = beta == 0.0 ? v : v + beta * *p;
        }
    });
}

// ---------------------------------------------------------------------------
// Triangular solve, diagonal-block part of the blocked right-looking
// algorithm: one workgroup per (descriptor, four 64-column RHS blocks, 64-row
// block rb) solves L_rb,rb X = B (or L_rb,rb^T X = B) in place.  Each wave
// takes one column block, each lane one column with its 64 rows in
// registers; the diagonal block of L sits in LDS and every lane reads the
// same element at the same time (a broadcast), so after the one barrier that
// publishes L the solve needs no synchronisation.  The host plans the row
// blocks in order and applies the off-diagonal updates between them as
// GEMMs, so every step is one wide launch.
// ---------------------------------------------------------------------------
static_assert(TM == PS_N, "the diagonal-block solve is 64 wide");

template <bool TR>
__global__ void __launch_bounds__(256, TR ? 2 : 1) trsm_diag_kernel(const TrsmDesc* __restrict__ descs,
                                                           const TrsmWork* __restrict__ work) {
    // the strictly lower part of the block, packed in solve order
    __shared__ __align__(16) double sL[PS_NP];
    const TrsmWork wk = work[blockIdx.x];
    const TrsmDesc d = descs[wk.desc];
    const int r0 = wk.pad0 * TM;
    const int mv = min(TM, d.n - r0);
    const int t = threadIdx.x;
    for (int e = t; e < TM * TM; e += 256) {
        const int r = e & 63, c = e >> 6;
        if (r <= c) continue;
        const double v = (r < mv) ? d.L[r0 + r + (int64_t)(r0 + c) * d.ldl] : 0.0;
        sL[packed_pos<TR>(r, c)] = v;
    }
    // the wave's 64 columns through one buffer resource whose range ends with
    // the matrix: a lane's column at byte offset lane * ldb * 8 (idle lanes
    // at an offset that reads as zero), row r at the instruction's immediate
    // offset r * 8.  Rows past mv read whatever lies there (or zero past the
    // matrix); they are never stored, and the solves below never let them
    // reach a live row.
    const int lane = t & 63, blk = __builtin_amdgcn_readfirstlane(wk.cb * 4 + (t >> 6));
    const int col = blk * TM + lane;
    const bool act = blk < wk.pad1 && col < d.nrhs;
    const int64_t first = (int64_t)blk * TM * d.ldb + r0;
    const int64_t span = ((int64_t)(d.nrhs - 1) * d.ldb + d.n) - first;
    // (range capped below the idle offset, which stays below 2^31 with the
    // row added; the host keeps 64 ldb * 8 bytes below the cap)
    constexpr int kCap = 0x7ffff000, kIdle = 0x7ffff800;
    const int nrec = (int)min((int64_t)kCap, max((int64_t)0, span) * 8);
    const __amdgpu_buffer_rsrc_t rX =
        __builtin_amdgcn_make_buffer_rsrc(d.B + first, (short)0, nrec, 0x00020000);
    const int voff = act ? lane * (int)d.ldb * 8 : kIdle;
    double x[TM];
#pragma unroll
    for (int r = 0; r < TM; ++r)
        x[r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rX, voff + r * 8, 0, 0));
    __syncthreads();
    // (the forward solve needs about 300 registers, so it runs one wave
    // per SIMD; the backward one fits two)
    packed_solve<TR>(x, sL, mv);
    // an opaque copy of the offset, so the 64 store addresses are formed
    // here (as immediates) rather than hoisted and held across the solve
    int vs = voff;
    asm volatile("" : "+v"(vs));
    if (mv == TM) {
#pragma unroll
        for (int r = 0; r < TM; ++r)
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(gemm_u32x2, x[r]), rX, vs + r * 8, 0, 0);
    } else {
#pragma unroll
        for (int r = 0; r < TM; ++r)
            if (r < mv)
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(gemm_u32x2, x[r]), rX, vs + r * 8, 0, 0);
    }
}

}  // namespace dev

void launch_gemm(const dev::GemmDesc* d, const dev::TileRef* tiles, int ntiles, bool fp32,
                 hipStream_t s) {
    if (ntiles <= 0) return;
    if (fp32)
        hipLaunchKernelGGL(dev::gemm_batched_kernel<float>, dim3(ntiles), dim3(256), 0, s, d, tiles);
    else
        hipLaunchKernelGGL(dev::gemm_batched_kernel<double>, dim3(ntiles), dim3(256), 0, s, d, tiles);
}

void launch_trsm_diag(const dev::TrsmDesc* d, const dev::TrsmWork* w, int nwork, bool trans,
                      hipStream_t s) {
    if (nwork <= 0) return;
    if (trans)
        hipLaunchKernelGGL(dev::trsm_diag_kernel<true>, dim3(nwork), dim3(256), 0, s, d, w);
    else
        hipLaunchKernelGGL(dev::trsm_diag_kernel<false>, dim3(nwork), dim3(256), 0, s, d, w);
}

}  // namespace amls
