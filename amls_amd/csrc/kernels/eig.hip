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
#include <type_traits>

namespace amls {
namespace dev {

constexpr int RS = 64;   // rows per workgroup in a tridiagonalisation step

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes, returned to every lane: pairs and quads by quad
// permutes, rows of 16 by row shifts, then the row totals carried across
// rows by the row broadcasts, so lane 63 ends with the total.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_mov<0xB1>(v);    // quad_perm 1,0,3,2
    v += dpp_mov<0x4E>(v);    // quad_perm 2,3,0,1
    v += dpp_mov<0x114>(v);   // row_shr:4
    v += dpp_mov<0x118>(v);   // row_shr:8
    v += dpp_mov<0x142>(v);   // row_bcast:15
    v += dpp_mov<0x143>(v);   // row_bcast:31
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
    // 256-thread block reduction into 4 slots of red that no thread of the
    // block reads or writes elsewhere in the kernel (so one barrier suffices)
    v = wave_sum(v);
    const int t = threadIdx.x;
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
    __shared__ double prt[4][64], pct[4][64];
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
    // staged in LDS at once (v and p as they are; the column entry until the
    // new reflector's entry replaces it), so the loads go out with the others
    // above instead of being deferred to their use after the reductions
    if (t < 128) {
        const double vv_t = (vok && hasPrev) ? svr[vidx] : 0.0;
        const double pv_t = (vok && hasPrev) ? spr[vidx] : 0.0;
        const double av_t = vok ? Ai[vidx] : 0.0;
        double* vd = t < 64 ? vr : vc;
        double* wd = t < 64 ? wr : wc;
        double* xd = t < 64 ? xr : xc;
        vd[t & 63] = vv_t;
        wd[t & 63] = pv_t;
        xd[t & 63] = av_t;
    }

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
    const double xnorm2 = block_sum(part, red + 4);
    const double alpha = xof(a1, v1, p1);
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (xnorm2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    // global stores are issued last: a barrier makes every wave wait for the
    // stores issued before it, so none may precede the barriers below
    auto leader = [&]() {
        if (!wk.w) return;
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
    };
    if (wk.y < 0) {
        leader();
        return;
    }
    // the tile vectors keep v and p as staged (zero outside the trailing
    // range): w = tau p + alpha v is formed where the update needs it, and
    // the new reflector's entries replace the column entries in xr / xc,
    // which the update does not read, so no barrier separates the two
    if (t < 128) {
        double* xd = t < 64 ? xr : xc;
        const double vv_t = (t < 64 ? vr : vc)[t & 63];
        const double pv_t = (t < 64 ? wr : wc)[t & 63];
        const double av_t = xd[t & 63];
        xd[t & 63] = vok ? (vidx == i + 1 ? 1.0 : xof(av_t, vv_t, pv_t) * scal) : 0.0;
    }
    // the vector entries of the thread's row and its 16 columns, read into
    // registers before the loop: the LDS writes of the loop would otherwise
    // order every read behind the previous element's write
    unsigned store = 0;   // tile elements to write back (bit k)
    {
        const int rr = t & 63, cq = t >> 6;
        const double vrr = vr[rr], wrr = tauP * wr[rr] + alphaP * vrr;
        double vcc[16], wcc[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            vcc[k] = vc[cq + 4 * k];
            wcc[k] = tauP * wc[cq + 4 * k] + alphaP * vcc[k];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int cc = cq + 4 * k;
            const int r = r0 + rr, c = c0 + cc;
            double a = ta[k];
            if (r < n && c > i && r >= c && hasPrev) {
                a -= vrr * wcc[k] + wrr * vcc[k];
                store |= 1u << k;
            }
            ta[k] = a;
            tile[rr * 65 + cc] = a;
        }
    }
    __syncthreads();
    // products with x_i: thread (quarter q, lane l) sums 16 terms of row l
    // and 16 of column l (the diagonal element only in the row sum), the
    // four quarters are added in LDS, and 128 threads add the row and column
    // totals into p_i
    {
        const int q = t >> 6, l = t & 63;
        double rs = 0.0, cs = 0.0;
        // two halves of eight terms, each half's LDS reads issued together
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double tr_[8], tc_[8], xc_[8], xr_[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = 16 * q + 8 * h + u;
                tr_[u] = tile[l * 65 + k];
                tc_[u] = tile[k * 65 + l];
                xc_[u] = xc[k];
                xr_[u] = xr[k];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = 16 * q + 8 * h + u;
                rs += tr_[u] * xc_[u];
                if (r0 + k != c0 + l) cs += tc_[u] * xr_[u];
            }
        }
        prt[q][l] = rs;
        pct[q][l] = cs;
    }
    __syncthreads();
    if (t < 64) {
        const int r = r0 + t;
        const double s = prt[0][t] + prt[1][t] + prt[2][t] + prt[3][t];
        if (r > i && r < n) atomicAdd(spw + r, s);
    } else if (t < 128) {
        const int cc = t - 64;
        const int c = c0 + cc;
        const double s = pct[0][cc] + pct[1][cc] + pct[2][cc] + pct[3][cc];
        if (c > i && c < n) atomicAdd(spw + c, s);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k)
        if (store & (1u << k)) {
            const int e = t + 256 * k;
            A[(r0 + (e & 63)) + (int64_t)(c0 + (e >> 6)) * lda] = ta[k];
        }
    leader();
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

// ---------------------------------------------------------------------------
// The last SYT_TAIL steps of the reduction in one workgroup per matrix.
// Steps i0 .. n-1 (n - i0 <= SYT_TAIL) of matrix blockIdx.x: the trailing
// block S = A[i0:n, i0:n] (lower triangle, mirrored) is loaded with the
// pending update of step i0 - 1 applied and stays in registers: 512
// threads as a 32 x 16 grid, thread (tr, tc) holding the 6 x 12 block
// S[6 tr + a][12 tc + b].  Each step is the unblocked Householder reduction
// with the update applied at once: no launch per step, no atomics, two
// barriers.
//   1. the owners of column k write it to LDS;
//   2. every wave forms the reflector (norm, beta, tau) from it, wave 0
//      stores the outputs; p = S v: each thread its six row partials,
//      summed over the 16 column threads of its DPP row, lane tc = 0
//      writes them;
//   3. every wave forms p.v itself, each thread w for its rows and columns
//      and applies S -= v w^T + w v^T to its block.
// Outputs follow the tiled steps exactly: d, e, tau, the reflector v_i (1
// in row i + 1) in column i of A below the diagonal, and the reflector of
// step i0 - 1, which step i0 would store.  Entries of S in rows or columns
// at or before the current step are never read again; v and w are zero
// there, so the updates leave them as they are.
// ---------------------------------------------------------------------------
constexpr int SYT_TAIL = 192;
constexpr int SYT_BR = 6, SYT_BC = 12;   // rows, columns of S per thread

// sum over the 16 lanes of each DPP row, returned to all of them
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_mov<0xB1>(v);    // quad_perm 1,0,3,2
    v += dpp_mov<0x4E>(v);    // quad_perm 2,3,0,1
    v += dpp_mov<0x141>(v);   // row_half_mirror: the other quad of each 8
    v += dpp_mov<0x140>(v);   // row_mirror: the other 8 of the row
    return v;
}

__global__ void __launch_bounds__(512) sytrd_tail_kernel(const EigDesc* __restrict__ descs, int i0,
                                                         double* __restrict__ sv_all,
                                                         double* __restrict__ sp_all,
                                                         const int64_t* __restrict__ soff) {
    __shared__ double xs[2 * SYT_TAIL], vl[SYT_TAIL], pl[SYT_TAIL], wl[SYT_TAIL];
    __shared__ double red[8];
    const EigDesc d = descs[blockIdx.x];
    const int n = d.n;
    const int m = n - i0;
    if (m <= 0) return;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int tr = t >> 4, tc = t & 15;
    const int R0 = SYT_BR * tr, C0 = SYT_BC * tc;
    const int64_t lda = d.lda;
    double* A = d.A;
    for (int r = t; r < SYT_TAIL; r += 512) {
        pl[r] = 0.0;
        wl[r] = 0.0;
        vl[r] = 0.0;
    }
    // pending update of step i0 - 1: w = tau p + alpha v, alpha = -tau^2 (p.v) / 2
    if (i0 > 0) {
        const double* vp = sv_all + 2 * soff[blockIdx.x] + (int64_t)((i0 + 1) & 1) * n + i0;
        const double* pp = sp_all + 3 * soff[blockIdx.x] + (int64_t)((i0 + 2) % 3) * n + i0;
        const double tauP = d.tau[i0 - 1];
        double part = 0.0;
        if (t < m) {
            const double v = vp[t], pv = pp[t];
            vl[t] = v;
            pl[t] = pv;
            part = pv * v;
            A[(i0 + t) + (int64_t)(i0 - 1) * lda] = v;   // reflector of step i0 - 1
        }
        part = wave_sum(part);
        if (lane == 0) red[wv] = part;
        __syncthreads();
        double pv = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) pv += red[q];
        const double alphaP = -0.5 * tauP * tauP * pv;
        if (t < m) wl[t] = tauP * pl[t] + alphaP * vl[t];
    }
    __syncthreads();
    // the thread's block of S, with the pending update
    double S[SYT_BR][SYT_BC];
    {
        double vc_[SYT_BC], wc_[SYT_BC];
#pragma unroll
        for (int b = 0; b < SYT_BC; ++b) {
            vc_[b] = vl[C0 + b];
            wc_[b] = wl[C0 + b];
        }
#pragma unroll
        for (int a = 0; a < SYT_BR; ++a) {
            const int r = R0 + a;
            const double vr = vl[r], wr = wl[r];
#pragma unroll
            for (int b = 0; b < SYT_BC; ++b) {
                const int c = C0 + b;
                double x = 0.0;
                if (r < m && c < m) {
                    const int hi = r > c ? r : c, lo = r > c ? c : r;
                    x = A[(i0 + hi) + (int64_t)(i0 + lo) * lda] - (vr * wc_[b] + wr * vc_[b]);
                }
                S[a][b] = x;
            }
        }
    }
    // one step; b0 = k % SYT_BC is a compile-time constant of the unrolled
    // loop below, so the column owners index their registers statically.
    // Column k goes to LDS (xs, double-buffered by step parity); every wave
    // then forms the reflector from it on its own, so v needs no second
    // publication and a step has two barriers.
    auto step = [&](int k, auto b0c) {
        constexpr int b0 = decltype(b0c)::value;
        const int i = i0 + k;
        double* xk = xs + (k & 1) * SYT_TAIL;
        if (tc == k / SYT_BC) {
#pragma unroll
            for (int a = 0; a < SYT_BR; ++a) xk[R0 + a] = S[a][b0];
        }
        __syncthreads();
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < SYT_TAIL / 64; ++q) {
            const int r = lane + 64 * q;
            const double x = xk[r];
            if (r >= k + 2 && r < m) part += x * x;
        }
        const double xnorm2 = wave_sum(part);
        const double alpha = k + 1 < m ? xk[k + 1] : 0.0;
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (xnorm2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        auto vof = [&](int r) { return r == k + 1 ? 1.0 : (r > k + 1 && r < m ? xk[r] * scal : 0.0); };
        if (wv == 0) {
            if (lane == 0) {
                d.d[i] = xk[k];
                if (k + 1 < m) {
                    d.e[i] = beta;
                    d.tau[i] = tau;
                }
            }
#pragma unroll
            for (int q = 0; q < SYT_TAIL / 64; ++q) {
                const int r = lane + 64 * q;
                if (r > k && r < m) A[(i0 + r) + (int64_t)i * lda] = vof(r);
            }
        }
        if (k + 1 >= m) return;
        // rows wholly at or before k: p and the update are not needed there
        // (the wave's 24 rows are skipped together)
        const bool live = R0 + SYT_BR - 1 > k;
        // p = S v for this thread's rows: six partials, summed over the 16
        // column threads of the DPP row
        double vc_[SYT_BC];
#pragma unroll
        for (int b = 0; b < SYT_BC; ++b) vc_[b] = vof(C0 + b);
        if (live) {
            double pr[SYT_BR];
#pragma unroll
            for (int a = 0; a < SYT_BR; ++a) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < SYT_BC; ++b) acc += S[a][b] * vc_[b];
                pr[a] = row16_sum(acc);
            }
            if (tc == 0) {
#pragma unroll
                for (int a = 0; a < SYT_BR; ++a) pl[R0 + a] = pr[a];
            }
        }
        __syncthreads();
        // p.v (every wave on its own), w for the rows and columns, update
        part = 0.0;
#pragma unroll
        for (int q = 0; q < SYT_TAIL / 64; ++q) {
            const int r = lane + 64 * q;
            if (r > k && r < m) part += pl[r] * vof(r);
        }
        const double alpha2 = -0.5 * tau * tau * wave_sum(part);
        if (!live) return;
        double wc_[SYT_BC];
#pragma unroll
        for (int b = 0; b < SYT_BC; ++b) {
            const int c = C0 + b;
            wc_[b] = c > k ? tau * pl[c] + alpha2 * vc_[b] : 0.0;
        }
        // row by row, so only one row's v and w are live beside the block
#pragma unroll
        for (int a = 0; a < SYT_BR; ++a) {
            const int r = R0 + a;
            const double vr = vof(r);
            const double wr = r > k ? tau * pl[r] + alpha2 * vr : 0.0;
#pragma unroll
            for (int b = 0; b < SYT_BC; ++b) S[a][b] -= vr * wc_[b] + wr * vc_[b];
        }
    };
    for (int kb = 0; kb < m; kb += SYT_BC) {
        step(kb, std::integral_constant<int, 0>());
#define AMLS_TAIL_STEP(B) \
        if (kb + B < m) step(kb + B, std::integral_constant<int, B>());
        AMLS_TAIL_STEP(1) AMLS_TAIL_STEP(2) AMLS_TAIL_STEP(3) AMLS_TAIL_STEP(4)
        AMLS_TAIL_STEP(5) AMLS_TAIL_STEP(6) AMLS_TAIL_STEP(7) AMLS_TAIL_STEP(8)
        AMLS_TAIL_STEP(9) AMLS_TAIL_STEP(10) AMLS_TAIL_STEP(11)
#undef AMLS_TAIL_STEP
    }
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

// NC Sturm counts at once (shifts x[0..NC-1]) for the multisection, by the
// three-term recurrence of the leading principal minors of T - x I:
//   p_j = (d_j - x) p_{j-1} - e_{j-1}^2 p_{j-2},  p_{-1} = 1, p_{-2} = 0;
// the number of eigenvalues below x is the number of sign changes along
// 1, p_0, ..., p_{n-1} (counted on the sign bits, so a zero, +0 or -0,
// still yields one change across j - 1 .. j + 1 since p_{j+1} = -e^2 p_{j-1}).
// Only one FMA per row is on the dependent chain (e^2 p_{j-2} is formed a
// step earlier), against reciprocal, Newton step and update in the ratio
// form q_j = p_j / p_{j-1}.  The rows are scaled by 2^-k, 2^k > |T| (exact),
// so |d - x| < 2 and e^2 < 1, and after every chunk of eight rows the pair
// (p_{j-1}, p_j) is rescaled by a power of two (exact): within a chunk
// |p| grows at most (1 + sqrt 2)^8 ~ 1.2e3.  Backward stability is that of
// the ratio form: each row's operations perturb d_j - x and e^2 by a few ulp.
// The rows are read a chunk ahead (two chunk buffers used in turn), so the
// loads wait behind eight steps.
// Pre: dd and ee already hold d 2^-k and (e 2^-k)^2 (the staged LDS copy).
template <bool Pre, int NC>
__device__ __forceinline__ void sturm_count_n(const double* __restrict__ dd,
                                              const double* __restrict__ ee, int n,
                                              const double* x, int k, int* cnt) {
    double y[NC], a[NC], b[NC];
    unsigned s[NC];
    int c[NC];
    auto sgn = [](double v) { return (unsigned)(__builtin_bit_cast(unsigned long long, v) >> 63); };
    const double d0 = Pre ? dd[0] : __builtin_amdgcn_ldexp(dd[0], -k);
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        y[q] = __builtin_amdgcn_ldexp(x[q], -k);
        a[q] = 1.0;
        b[q] = d0 - y[q];
        s[q] = sgn(b[q]);
        c[q] = (int)s[q];
    }
    constexpr int CH = 8;
    auto fetch = [&](double* dv, double* ev, int j0) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int r = min(j0 + u, n - 1);
            if (Pre) {
                dv[u] = dd[r];
                ev[u] = ee[r - 1];
            } else {
                dv[u] = __builtin_amdgcn_ldexp(dd[r], -k);
                const double e = __builtin_amdgcn_ldexp(ee[r - 1], -k);
                ev[u] = e * e;
            }
        }
    };
    auto step = [&](double dj, double e2) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double p = fma(dj - y[q], b[q], -e2 * a[q]);
            a[q] = b[q];
            b[q] = p;
            const unsigned t = sgn(p);
            c[q] += (int)(t ^ s[q]);
            s[q] = t;
        }
    };
    auto chunk = [&](const double* dv, const double* ev) {
#pragma unroll
        for (int u = 0; u < CH; ++u) step(dv[u], ev[u]);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const int m = max(__builtin_amdgcn_frexp_exp(a[q]), __builtin_amdgcn_frexp_exp(b[q]));
            a[q] = __builtin_amdgcn_ldexp(a[q], -m);
            b[q] = __builtin_amdgcn_ldexp(b[q], -m);
        }
    };
    if (n > 1) {
        double d1[CH], e1[CH], d2[CH], e2[CH];
        fetch(d1, e1, 1);
        int j0 = 1;
        // two chunks per pass, buffers used in turn: rows j0.. in d1,
        // j0 + CH.. fetched into d2 while d1's rows are stepped
        for (; j0 + 2 * CH <= n; j0 += 2 * CH) {
            fetch(d2, e2, j0 + CH);
            chunk(d1, e1);
            fetch(d1, e1, j0 + 2 * CH);
            chunk(d2, e2);
        }
        if (j0 + CH <= n) {
            fetch(d2, e2, j0 + CH);
            chunk(d1, e1);
            j0 += CH;
#pragma unroll
            for (int u = 0; u < CH; ++u)
                if (j0 + u < n) step(d2[u], e2[u]);
        } else {
#pragma unroll
            for (int u = 0; u < CH; ++u)
                if (j0 + u < n) step(d1[u], e1[u]);
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) cnt[q] = c[q];
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
        const double x = thr[m];
        int below;
        sturm_count_n<false, 1>(dd, ee, d.n, &x, __builtin_amdgcn_frexp_exp(tn), &below);
        counts[m] = min(d.n - below, d.kcap);
    }
}

// lam[j] = (j+1)-th largest eigenvalue, j < k; one wave per work item
// (x = matrix, y = j; the host pads each matrix's items to whole workgroups
// of four with y = -1, so a workgroup stages one tridiagonal in LDS).
// Multisection: each round the 64 lanes evaluate the Sturm count at 64
// interior points of the current bracket, which shrinks the bracket 65-fold
// per round (instead of 2-fold for plain bisection).  One shift per lane is
// the fastest measured: the recurrence is bound by instruction issue, so
// two or four shifts per lane (129 or 257 sections) cost more per round
// than the rounds they save (0.33 ms against 0.37 and 0.53 ms per launch).
__global__ void __launch_bounds__(256) bisect_kernel(const EigDesc* __restrict__ descs,
                                                     const int2* __restrict__ work, int nwork,
                                                     int useLds) {
    extern __shared__ double sde[];
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const EigDesc d = descs[work[blockIdx.x * 4].x];
    const double *dd, *ee;
    stage_tridiag(d, sde, useLds, dd, ee);
    const int n = d.n;
    double lo, hi, pivmin, tn;
    gersh_wave(dd, ee, n, lane, lo, hi, pivmin, tn);
    const int k = __builtin_amdgcn_frexp_exp(tn);   // 2^(k-1) <= |T| < 2^k
    if (useLds) {
        // the staged copy becomes what the Sturm counts read: d 2^-k and
        // (e 2^-k)^2 (every wave of the block has its bounds by now)
        __syncthreads();
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            sde[r] = __builtin_amdgcn_ldexp(sde[r], -k);
            const double e = __builtin_amdgcn_ldexp(sde[n + r], -k);
            sde[n + r] = e * e;
        }
        __syncthreads();
    }
    if (wid >= nwork) return;
    const int2 wk = work[wid];
    if (wk.y < 0) return;
    const int j = wk.y;
    const int a = n - 1 - j;    // ascending index
    constexpr int NC = 1, NP = 64 * NC;   // shifts per lane, points per round
    constexpr double SEC = NP + 1.0;
    for (int round = 0; round < 24; ++round) {
        const double tol = 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin;
        if (hi - lo <= tol) break;
        const double w = hi - lo;
        double x[NC];
        int c[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) x[q] = lo + w * (double)(lane + 1 + 64 * q) / SEC;
        // separate calls, so the staged copy is read with LDS instructions
        if (useLds)
            sturm_count_n<true, NC>(sde, sde + n, n, x, k, c);
        else
            sturm_count_n<false, NC>(d.d, d.e, n, x, k, c);
        // counts grow with x: points with c <= a form a prefix of the NP
        int np = 0;
#pragma unroll
        for (int q = 0; q < NC; ++q) np += __popcll(__ballot(c[q] <= a));
        const double nlo = np > 0 ? lo + w * (double)np / SEC : lo;
        const double nhi = np < NP ? lo + w * (double)(np + 1) / SEC : hi;
        if (!(nhi < hi || nlo > lo)) break;   // no progress at this resolution
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) d.lam[j] = 0.5 * (lo + hi);
}

// Inverse iteration.  Thread per eigenvalue; a thread whose eigenvalue starts
// a cluster (gap to the previous one above ctol*|T|) computes the whole
// cluster in order, orthogonalising each vector against the earlier ones.
//
// Every recurrence below walks the rows in order with a loop-carried
// dependence and one lane carries each chain, so a step costs instruction
// issue plus memory latency:
//  - scratch of matrix m holds three arrays of 16-byte records, interleaved
//    by eigenvalue (record (r, j) of array q at byte ((q n + r) k + j) 16):
//    R0 = (u1, u2) of row r of U, R1 = (1 / pivot, multiplier),
//    R2 = (interchange flag, x); a step loads whole records, so it issues
//    few memory instructions and a wave's accesses are contiguous;
//  - scratch and Y are addressed through buffer instructions: per-lane
//    column offset plus a row offset shared by the step (the host keeps
//    48 n k below 2^31);
//  - each step's inputs are fetched D steps ahead into a register ring
//    (slot s mod D, static after unrolling), with D chosen so the
//    instructions in flight stay within the wave's 63-entry memory counter;
//  - the LU factors store the reciprocal of each pivot (hardware reciprocal
//    and two Newton steps), so the solves multiply instead of divide.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
struct d2 { double a, b; };
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ __forceinline__ d2 buf_ld2(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
}
__device__ __forceinline__ void buf_st(double x, __amdgpu_buffer_rsrc_t r, int vo, int so) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, vo, so, 0);
}
__device__ __forceinline__ void buf_st2(double x, double y, __amdgpu_buffer_rsrc_t r, int vo, int so) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, d2{x, y}), r, vo, so, 0);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int64_t bytes) {
    // raw buffer (stride 0), dword 3 as for every gfx9 target
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double fast_rcp(double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    return fma(r, fma(-b, r, 1.0), r);
}

// Run step(s, in) for s = 0 .. count-1, where in[0..NA) was filled by
// load(s, in) D steps earlier (count is uniform over the wave).  For loops
// that only load.
template <int D, int NA, typename Load, typename Step>
__device__ __forceinline__ void prefetched(int count, Load load, Step step) {
    double buf[D][NA];
#pragma unroll
    for (int u = 0; u < D; ++u)
        if (u < count) load(u, buf[u]);
    int c = 0;
    // whole chunks whose every prefetch target exists: no per-step guards
    for (; c + 2 * D <= count; c += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            step(c + u, buf[u]);
            load(c + u + D, buf[u]);
        }
    }
    // the last fewer than 2 D steps
    for (; c < count; c += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int s = c + u;
            if (s < count) {
                step(s, buf[u]);
                if (s + D < count) load(s + D, buf[u]);
            }
        }
    }
}

// Loops that load and store: step(s, in, out) computes the values that
// store(s, out) writes, s = 0 .. count-1.  Chunks of D steps alternate
// between two register buffers, and every memory instruction is issued one
// chunk of work before anything waits on it: the loads of chunk c + 1 and
// the stores of chunk c - 1 go out before chunk c is computed.  (The
// compiler's wait at the loop head drains the memory counter; with this
// order nothing then in flight is younger than a chunk of work.)  Loads
// beyond the last step re-read step count - 1 instead of branching.
template <int D, int NA, int NO, typename Load, typename Step, typename Store>
__device__ __forceinline__ void pipelined(int count, Load load, Step step, Store store) {
    double A[D][NA], B[D][NA], oA[D][NO], oB[D][NO];
    if (count <= 0) return;
    auto ld = [&](int s, double* in) { load(s < count ? s : count - 1, in); };
#pragma unroll
    for (int u = 0; u < D; ++u) ld(u, A[u]);
    int c = 0;
    for (; c + 2 * D <= count; c += 2 * D) {
#pragma unroll
        for (int u = 0; u < D; ++u) load(c + D + u, B[u]);
        if (c > 0) {
#pragma unroll
            for (int u = 0; u < D; ++u) store(c - D + u, oB[u]);
        }
#pragma unroll
        for (int u = 0; u < D; ++u) step(c + u, A[u], oA[u]);
#pragma unroll
        for (int u = 0; u < D; ++u) ld(c + 2 * D + u, A[u]);
#pragma unroll
        for (int u = 0; u < D; ++u) store(c + u, oA[u]);
#pragma unroll
        for (int u = 0; u < D; ++u) step(c + D + u, B[u], oB[u]);
    }
    if (c > 0) {
#pragma unroll
        for (int u = 0; u < D; ++u) store(c - D + u, oB[u]);
    }
    // fewer than 2 D steps left; A holds steps c .. c + D - 1
#pragma unroll
    for (int u = 0; u < D; ++u) ld(c + D + u, B[u]);
#pragma unroll
    for (int u = 0; u < D; ++u)
        if (c + u < count) {
            step(c + u, A[u], oA[u]);
            store(c + u, oA[u]);
        }
#pragma unroll
    for (int u = 0; u < D; ++u)
        if (c + D + u < count) {
            step(c + D + u, B[u], oB[u]);
            store(c + D + u, oB[u]);
        }
}

// dd / ee: the tridiagonal, in LDS (Lds) or global memory.
template <bool Lds>
__device__ __forceinline__ void stein_body(const EigDesc& d, const double* dd, const double* ee,
                                           int2 wk) {
    const int n = d.n;
    const int K = d.k;
    const int j0 = wk.y;
    double lo, hi, pivmin, tn;
    gersh(dd, ee, n, lo, hi, pivmin, tn);
    const double ctol = 1e-7 * tn;
    if (j0 > 0 && fabs(d.lam[j0 - 1] - d.lam[j0]) <= ctol) return;   // not a leader
    int j1 = j0 + 1;
    while (j1 < K && fabs(d.lam[j1 - 1] - d.lam[j1]) <= ctol) ++j1;
    const __amdgpu_buffer_rsrc_t ws = buf_rsrc(d.work, (int64_t)48 * n * K);
    const __amdgpu_buffer_rsrc_t ys = buf_rsrc(d.Y, (int64_t)8 * n * K);
    const int vo = 16 * j0;
    const int rowB = 16 * K;
    const int a1 = n * rowB, a2 = 2 * n * rowB;   // arrays R1, R2 (R0 at 0)
    auto vr = [&](int r) { return vo + r * rowB; };
    const double tiny = DBL_EPSILON * fmax(tn, DBL_MIN);
    auto guard = [&](double p) { return fabs(p) < tiny ? copysign(tiny, p == 0.0 ? 1.0 : p) : p; };
    for (int j = j0; j < j1; ++j) {
        const double lam = d.lam[j] + (j - j0) * 10.0 * DBL_EPSILON * tn;
        // LU with partial pivoting of T - lam I (the second superdiagonal of
        // the running row is zero at every step), stored for the solves
        double udk = dd[0] - lam;
        double u1k = n > 1 ? ee[0] : 0.0;
        pipelined<8, 3, 5>(
            n - 1,
            [&](int k, double* in) {
                in[0] = ee[k];
                in[1] = dd[k + 1];
                in[2] = k + 2 < n ? ee[k + 1] : 0.0;
            },
            [&](int, const double* in, double* out) {
                const double sub = in[0], dg = in[1] - lam, sp = in[2];
                const bool piv = !(fabs(udk) >= fabs(sub));
                const double rp = fast_rcp(piv ? sub : guard(udk));
                const double l = (piv ? udk : sub) * rp;
                out[0] = piv ? dg : u1k;
                out[1] = piv ? sp : 0.0;
                out[2] = rp;
                out[3] = l;
                out[4] = piv ? 1.0 : 0.0;
                const double nud = piv ? u1k - l * dg : dg - l * u1k;
                u1k = piv ? -l * sp : sp;
                udk = nud;
            },
            [&](int k, const double* out) {
                buf_st2(out[0], out[1], ws, vr(k), 0);
                buf_st2(out[2], out[3], ws, vr(k), a1);
                buf_st(out[4], ws, vr(k), a2);
            });
        const double rlast = fast_rcp(guard(udk));
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
            // forward: row interchanges and multipliers; R2 of row k + 1
            // brings x(k + 1) and the next step's interchange flag
            double xk;
            if (it == 0) {
                xk = start();
                pipelined<10, 2, 1>(
                    n - 1,
                    [&](int k, double* in) {
                        in[0] = buf_ld(ws, vr(k), a1 + 8);
                        in[1] = buf_ld(ws, vr(k), a2);
                    },
                    [&](int, const double* in, double* out) {
                        const double xn = start();
                        const bool sw = in[1] != 0.0;
                        const double a = sw ? xn : xk, b = sw ? xk : xn;
                        out[0] = a;
                        xk = b - in[0] * a;
                    },
                    [&](int k, const double* out) { buf_st(out[0], ws, vr(k), a2 + 8); });
            } else {
                const d2 r0 = buf_ld2(ws, vr(0), a2);
                double pvk = r0.a;
                xk = r0.b * sc;
                pipelined<10, 3, 1>(
                    n - 1,
                    [&](int k, double* in) {
                        in[0] = buf_ld(ws, vr(k), a1 + 8);
                        const d2 r2 = buf_ld2(ws, vr(k + 1), a2);
                        in[1] = r2.a;
                        in[2] = r2.b;
                    },
                    [&](int, const double* in, double* out) {
                        const double xn = in[2] * sc;
                        const bool sw = pvk != 0.0;
                        const double a = sw ? xn : xk, b = sw ? xk : xn;
                        out[0] = a;
                        xk = b - in[0] * a;
                        pvk = in[1];
                    },
                    [&](int k, const double* out) { buf_st(out[0], ws, vr(k), a2 + 8); });
            }
            // back substitution with U (bandwidth 3), accumulating the norm
            // and the sign of the largest entry
            double nrm = 0.0, amax = 0.0, sgn = 1.0;
            auto acc = [&](double v) {
                nrm += v * v;
                if (fabs(v) > amax) { amax = fabs(v); sgn = v < 0.0 ? -1.0 : 1.0; }
            };
            double x1 = xk * rlast;
            buf_st(x1, ws, vr(n - 1), a2 + 8);
            acc(x1);
            double x2 = 0.0;
            pipelined<8, 4, 1>(
                n - 1,
                [&](int s_, double* in) {
                    const int k = n - 2 - s_;
                    const d2 u = buf_ld2(ws, vr(k), 0);
                    in[0] = u.a;
                    in[1] = u.b;
                    in[2] = buf_ld(ws, vr(k), a1);
                    in[3] = buf_ld(ws, vr(k), a2 + 8);
                },
                [&](int, const double* in, double* out) {
                    const double v = (in[3] - in[0] * x1 - in[1] * x2) * in[2];
                    out[0] = v;
                    acc(v);
                    x2 = x1;
                    x1 = v;
                },
                [&](int s_, const double* out) { buf_st(out[0], ws, vr(n - 2 - s_), a2 + 8); });
            // orthogonalise against earlier members of the cluster (modified
            // Gram-Schmidt; the last subtraction pass recomputes the norm)
            for (int q = j0; q < j; ++q) {
                const int yo = 8 * q * n;
                double dot = 0.0;
                prefetched<16, 2>(
                    n,
                    [&](int r, double* in) {
                        in[0] = buf_ld(ws, vr(r), a2 + 8);
                        in[1] = buf_ld(ys, yo, 8 * r);
                    },
                    [&](int, const double* in) { dot += in[0] * in[1]; });
                const bool lastq = q + 1 == j;
                if (lastq) { nrm = 0.0; amax = 0.0; sgn = 1.0; }
                pipelined<10, 2, 1>(
                    n,
                    [&](int r, double* in) {
                        in[0] = buf_ld(ws, vr(r), a2 + 8);
                        in[1] = buf_ld(ys, yo, 8 * r);
                    },
                    [&](int, const double* in, double* out) {
                        const double v = in[0] - dot * in[1];
                        out[0] = v;
                        if (lastq) acc(v);
                    },
                    [&](int r, const double* out) { buf_st(out[0], ws, vr(r), a2 + 8); });
            }
            sc = sgn / sqrt(nrm);
        }
        const int yo = 8 * j * n;
        pipelined<16, 1, 1>(
            n, [&](int r, double* in) { in[0] = buf_ld(ws, vr(r), a2 + 8); },
            [&](int, const double* in, double* out) { out[0] = in[0] * sc; },
            [&](int r, const double* out) { buf_st(out[0], ys, yo, 8 * r); });
    }
}

__global__ void __launch_bounds__(64) stein_kernel(const EigDesc* __restrict__ descs,
                                                   const int2* __restrict__ work, int nwork,
                                                   int useLds) {
    // every 64-item block belongs to one matrix (the host pads per matrix);
    // its tridiagonal is staged in LDS when it fits (useLds)
    extern __shared__ double sde[];
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const EigDesc d = descs[work[blockIdx.x * blockDim.x].x];
    const int n = d.n;
    if (useLds) {
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            sde[r] = d.d[r];
            sde[n + r] = r + 1 < n ? d.e[r] : 0.0;
        }
        __syncthreads();
    }
    if (id >= nwork) return;
    const int2 wk = work[id];
    if (wk.y < 0) return;
    if (useLds) stein_body<true>(d, sde, sde + n, wk);
    else stein_body<false>(d, d.d, d.e, wk);
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

// Column i of every matrix: rows <= i := 0, row i + 1 := 1 (column n - 1:
// all zero), so the column holds the reflector vector v_i exactly.  The
// tridiagonal is kept separately, so nothing else reads these entries.
// Block (column block of 64, matrix).
__global__ void __launch_bounds__(256) reflector_prep_kernel(const EigDesc* __restrict__ descs) {
    const EigDesc d = descs[blockIdx.y];
    const int n = d.n;
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    if (c >= n) return;
    double* col = d.A + (int64_t)c * d.lda;
    const int last = min(c + 1, n - 1);
    for (int r = threadIdx.x >> 6; r <= last; r += 4) col[r] = (r == c + 1) ? 1.0 : 0.0;
}

// ---------------------------------------------------------------------------
// Y := Q Y, one wave per (matrix, pair of columns): the wave keeps its two
// columns in registers (lane l owns rows l + 64 m) and applies the
// reflectors H_{n-2}, ..., H_0 in turn.  The only cross-lane step per
// reflector is a wave reduction of two dot products (DPP, no workgroup
// barrier), and each reflector's column is fetched two reflectors ahead.
// The reflector columns must hold v_i in full (reflector_prep_kernel): 0 in
// rows <= i, 1 in row i + 1, so a column is loaded as it stands; rows at or
// beyond n read as zero (their buffer offset lies past the buffer's end).
// ---------------------------------------------------------------------------
constexpr int ORW_GROUP = 8;   // register rows per skip test (512 matrix rows)

template <int MR>
__global__ void __launch_bounds__(64) ormtr_wave_kernel(const EigDesc* __restrict__ descs,
                                                        const int4* __restrict__ work) {
    const int4 wk = work[blockIdx.x];   // (matrix, first column, columns <= 2)
    const EigDesc d = descs[wk.x];
    const int n = d.n;
    const int l = threadIdx.x;
    const int ldaB = 8 * (int)d.lda;
    const __amdgpu_buffer_rsrc_t ar = buf_rsrc(d.A, (int64_t)ldaB * n);
    int voff[MR];
    double y0[MR], y1[MR];
    const double* Y0 = d.Y + (int64_t)wk.y * n;
    const double* Y1 = Y0 + n;
    const bool two = wk.z > 1;
#pragma unroll
    for (int m = 0; m < MR; ++m) {
        const int r = l + 64 * m;
        voff[m] = r < n ? 8 * r : 0x7ffffff0;
        y0[m] = r < n ? Y0[r] : 0.0;
        y1[m] = (two && r < n) ? Y1[r] : 0.0;
    }
    auto load = [&](int i, double (&v)[MR], double& tau) {
        if (i < 0) return;
#pragma unroll
        for (int m = 0; m < MR; ++m) v[m] = buf_ld(ar, voff[m], i * ldaB);
        tau = d.tau[i];
    };
    auto step = [&](int i, const double (&v)[MR], double tau) {
        if (i < 0) return;
        double p0 = 0.0, p1 = 0.0;
#pragma unroll
        for (int g = 0; g < MR; g += ORW_GROUP) {
            if (64 * (g + ORW_GROUP) - 1 > i) {   // some row of the group lies below i
#pragma unroll
                for (int m = g; m < g + ORW_GROUP && m < MR; ++m) {
                    p0 += v[m] * y0[m];
                    p1 += v[m] * y1[m];
                }
            }
        }
        const double c0 = tau * wave_sum(p0), c1 = tau * wave_sum(p1);
#pragma unroll
        for (int g = 0; g < MR; g += ORW_GROUP) {
            if (64 * (g + ORW_GROUP) - 1 > i) {
#pragma unroll
                for (int m = g; m < g + ORW_GROUP && m < MR; ++m) {
                    y0[m] -= c0 * v[m];
                    y1[m] -= c1 * v[m];
                }
            }
        }
    };
    double va[MR], vb[MR], ta = 0.0, tb = 0.0;
    load(n - 2, va, ta);
    load(n - 3, vb, tb);
    for (int i = n - 2; i >= 0; i -= 2) {
        step(i, va, ta);
        load(i - 2, va, ta);
        step(i - 1, vb, tb);
        load(i - 3, vb, tb);
    }
    double* W0 = d.Y + (int64_t)wk.y * n;
    double* W1 = W0 + n;
#pragma unroll
    for (int m = 0; m < MR; ++m) {
        const int r = l + 64 * m;
        if (r < n) {
            W0[r] = y0[m];
            if (two) W1[r] = y1[m];
        }
    }
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

int sytrd_tail_order() { return dev::SYT_TAIL; }

void launch_sytrd_tail(const dev::EigDesc* d, int nmat, int i0, double* sv, double* sp,
                       const int64_t* soff, hipStream_t s) {
    if (nmat <= 0) return;
    hipLaunchKernelGGL(dev::sytrd_tail_kernel, dim3(nmat), dim3(512), 0, s, d, i0, sv, sp, soff);
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
    (void)woff;
    hipLaunchKernelGGL(dev::stein_kernel, dim3((nwork + 63) / 64), dim3(64), useLds ? lds : 0, s, d,
                       work, nwork, useLds);
}

void launch_ormtr(const dev::EigDesc* d, const int4* work, int nwork, int ldsDoubles, hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::ormtr_kernel, dim3(nwork), dim3(256), (size_t)ldsDoubles * 8, s, d, work);
}

bool ormtr_wave_fits(int nmax) { return nmax <= 2048; }

void launch_ormtr_wave(const dev::EigDesc* d, int nmat, const int4* work, int nwork, int nmax,
                       hipStream_t s) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(dev::reflector_prep_kernel, dim3((nmax + 63) / 64, nmat), dim3(256), 0, s, d);
    if (nmax <= 512)
        hipLaunchKernelGGL(dev::ormtr_wave_kernel<8>, dim3(nwork), dim3(64), 0, s, d, work);
    else if (nmax <= 1024)
        hipLaunchKernelGGL(dev::ormtr_wave_kernel<16>, dim3(nwork), dim3(64), 0, s, d, work);
    else if (nmax <= 1536)
        hipLaunchKernelGGL(dev::ormtr_wave_kernel<24>, dim3(nwork), dim3(64), 0, s, d, work);
    else
        hipLaunchKernelGGL(dev::ormtr_wave_kernel<32>, dim3(nwork), dim3(64), 0, s, d, work);
}

}  // namespace amls
