// Unit-triangular solves of 64 unknowns held in registers, one right-hand
// side per lane, against a 64 x 64 unit lower triangular block L whose
// strictly lower part sits in LDS packed in the order the solve consumes it:
//   forward  (L x = b):   pair (j, i > j) in step j uses L(i, j),
//                          p = j 63 - j (j - 1) / 2 + (i - j - 1);
//   backward (L^T x = b): pair (j, i < j) in step j (j descending) uses
//                          L(j, i), p = 2016 - j (j + 1) / 2 + i.
// Every lane reads the same element at the same time (an LDS broadcast).
// The values stream through a ring of PS_D registers, each read PS_D pairs
// before its use; scheduling barriers every 8 pairs stop the compiler from
// hoisting all reads ahead, which would need more registers than a wave has.
#pragma once

namespace amls {
namespace dev {

constexpr int PS_N = 64;                          // unknowns
constexpr int PS_NP = PS_N * (PS_N - 1) / 2;      // packed entries
constexpr int PS_D = 16;                          // ring depth
typedef unsigned int ps_u32x2 __attribute__((ext_vector_type(2)));

// packed position of L(r, c), r > c
template <bool TR>
__device__ __forceinline__ int packed_pos(int r, int c) {
    return TR ? PS_NP - r * (r + 1) / 2 + c : c * (PS_N - 1) - c * (c - 1) / 2 + (r - c - 1);
}

// x := L^-1 x (TR = false) or L^-T x (TR = true).  live: the unknowns
// [0, live) are the real ones; forward, unknown i only reads x_j with j < i,
// so the rest never reach them; backward, steps j >= live use x_j = 0.  L's
// entries in rows or columns >= live must be zero.
template <bool TR>
__device__ __forceinline__ void packed_solve(double (&x)[PS_N], const double* sL, int live) {
    double buf[PS_D];
#pragma unroll
    for (int q = 0; q < PS_D; ++q) buf[q] = sL[q];
    if (!TR) {
#pragma unroll
        for (int j = 0; j < PS_N - 1; ++j) {
            const double xj = x[j];
#pragma unroll
            for (int i = j + 1; i < PS_N; ++i) {
                const int p = j * (PS_N - 1) - j * (j - 1) / 2 + (i - j - 1);
                const double l = buf[p % PS_D];
                if (p + PS_D < PS_NP) buf[p % PS_D] = sL[p + PS_D];
                x[i] -= l * xj;
                if ((p & 7) == 7) __builtin_amdgcn_sched_barrier(0);
            }
        }
    } else {
        // L^T upper unit: once x_j is final, x_i -= L(j, i) x_j for i < j
#pragma unroll
        for (int j = PS_N - 1; j > 0; --j) {
            const double xj = j < live ? x[j] : 0.0;
#pragma unroll
            for (int i = 0; i < j; ++i) {
                const int p = PS_NP - j * (j + 1) / 2 + i;
                const double l = buf[p % PS_D];
                if (p + PS_D < PS_NP) buf[p % PS_D] = sL[p + PS_D];
                x[i] -= l * xj;
                if ((p & 7) == 7) __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
}

}  // namespace dev
}  // namespace amls
