// Device-side descriptors and helpers shared by the HIP kernels (gfx950).
//
// Every batched kernel takes an array of small descriptors that point into
// the context's device arena; the host builds the arrays once per phase and
// uploads them, so a launch carries only (descriptor pointer, work list).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace amls {
namespace dev {

// C = alpha1 * op(A) op(B) + alpha2 * op(A2) op(B2) + beta * C
// op(A) is m x k, op(B) is k x n; ta/tb = 1 means the stored matrix is the
// transpose.  lower = 1 writes only elements with row >= col (after the
// row/col offsets below are removed: the test is on tile coordinates).
struct GemmDesc {
    const double* A;
    const double* B;
    const double* A2;
    const double* B2;
    double* C;
    int64_t lda, ldb, lda2, ldb2, ldc;
    int32_t m, n, k, k2;
    int32_t ta, tb, ta2, tb2;
    double alpha1, alpha2, beta;
    int32_t lower;      // write only the lower triangle of C (gr >= gc)
    int32_t triA;       // op(A) of the first product is 1: lower, 2: upper
                        // triangular (square, m = k): its zero half is skipped
    int32_t kchunk;     // > 0: split-k, tile work item `pad` takes k in
                        // [pad kchunk, (pad + 1) kchunk) and adds its part to C
                        // atomically (C must already hold beta C)
    int32_t pad0;
    const double* rsC;  // non-null: row i of the product is scaled by rsC[i]
                        // before it is stored (beta = 0, no split-k)
};

// One 64 x 64 output tile of descriptor `desc`.
struct TileRef {
    int32_t desc, tm, tn, pad;
};

// Left triangular solve in place: op(L) X = B with L unit lower triangular
// (strictly lower part of an LDL^T factor, leading dimension ldl).
// trans = 0: L X = B (forward), trans = 1: L^T X = B (backward).
struct TrsmDesc {
    const double* L;
    double* B;
    int64_t ldl, ldb;
    int32_t n, nrhs;
    int32_t trans;
    int32_t rhsLower;   // forward solve with a lower-triangular RHS (e.g. I): rows above
                        // each 64-column block start zero and stay zero
};

struct TrsmWork {
    int32_t desc, cb, pad0, pad1;   // cb = group of four 64-column blocks of the
                                    // right-hand side, pad0 = 64-row block solved by
                                    // trsm_diag_kernel, pad1 = column blocks in play
};

// Partial LDL^T of a front: columns [0, n) of the f x f lower-stored front F
// (leading dimension f) are eliminated; the trailing (f-n) x (f-n) block
// becomes the Schur complement.  W is an f x 64 scratch panel.
struct FrontDesc {
    double* F;
    double* W;
    int64_t f;
    int32_t n, node;
};

// dst[rowmap[s] (or s), colmap[t] (or coloff + t)] += src[s, t]
struct ExtAddDesc {
    const double* src;
    double* dst;
    const int64_t* rowmap;   // null: identity
    int64_t lds, ldd;
    int32_t rows, cols;
    int32_t coloff;          // used when symmetric == 0
    int32_t symmetric;       // 1: square src, colmap = rowmap, lower part only
    int32_t rowoff;          // subtracted from every rowmap entry (dst row = rowmap[s] - rowoff)
    int32_t pad;
};

// Symmetric tridiagonal eigen-problem bookkeeping for one dense matrix.
struct EigDesc {
    double* A;        // n x n, leading dimension lda (reflectors on exit)
    double* d;        // n
    double* e;        // n  (e[i] = T(i+1, i))
    double* tau;      // n
    double* lam;      // selected eigenvalues (descending), capacity kcap
    double* Y;        // n x kcap eigenvectors (ld n)
    double* work;     // inverse-iteration scratch: 6 * n * k doubles (interleaved)
    int64_t lda;
    int32_t n, k;     // k = number selected
    int32_t node, kcap;
};

// dst(i,j) = alpha * rs[i] * cs[j] * S(i,j); mode 0: S = src, 1: S = src^T,
// 2: S = symmetric completion of src's lower triangle.
struct Copy2dDesc {
    const double* src;
    double* dst;
    const double* rs;
    const double* cs;
    int64_t lds, ldd;
    int32_t rows, cols;
    int32_t mode, pad;  // mode 0 copy, 1 transpose, 2 symmetric completion
                        // from the lower triangle, 3 fill with alpha, 4 unit
                        // lower triangle of src (strict lower part, 1 on the
                        // diagonal, 0 above: L of an LDL^T factor)
    double alpha;
};

// scatter = 0: dst(r, c) = src(rows_idx[r], c); 1: dst(rows_idx[r], c) = src(r, c)
struct GatherDesc {
    const double* src;
    double* dst;
    const int64_t* rows_idx;
    int64_t lds, ldd;
    int32_t rows, cols;
    int32_t scatter, pad;
};

// dst(r, c) = src(ri ? ri[r] : r, ci ? ci[c] : c) for r < rows, c < cols
// (int32 index lists: the stored-column lists of the reference's ColComp)
struct Gather2dDesc {
    const double* src;
    double* dst;
    const int32_t* ri;
    const int32_t* ci;
    int64_t lds, ldd;
    int32_t rows, cols;
};

// Small per-matrix vector operations (one block per descriptor):
//   op 0: dst[i] = 1 / sqrt(src[i + i*ld])       (D^{-1/2} from an LDL^T diagonal)
//   op 1: dst[i] = 1 / src[i], dst2[i] = sqrt(1 / src[i])
//   op 2: dst[i + i*ld] = 1
//   op 3: dst[i] = 1 / sqrt(src[i])
//   op 4: dst[i] = sqrt(1 / src[i])
//   op 5: dst[i] = 1 / src[i + i*ld]             (D^{-1} from an LDL^T diagonal)
//   op 6: dst[i] = sqrt(src[i + i*ld])           (D^{1/2} from an LDL^T diagonal)
//   op 7: dst[i] = 1 / src[i] + a                (shift-invert Ritz values)
//   op 8: dst[i + i*ld] += a                     (diagonal shift)
struct VecDesc {
    const double* src;
    double* dst;
    double* dst2;
    int64_t ld;
    int32_t n, op;
    double a;
};

// Shift-invert Lanczos on one dense symmetric S (n x n, full, ld n): runs
// steps state[0] .. ncv-1 (stopping early at a breakdown, beta < 1e-13,
// which sets state[1]); U holds the Lanczos vectors (n x cap, ld n, column
// j = u_j), alpha / beta the tridiagonal (beta[j] couples u_j and u_{j+1}).
// u0 (unnormalised) starts the run when state[0] = 0.
struct LanczosDesc {
    const double* S;
    double* U;
    double* alpha;
    double* beta;
    const double* u0;
    int32_t* state;   // [0] steps done, [1] breakdown flag
    int32_t n, ncv;
    double* W;        // n: the step's vector w
    double* C;        // min(n, 1536): reorthogonalisation coefficients
    double* part;     // 64: per-block partial sums of |w|^2
};

// Statistics of a front's leading n x n block before its factorisation:
// stat[2 node] = non-zero entries of the symmetric block (both triangles,
// counted from the stored lower one), stat[2 node + 1] = max |diagonal|.
struct FrontStatDesc {
    const double* F;
    int64_t f;
    int32_t n, node;
};

}  // namespace dev
}  // namespace amls
