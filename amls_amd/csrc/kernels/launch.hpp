// Host-callable launchers for the batched HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.hpp"

namespace amls {

void launch_gemm(const dev::GemmDesc* d, const dev::TileRef* tiles, int ntiles, bool fp32,
                 hipStream_t s);
// Diagonal 64 x 64 block solves of a blocked unit-triangular solve: work item
// (descriptor, RHS column block, -, -) with pad0 = the row block; the caller
// applies the off-diagonal updates between row blocks.
// one launch per direction: the work items of a launch all solve with
// trans == `trans`
void launch_trsm_diag(const dev::TrsmDesc* d, const dev::TrsmWork* w, int nwork, bool trans,
                      hipStream_t s);
// Panel s0 of the fronts in `list` (indices into d): diagonal block, then the
// rows below it in chunks of 256 (rows = (front index, chunk) work items).
void launch_ldlt_panel(const dev::FrontDesc* d, const int32_t* list, int count, int s0, int* err,
                       const int2* rows, int nrows, hipStream_t s);
void launch_extadd(const dev::ExtAddDesc* d, const int4* work, int nwork, hipStream_t s);
// Non-zero count and largest |diagonal| of the fronts' leading blocks: work
// items (front index, 64-column block); stat (2 per node, cleared first).
void launch_front_stats(const dev::FrontStatDesc* d, const int2* work, int nwork, double* stat,
                        hipStream_t s);
// Shift-invert Lanczos steps of every descriptor's node (order n <=
// lanczos_max_order(), ncv <= lanczos_max_ncv()): init normalises the start
// vectors of nodes at step 0; one step is six launches over the device.
int lanczos_max_order();
int lanczos_max_ncv();
void launch_lanczos_init(const dev::LanczosDesc* d, int count, hipStream_t s);
void launch_lanczos_step(const dev::LanczosDesc* d, int count, int nmax, int jmax, hipStream_t s);
void launch_scatter(const int64_t* src, const int64_t* tgt, const double* vals, double* dst,
                    int64_t count, double scale, bool accumulate, hipStream_t s);
void launch_fill(double* p, int64_t count, double v, hipStream_t s);
// out(r, c) += src(r, c) over rows x cols (column-major, leading dims ldo, lds)
void launch_add2d(double* out, int64_t ldo, const double* src, int64_t lds, int64_t rows,
                  int64_t cols, hipStream_t s);
void launch_set_i32(int* p, int v, hipStream_t s);   // *p = v, as a kernel (graph-safe)
void launch_noop(hipStream_t s);   // empty kernel (dispatch-gap calibration)
void launch_vecop(const dev::VecDesc* d, int count, hipStream_t s);
void launch_copy2d(const dev::Copy2dDesc* d, const int4* work, int nwork, hipStream_t s);
void launch_gather(const dev::GatherDesc* d, const int4* work, int nwork, hipStream_t s);
// work items (descriptor, 64-row block, 64-column block)
void launch_gather2d(const dev::Gather2dDesc* d, const int4* work, int nwork, hipStream_t s);
// x[0 .. n) = reproducible uniform values in [-1, 1) from seed
void launch_fill_hash(double* x, int64_t n, uint64_t seed, hipStream_t s);
// res[j] = || R(:, j) - lam[j] V(:, j) || for j < m (q x m blocks, ld q)
void launch_ritz_resid(const double* R, const double* V, const double* lam, int64_t q, int m,
                       double* res, hipStream_t s);
// Step i of the tiled tridiagonalisation; sv holds 2 soff-strided vectors
// of n per matrix and sp 3 (soff[m] = element offset / stride unit n_m).
void launch_sytrd_step(const dev::EigDesc* d, const int4* work, int nwork, int i, double* sv,
                       double* sp, const int64_t* soff, hipStream_t s);
// Step i of one matrix (descriptor by value; sv 2n, sp 3n doubles).
void launch_sytrd_step1(const dev::EigDesc& d, int i, double* sv, double* sp, hipStream_t s);
// Steps i0 .. n-1 of every matrix in one workgroup each (n - i0 must not
// exceed sytrd_tail_order()); matrices with n <= i0 are skipped.
int sytrd_tail_order();
void launch_sytrd_tail(const dev::EigDesc* d, int nmat, int i0, double* sv, double* sp,
                       const int64_t* soff, hipStream_t s);
// counts[m] = number of eigenvalues of matrix m >= thr[m], capped at kcap
// (nmax: the largest n)
void launch_eig_count(const dev::EigDesc* d, int nmat, const double* thr, int32_t* counts,
                      int nmax, hipStream_t s);
// work: (matrix, j) padded per matrix to multiples of 4 with j = -1
void launch_bisect(const dev::EigDesc* d, const int2* work, int nwork, int nmax, hipStream_t s);
// work: (matrix, j) padded per matrix to multiples of 64 with j = -1
void launch_stein(const dev::EigDesc* d, const int2* work, int nwork, const int64_t* woff,
                  int nmax, hipStream_t s);
// work: (matrix, first column, column count <= 8); ldsDoubles >= n * count
void launch_ormtr(const dev::EigDesc* d, const int4* work, int nwork, int ldsDoubles, hipStream_t s);
// Wave-per-column-pair variant for matrices of order <= 2048 (ormtr_wave_fits):
// work items (matrix, first column, columns <= 2); overwrites the entries of
// A above each reflector with the reflector's own zeros and unit.
bool ormtr_wave_fits(int nmax);
void launch_ormtr_wave(const dev::EigDesc* d, int nmat, const int4* work, int nwork, int nmax,
                       hipStream_t s);
// y(:, c) = A x(:, c) for CSR A (n x n), column-major x, y with leading dimension n
void launch_spmm(int64_t n, const int64_t* rowptr, const int64_t* colidx, const double* vals,
                 const double* x, double* y, int ncols, int64_t ld, hipStream_t s);
// Y = A X with X, Y row-major (n x ncols, leading dimension ncols)
void launch_spmm_rm(int64_t n, const int64_t* rowptr, const int64_t* colidx, const double* vals,
                    const double* x, double* y, int ncols, hipStream_t s);
// Y = A X with X row-major (3 nb x ncols) and A in 3 x 3 blocks grouped
// by four block rows (merged column lists): qptr (nq + 1) into ucol
// (3 c | row mask << 28), qvptr (nq + 1) into val (9 doubles per block,
// row-major).  Y row-major (ldy = 0) or column-major with leading
// dimension ldy.
void launch_spmm_bsr3q(int64_t nq, int64_t nb, const int64_t* qptr, const uint32_t* ucol,
                       const int64_t* qvptr, const double* val, const double* x, double* y, int ncols,
                       hipStream_t s, int64_t ldy = 0);
// out[c] = sum_r x(r, c) * y(r, c)
void launch_coldot(int64_t n, const double* x, const double* y, int ncols, int64_t ld, double* out,
                   hipStream_t s);
// M-normalisation per problem segment: nrm(s, c) = x_s(:, c)^T y_s(:, c) over
// rows [seg[s], seg[s+1]), then x_s(:, c) /= sqrt(nrm(s, c)) (and y_s(:, c)
// too if scaleY); nrm is nseg x ncols
void launch_colnorm_seg(const int64_t* seg, int nseg, double* x, double* y, int ncols,
                        int64_t ld, double* nrm, hipStream_t s, bool scaleY = false);
// out[s ncols + c] = x_s(:, c)^T y_s(:, c) over rows [seg[s], seg[s+1])
void launch_coldot_seg(const int64_t* seg, int nseg, const double* x, const double* y, int ncols,
                       int64_t ld, double* out, hipStream_t s);
// x(r, c) *= 1 / sqrt(max(nrm[c], tiny))
void launch_colscale_rsqrt(int64_t n, double* x, int ncols, int64_t ld, const double* nrm,
                           hipStream_t s);
// dst(perm[j], c) = src(j, c)  (un-permute rows)
void launch_permute_rows(int64_t n, const int64_t* perm, const double* src, double* dst, int ncols,
                         int64_t ld, hipStream_t s);

// Element sensitivities of eigenvalues: gs(e, j) = x_je^T K0 x_je and
// gr(e, j) = -theta_j x_je^T M0 x_je (column-major nelem x nev; either may be
// null); conn is nelem x npe row-major, -1 for a constrained dof;
// npe <= adjoint_max_npe().
int adjoint_max_npe();
void launch_adjoint(int64_t n, int64_t nelem, int npe, const int64_t* conn, const double* K0,
                    const double* M0, int nev, const double* theta, const double* Z, double* gs,
                    double* gr, hipStream_t s);

}  // namespace amls
