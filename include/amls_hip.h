/*
 * amls_hip.h - C-ABI of libamls_hip.so, the MI355X (gfx950) implementation of
 * the AMLS numeric hot path: block Gaussian elimination down a nested-
 * dissection tree, local eigenproblems, condensation, the reduced eigen-
 * problem, back-transformation and M-normalisation, plus an FP64 subspace-
 * iteration Rayleigh-Ritz stage.
 *
 * Conventions (those of the reference's C++ interfaces):
 *   - matrices are symmetric and given in CSC with int64 column pointers and
 *     row indices, both triangles stored (Eigen::SparseMatrix<double,
 *     ColMajor, int64_t>, src/core/types.hpp:11-21);
 *   - dense blocks are column-major double;
 *   - perm[new] = old (src/partition/partitioner.hpp:8-21);
 *   - every entry point returns an amls_status; the text of the last failure
 *     is available from amls_hip_last_error().  The C++ reference throws
 *     std::invalid_argument / std::runtime_error instead
 *     (src/amls/solver.cpp:107-111, 148-149, 254-256;
 *     src/linalg/sparse_direct.cpp:13-22).
 *   - a context is externally synchronised: one host thread at a time.
 */
#ifndef AMLS_HIP_H
#define AMLS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum amls_status {
    AMLS_OK = 0,
    AMLS_E_ARG = 1,      /* invalid argument (nev <= 0, empty matrix, ...) */
    AMLS_E_NOT_SPD = 2,  /* non-positive or non-finite pivot in a node LDL^T */
    AMLS_E_NOCONV = 3,   /* eigen-solve could not deliver what was asked */
    AMLS_E_HIP = 4,      /* HIP runtime error */
    AMLS_E_OOM = 5,      /* device allocation failed */
    AMLS_E_STATE = 6,    /* call out of order, or no device in this context */
    AMLS_E_INTERNAL = 7
} amls_status;

typedef struct amls_ctx amls_ctx;

/*
 * Options of AmlsSolver::solve.  The first eleven fields are AmlsOptions of
 * src/amls/solver.hpp:18-30 with the same meaning and defaults; the rest
 * select the GPU forms.
 */
typedef struct amls_options {
    int32_t nLevels;            /* 6 */
    int64_t leafSize;           /* 500 */
    double thresholdFactor;     /* 2.0  (eta, lambda_cut = eta * lambda_hat) */
    int64_t fixedKeep;          /* 0: two-pass threshold truncation */
    int64_t nev;                /* 100 */
    int64_t denseProjectedMax;  /* 5000 (diagnostic only, as in the reference) */
    double shift;               /* 0.0: K <- K + shift*M */
    int32_t twoPass;            /* 1 */
    int32_t numThreads;         /* 0 (host threads; the GPU path ignores it) */
    int32_t recoverVectors;     /* 1 */
    int32_t verbosity;          /* 1 */
    int32_t precision;          /* 0: all FP64; 1: FP32 MFMA for the interface
                                   Schur, M-congruence and projection GEMMs */
    int32_t siSteps;            /* subspace-iteration Rayleigh-Ritz steps (0 = plain AMLS) */
    int64_t siVectors;          /* subspace width q (0: max(2*nev, nev+8), capped at p) */
    int32_t silMode;            /* local solver switch, EigenpairSolver::silMode
                                   (src/amls/eigenpair_solver.hpp:28-29): 0 = shift-
                                   invert Lanczos for nodes of order >= 200 whose K-hat
                                   has < 0.15 n^2 non-zeros, 1 = always, -1 = never */
} amls_options;

/* Fills *opt with the defaults above. */
amls_status amls_options_default(amls_options* opt);

/* device < 0 creates a host-only context: partition and symbolic analysis
 * work, every GPU phase returns AMLS_E_STATE. */
amls_status amls_hip_create(int32_t device, uint32_t flags, amls_ctx** out);
amls_status amls_hip_destroy(amls_ctx* ctx);
const char* amls_hip_last_error(const amls_ctx* ctx);

/*
 * Analysis + upload.  Partitions |K|+|M| (nested dissection, nLevels,
 * leafSize; coords = optional n x 3 row-major dof coordinates that enable
 * plane separators), permutes, runs the symbolic analysis and copies the
 * values to the device.  Replaces step 1 of AmlsSolver::solve
 * (src/amls/solver.cpp:124-134: MetisPartitioner::recursiveBisection +
 * permuteSymmetric) and GaussianElimination::precompute_/symbolicFill_
 * (src/amls/gaussian_elimination.cpp:183-270).  shift: K <- K + shift*M
 * (src/amls/solver.cpp:114-122).
 */
amls_status amls_hip_set_problem(amls_ctx* ctx, int64_t n, const int64_t* Kcolptr,
                                 const int64_t* Krowidx, const double* Kval,
                                 const int64_t* Mcolptr, const int64_t* Mrowidx,
                                 const double* Mval, const double* coords, int32_t nLevels,
                                 int64_t leafSize, double shift);

/*
 * A batch of independent pencils solved together (config C4: many small
 * problems).  Problem s is given by the s-th entry of each pointer array;
 * coords may be null, or hold null entries.  Each problem is partitioned on
 * its own; the trees form one forest whose nodes of equal height share every
 * level-batched launch, and the reduced problems, the subspace iteration's
 * dense steps and the normalisation run once per problem, batched.  Rows of
 * problem s occupy [rowStart[s], rowStart[s+1]) of the concatenated ordering.
 * The reference solves one pencil per AmlsSolver::solve call; this entry has
 * no counterpart there.
 */
amls_status amls_hip_set_problem_batch(amls_ctx* ctx, int64_t count, const int64_t* n,
                                       const int64_t* const* Kcolptr,
                                       const int64_t* const* Krowidx, const double* const* Kval,
                                       const int64_t* const* Mcolptr,
                                       const int64_t* const* Mrowidx, const double* const* Mval,
                                       const double* const* coords, int32_t nLevels,
                                       int64_t leafSize, double shift);
/* count problems; rowStart (count+1) and nodeStart (count+1, post-order
 * positions of the forest) may be null. */
amls_status amls_hip_batch_info(const amls_ctx* ctx, int64_t* count, int64_t* rowStart,
                                int64_t* nodeStart);

/* Tree in post-order positions p = 0..nNodes-1 (the elimination order of
 * SeparatorTree::postOrder, src/amls/separator_tree.cpp:33-38). */
amls_status amls_hip_tree_info(const amls_ctx* ctx, int64_t* nNodes, int64_t* maxLevel,
                               int64_t* n);
/* Any output pointer may be null.  treeId = preorder id of the node at p,
 * offset/size = its range in the permuted ordering, parent = parent position
 * (-1 at the root), level = 0 for leaves, boundary sizes b_p, perm[new] = old. */
amls_status amls_hip_tree_get(const amls_ctx* ctx, int64_t* treeId, int64_t* offset,
                              int64_t* size, int64_t* parent, int64_t* level,
                              int64_t* bsize, int64_t* perm);
/* Bisection method of the following set_problem(_batch) calls: 0 (default)
 * plane separators where coordinates allow, else for graphs of average
 * degree <= 32 a spectral cut (Fiedler vector by Lanczos, median split,
 * smaller cut boundary) with Fiduccia-Mattheyses vertex-separator
 * refinement, else BFS level sets from a pseudo-peripheral vertex; graph
 * separators drop the vertices adjacent to one side only; 1 BFS level sets
 * only, unrefined.  Stands in for MetisPartitioner's
 * METIS_ComputeVertexSeparator (src/partition/metis_partitioner.cpp:80-103),
 * keeping its recursion and stop rules (:110-179). */
amls_status amls_hip_set_partition(amls_ctx* ctx, int32_t method);
/* Host set-up time of the last set_problem(_batch), seconds, t[4]:
 * t[0] canonical columns, graph of |K| + |M|, nested dissection and symbolic
 * tables; t[1] device tables (assembly maps, permuted CSR, 3 x 3 block
 * layouts); t[2] upload to the device; t[3] total.  Host-only contexts do
 * t[0] only.  The breakdown of the reference's "partition" timing key
 * (src/amls/solver.cpp:124-134), which covers only t[0]'s partition step. */
amls_status amls_hip_setup_timing(const amls_ctx* ctx, double* t);

/*
 * Block Gaussian elimination of every tree level (leaves first): per node a
 * partial LDL^T of its front, X = Khat^-1 K_iB, the K Schur complement and
 * the M congruence update, passed to the parent by extend-add.  Replaces
 * GaussianElimination::run/runLevel/runDag (src/amls/gaussian_elimination.cpp
 * :458-866).  precision: 0 = FP64, 1 = mixed (other values: AMLS_E_ARG).
 * The first call for a problem and precision launches eagerly, the second
 * captures the launch sequence as a HIP graph and later calls replay it (K
 * and M values are read from device memory at every call); with profiling
 * on, launches are eager.
 */
amls_status amls_hip_ge(amls_ctx* ctx, int32_t precision);

/* Standard-form local pencils C_i = (L sqrt(D))^-1 Mhat (L sqrt(D))^-T and
 * their tridiagonal reduction, for all nodes (after amls_hip_ge); for the
 * nodes the local solver switch sends to shift-invert Lanczos, also
 * S_i = R^T (Khat - sigma Mhat)^-1 R with Mhat = R R^T and
 * sigma = -max(1e-10, 1e-9 max |Khat_jj|) (sil::ShiftInvertLanczos,
 * src/amls/shift_invert_lanczos.hpp:32-41, 128-135, 175-178). */
amls_status amls_hip_spectral_prepare(amls_ctx* ctx);

/* The local solver switch for the following amls_hip_spectral_prepare /
 * amls_hip_spectral calls (mode as amls_options.silMode; amls_hip_solve and
 * amls_hip_solve_numeric set it from their options). */
amls_status amls_hip_set_sil_mode(amls_ctx* ctx, int32_t mode);
/* Per node (post-order) after amls_hip_spectral: path[p] = 0 dense solver,
 * 1 shift-invert Lanczos, 2 Lanczos tried and the dense solver used instead
 * (EigenpairSolver's fallbacks); steps[p] = Lanczos steps m of the Ritz
 * pairs (0 on the dense path).  Either pointer may be null. */
amls_status amls_hip_get_sil(amls_ctx* ctx, int32_t* path, int64_t* steps);

/*
 * Local eigenpairs and truncation for every node: kind 0 = FixedCount(value)
 * (value <= 0 keeps all), kind 1 = Threshold(lambda_cut = value).  k_out
 * (nNodes, post-order) receives k_i.  Replaces
 * SubspaceCondensation::spectralLevel + EigenpairSolver + the truncation
 * policies (src/amls/subspace_condensation.cpp:45-100,
 * src/amls/eigenpair_solver.hpp:32-63, src/amls/spectral_truncation.hpp:35-88).
 * Nodes on the shift-invert Lanczos path follow sil::solveIndex /
 * solveInterval (src/amls/shift_invert_lanczos.hpp:126-203): the same
 * ncv rules, start vector (glibc's rand() after srand(12345), as Eigen's
 * Vec::Random draws it), kept Ritz values and dense fallbacks; within such a
 * node the modes are in ascending order.
 */
amls_status amls_hip_spectral(amls_ctx* ctx, int32_t kind, double value, int64_t* k_out);

/*
 * amls_hip_spectral with pinned counts: node p (post-order) keeps exactly
 * min(k_in[p], n_p) modes, its lowest, from the dense reduction (no node
 * takes the shift-invert Lanczos path).  For parity runs: pinning the k_i of
 * another implementation's run removes the discreteness of the truncation
 * (a local eigenvalue within rounding of lambda_cut).  k_out (may be null)
 * receives the counts kept.  No counterpart in the reference, whose
 * truncation policies take one count or one cut for all nodes.
 */
amls_status amls_hip_spectral_pinned(amls_ctx* ctx, const int64_t* k_in, int64_t* k_out);

/*
 * Margin of the last threshold truncation (amls_hip_spectral kind 1, or
 * pass B of a solve): the smallest min(|w_k - cut|, |w_k+1 - cut|) / cut over
 * the nodes, w_k the node's last kept and w_k+1 its first dropped local
 * eigenvalue (Ritz value on the shift-invert Lanczos path), and the node
 * (post-order) where it occurs.  A margin near rounding flags a truncation
 * that another implementation may decide the other way.  After a fixed-count
 * or pinned selection: margin = +infinity, node = -1.
 */
amls_status amls_hip_truncation_margin(amls_ctx* ctx, double* margin, int64_t* node);

/* Projection of M onto the retained modes, all levels; returns p.
 * Replaces SubspaceCondensation::projectLevel/projectNode_/finalize
 * (src/amls/subspace_condensation.cpp:102-261). */
amls_status amls_hip_project(amls_ctx* ctx, int32_t precision, int64_t* p_out);

/* Lowest `want` eigenpairs of (diag mu, Mtilde); theta (want) to the host.
 * vectors = 0 computes eigenvalues only.  Replaces solveProjected
 * (src/amls/solver.cpp:145-161). */
amls_status amls_hip_reduced_solve(amls_ctx* ctx, int64_t want, int32_t vectors,
                                   double* theta);

/* z = L^-T S q, un-permute, M-normalise; Z (n x want, original ordering,
 * column-major) to the host if non-null.  Replaces
 * SubspaceCondensation::backtransform and the normalisation loop
 * (src/amls/subspace_condensation.cpp:263-319, src/amls/solver.cpp:302-311). */
amls_status amls_hip_backtransform(amls_ctx* ctx, double* Z);

/*
 * Subspace iteration with the AMLS block factors as the inverse
 * (Q <- K^-1 M Q through L, Khat, L^T): `steps` applications with the
 * columns scaled to unit M-norm between them, then one FP64 Rayleigh-Ritz
 * with the original K and M that keeps the lowest nev pairs.  The q start columns are
 * the lowest q Ritz vectors of the last reduced solve.  theta (nev) and
 * Z (n x nev, M-normalised) to the host if non-null.  No counterpart in the
 * reference's sources (README.md:14 claims it); algorithm AMLS-SIMb of
 * papers/Improving_eigenpairs_of_automated_multilevel_substructuring_with_
 * subspace_iterations.md:384-427.  amls_hip_solve / amls_hip_solve_numeric
 * with siSteps > 0 start from the whole AMLS subspace instead when
 * siVectors >= p (q = p, no reduced solve).
 */
amls_status amls_hip_subspace_iter(amls_ctx* ctx, int32_t steps, int64_t q, int64_t nev,
                                   double* theta, double* Z);

/*
 * Solver of the final Rayleigh-Ritz problem's nev largest eigenpairs (the
 * q x q matrix C = D^-1/2 L^-1 (Y^T M Y) L^-T D^-1/2 with Y^T K Y = L D L^T):
 * mode 0 (default) uses Chebyshev-filtered subspace iteration on a block of
 * at most 192 columns when one problem is solved, nev <= 144 and q is at
 * least four block widths, and falls back to the dense reduction when it
 * does not converge (Ritz residuals above 1e-11 of the largest eigenvalue);
 * mode 1 always takes the dense reduction.  get_rr_info: the filter
 * iterations of the last Rayleigh-Ritz (0 when the dense path gave the
 * result) and its largest relative Ritz residual.  Either pointer may be
 * null.  No counterpart in the reference (its subspace iteration is absent
 * from the sources).
 */
amls_status amls_hip_set_rr_mode(amls_ctx* ctx, int32_t mode);
amls_status amls_hip_get_rr_info(amls_ctx* ctx, int32_t* iters, double* residual);

/*
 * The whole pipeline of AmlsSolver::solve (src/amls/solver.cpp:100-319),
 * including partition (coords may be null), two-pass truncation, the reduced
 * solve, vector recovery and the optional subspace iteration.
 * theta: nev (ascending); Z: n x nev or null; p_out: reduced size;
 * timing: 8 doubles {partition, gauss, condense, reduced, backtransform,
 * subspace, total, p0}.
 */
amls_status amls_hip_solve(amls_ctx* ctx, int64_t n, const int64_t* Kcolptr,
                           const int64_t* Krowidx, const double* Kval, const int64_t* Mcolptr,
                           const int64_t* Mrowidx, const double* Mval, const double* coords,
                           const amls_options* opt, double* theta, double* Z, int64_t* p_out,
                           double* timing);

/*
 * The numeric part of amls_hip_solve on the problem of the last
 * amls_hip_set_problem (values already in device memory): everything after
 * the partition.  theta / Z / p_out / timing as in amls_hip_solve (timing[0]
 * = 0); Z may be null to keep the vectors in device memory only.  Lets a
 * caller time the solve with its inputs resident in HBM, or re-solve after
 * changing options.
 */
amls_status amls_hip_solve_numeric(amls_ctx* ctx, const amls_options* opt, double* theta,
                                   double* Z, int64_t* p_out, double* timing);

/*
 * amls_hip_set_problem_batch + amls_hip_solve_numeric in one call.  For a
 * batch of `count` problems (here and in amls_hip_solve_numeric after
 * amls_hip_set_problem_batch): theta holds count x nev values, problem-major
 * (theta[s * nev + j]); Z is (sum n_s) x nev, column-major, the rows of
 * problem s at rowStart[s]; p_out receives count reduced sizes; timing as
 * above (timing[7] = total pass-A size).  The subspace iteration of a batch
 * starts from the lowest q Ritz vectors of every problem (siVectors is capped
 * at the smallest p_s).  In a batch, amls_hip_spectral applies its value to
 * every problem, amls_hip_project returns the total p, amls_hip_reduced_solve
 * returns count x want values and amls_hip_get_reduced the block-diagonal
 * Mtilde.
 */
amls_status amls_hip_solve_batch(amls_ctx* ctx, int64_t count, const int64_t* n,
                                 const int64_t* const* Kcolptr, const int64_t* const* Krowidx,
                                 const double* const* Kval, const int64_t* const* Mcolptr,
                                 const int64_t* const* Mrowidx, const double* const* Mval,
                                 const double* const* coords, const amls_options* opt,
                                 double* theta, double* Z, int64_t* p_out, double* timing);

/*
 * Eigenvalue sensitivities with respect to per-element stiffness and density
 * scales, for pencils assembled as K = sum_e s_e P_e^T Ke0 P_e and
 * M = sum_e r_e P_e^T Me0 P_e from one reference element pair:
 *   grad_s(e, j) = d theta_j / d s_e =                    x_je^T Ke0 x_je
 *   grad_r(e, j) = d theta_j / d r_e = -(theta_j - shift) x_je^T Me0 x_je
 * for M-normalised eigenvectors x_j gathered on element e's dofs.  theta is
 * what the solver returned for the problem set on the context: with a shift
 * (K <- K + shift M) it is lambda + shift, and the shift term's own
 * dependence on r_e is included.
 * conn: nelem x npe row-major dof indices in the original ordering (-1 for a
 * constrained dof); Ke0, Me0: npe x npe column-major, symmetric; theta: nev.
 * Z: n x nev column-major, or null to use the eigenvectors of the last solve
 * still in device memory (amls_hip_solve / _solve_numeric with vectors,
 * amls_hip_backtransform or amls_hip_subspace_iter).  grad_s, grad_r:
 * nelem x nev column-major host buffers, either may be null.  npe <= 96.
 * No counterpart in the reference's sources (the adjoint is new).
 */
amls_status amls_hip_adjoint(amls_ctx* ctx, int64_t nelem, int32_t npe, const int64_t* conn,
                             const double* Ke0, const double* Me0, int64_t nev,
                             const double* theta, const double* Z, double* grad_s,
                             double* grad_r);

/*
 * Split-k tuning of the batched GEMM (no counterpart in the reference, whose
 * GEMMs are OpenBLAS calls).  amls_hip_autotune times the split choices of
 * long-k products C = A^T B on this device (output orders 256 .. 2048, full
 * and lower-triangular, k = 8192 and 32768) and installs the fastest split
 * per shape as the context's table; table (capacity int64 entries) receives
 * rows of (output tiles, k, split) and *rows their number.  Later products
 * with k >= 8192 take the split of the nearest row.  amls_hip_set_tuning
 * installs a saved table (rows = 0 restores the built-in rule); it works on
 * host-only contexts too.
 */
amls_status amls_hip_autotune(amls_ctx* ctx, int64_t* table, int64_t capacity, int64_t* rows);
amls_status amls_hip_set_tuning(amls_ctx* ctx, const int64_t* table, int64_t rows);

/* ---- measurement -------------------------------------------------------
 * enable != 0 brackets every launch with HIP events on the stream it runs
 * on (grouped by kernel class) and clears the counters; enable = 0 stops the
 * event timing and clears the counters.  Launch counts and the algorithmic
 * flops / bytes of each class (from the launch descriptors: 2mnk and
 * (mk + kn + 2mn) words per GEMM, n(n-1) r for a unit triangular solve, two
 * passes over the trailing block per tridiagonalisation step, 2 nnz r for a
 * CSR product) are always counted.  Enabling also times 512 back-to-back
 * empty kernels as the class "launch_overhead" (the dispatch gap included by
 * classes timed as one bracket around many launches: sytrd_step and
 * sytrd_batch).  enable = 2 keys each class by the solve phase it ran in
 * ("ge/gemm_f64", "si/spmm", ...; phases ge, spectral, select, project,
 * reduced, backtransform, si); enable = 3 also splits GEMM batches by their
 * largest m, n, k, descriptor count, operand form and tile count.  No
 * counterpart in the reference. */
amls_status amls_hip_profile(amls_ctx* ctx, int32_t enable);

/* Chrome trace (the reference's AMLS_TRACE, src/core/trace.hpp:82-181):
 * path != null starts recording into that file (previous spans dropped);
 * path == null writes it and stops (AMLS_E_STATE when none is in progress;
 * *events, if given, receives the number of spans).  The trace holds the
 * host phases (tid 0: set-up analysis / tables / upload, GE, CONDENSE,
 * REDUCED, BACKTRANSFORM, SUBSPACE, SOLVE) and every GPU launch with its
 * solve phase (tid 1 main stream, tid 2 second stream), on the host's
 * steady clock in microseconds; tracing brackets every launch with events.
 * A context created while AMLS_TRACE=<file> is set records from the start
 * and writes the file when it is destroyed. */
amls_status amls_hip_trace(amls_ctx* ctx, const char* path, int64_t* events);
/* Class `index` (0-based, sorted by name); AMLS_E_ARG past the last one. */
amls_status amls_hip_kernel_stats(amls_ctx* ctx, int32_t index, char* name, int32_t name_len,
                                  int64_t* launches, double* ms, double* flops, double* bytes);

/* ---- inspection (tests) ---------------------------------------------- */
/* Node p after GE: the factored K block (L strict lower, D diagonal, n_p x n_p)
 * and the final Mhat_pp (lower triangle valid). */
amls_status amls_hip_get_node_factor(amls_ctx* ctx, int64_t p, double* LD, double* Mhat);
/* Node p after amls_hip_spectral: mu (k_p) and Z (n_p x k_p). */
amls_status amls_hip_get_local(amls_ctx* ctx, int64_t p, double* mu, double* Z);
/* After amls_hip_project: Mtilde (p x p, lower triangle valid, modes in
 * post-order node order) and the mode offset of each node (nNodes). */
amls_status amls_hip_get_reduced(amls_ctx* ctx, double* Mtilde, int64_t* moff);

/* ---- operator-granular entry points (host buffers, one matrix) ------- */
/*
 * The GEMM behind every ColCompOps method (src/linalg/colcomp.hpp:124-171,
 * backend OpenBlasColCompOps :174-423): C = alpha op(Ag) op(Bg) + beta C
 * with the gathered operands Ag(i, j) = A(aRowIdx[i], aColIdx[j]) and
 * Bg(i, j) = B(bRowIdx[i], bColIdx[j]) (a null list is the identity).  A is
 * the stored lda x aCols block, B the stored ldb x bCols block; Ag is
 * (transA ? k : m) x (transA ? m : k), Bg is (transB ? n : k) x
 * (transB ? k : n).  The lists are the int32 stored-column lists of the
 * reference's ColComp (`cols`): subMul(A, B, C) on compressed blocks is
 * A.data -= B.data C(B.cols, A.cols), i.e. alpha = -1, beta = 1,
 * bRowIdx = B.cols, bColIdx = A.cols; the transposed terms (subTransMul,
 * addTransMul) gather both operands.  precision 0 / 1 as for amls_hip_gemm.
 */
amls_status amls_hip_gemm_gather(amls_ctx* ctx, int32_t precision, int32_t transA, int32_t transB,
                                 int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                                 int64_t lda, int64_t aCols, const int32_t* aRowIdx,
                                 const int32_t* aColIdx, const double* B, int64_t ldb, int64_t bCols,
                                 const int32_t* bRowIdx, const int32_t* bColIdx, double beta,
                                 double* C, int64_t ldc);

/*
 * ISparseDirectSolver::factorize / solveWithFactor
 * (src/linalg/sparse_direct.hpp:13-64, .cpp:10-63) with the factor kept on
 * the device.  A: n x n CSC (int64 indices), its lower triangle used, as
 * Eigen's LDLT and SimplicialLDLT read it.  pivoting 1: DenseLDLTSolver's
 * Eigen::LDLT diagonal pivoting, P A P^T = L D L^T with the pivot order
 * Eigen picks (the first largest remaining |A_ii| at each step: its
 * factorisation is left-looking, so the order depends on the input
 * diagonal alone); pivoting 0: no permutation.  A non-positive or
 * non-finite D entry is AMLS_E_NOT_SPD (the reference's "P10" check).
 * factor_solve: X = A^-1 B (n x nrhs); factor_get: LD (n x n, L strict
 * lower, D diagonal, of P A P^T) and perm (perm[new] = old), either may be
 * null.  factor_free releases the handle (no context needed).
 */
typedef struct amls_factor amls_factor;
amls_status amls_hip_ldlt_factor(amls_ctx* ctx, int64_t n, const int64_t* colptr,
                                 const int64_t* rowidx, const double* val, int32_t pivoting,
                                 amls_factor** out);
amls_status amls_hip_ldlt_factor_solve(amls_ctx* ctx, const amls_factor* f, int64_t nrhs,
                                       const double* B, double* X);
amls_status amls_hip_ldlt_factor_get(amls_ctx* ctx, const amls_factor* f, double* LD,
                                     int64_t* perm);
amls_status amls_hip_ldlt_factor_free(amls_factor* f);

/*
 * IDenseEigensolver (src/linalg/dense_eigensolver.hpp:10-31) on the pencil
 * K x = lambda M x, K and M n x n column-major (lower triangles used), M SPD,
 * as LAPACK ?sygvd / ?sygvx with itype 1, uplo 'L' compute it
 * (dense_eigensolver.cpp:11-133): range 0 all pairs (solveAll), 1 lambda in
 * (lo, hi] (solveInterval, RANGE 'V'), 2 the il-th to iu-th lowest, 1-based
 * (solveIndex, RANGE 'I', iu clamped to n).  *m receives the count, w (m)
 * the eigenvalues ascending, V (n x m, may be null) M-orthonormal vectors.
 * n = 0, hi <= lo or iu < il give m = 0 as in the reference; an M that is
 * not positive definite is AMLS_E_NOT_SPD.
 */
amls_status amls_hip_dense_eig(amls_ctx* ctx, int64_t n, const double* K, const double* M,
                               int32_t range, double lo, double hi, int64_t il, int64_t iu,
                               int64_t* m, double* w, double* V);

/* C = alpha op(A) op(B) + beta C (column-major; trans 0/1; precision 0/1).
 * The GEMM of OpenBlasColCompOps::dgemm (src/linalg/colcomp.hpp:355-364). */
amls_status amls_hip_gemm(amls_ctx* ctx, int32_t precision, int32_t transA, int32_t transB,
                          int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                          int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                          int64_t ldc);
/* A = L D L^T (A n x n SPD, lower used) and X = A^-1 B (n x nrhs).  LD
 * receives L (strict lower) and D (diagonal).  DenseLDLTSolver::factorize +
 * solveWithFactor (src/linalg/sparse_direct.cpp:38-63). */
amls_status amls_hip_ldlt_solve(amls_ctx* ctx, int64_t n, const double* A, int64_t nrhs,
                                const double* B, double* LD, double* X);
/* amls_hip_ldlt_solve without any device allocation, on device buffers:
 * A (n x n, lower used) is overwritten by L (strict lower) and D (diagonal),
 * B (n x nrhs) by A^-1 B; the launches and results are those of
 * amls_hip_ldlt_solve.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Its scratch and launch
 * descriptors live in one 8 MiB array per device, held until the process
 * ends; sizes that do not fit it (n up to about 4,000 fit) return
 * AMLS_E_ARG.  Calls that overlap on one device race on that array. */
amls_status amls_hip_ldlt_solve_noalloc(amls_ctx* ctx, int64_t n, double* A, int64_t nrhs,
                                        double* B);
/* Largest k eigenpairs of a symmetric A (n x n, lower used): lam descending,
 * V n x k orthonormal. */
amls_status amls_hip_syevx_top(amls_ctx* ctx, int64_t n, const double* A, int64_t k,
                               double* lam, double* V);

/*
 * Out-of-core contribution store: ContributionStore of
 * src/amls/contribution_store.hpp:23-124 (append-only, one file per target
 * block, <dir>/t<target>.bin, records "int64 rows, int64 cols, rows*cols
 * doubles column-major" in native byte order, the reference's layout).
 *
 * store_open creates dir.  device >= 0: the blocks passed to append and
 * read_sum are device pointers of that GPU; they are staged through two
 * pinned host buffers of staging_bytes each (0: 64 MiB), and read_sum adds
 * the records on the GPU, on a stream of the store's own that is ordered
 * after the null stream (work on other streams must be finished by the
 * caller).  device < 0: host pointers, no HIP call.  Every call returns when
 * its data are on disk / in the output block.
 *
 * append(target, m, rows, cols, ld): ContributionStore::append (:34-49);
 * appends to one target serialise, any number of threads may append.
 * read_sum: ContributionStore::readSum (:53-92): out (rows x cols, leading
 * dimension ld) = the records of target summed in append order (one
 * addition per element and record: the host's element-wise sum bit for
 * bit); *found = 0 and out untouched when the target has no file; a
 * truncated trailing record is ignored; a record of another shape is
 * AMLS_E_ARG; consume != 0 deletes the file afterwards.  shape: rows / cols
 * of the first complete record and the record count (-1: no file), to size
 * the output.  remove / reset: :94-103 (reset deletes every file this store
 * wrote; close does too, as the reference's destructor).  disk_bytes: the
 * payload bytes appended since the last reset (:106).  last_error: text of
 * the store's last failure (valid until its next failing call).
 */
typedef struct amls_store amls_store;
amls_status amls_hip_store_open(const char* dir, int32_t device, int64_t staging_bytes,
                                amls_store** out);
amls_status amls_hip_store_close(amls_store* st);
const char* amls_hip_store_last_error(amls_store* st);
amls_status amls_hip_store_append(amls_store* st, int64_t target, const double* m, int64_t rows,
                                  int64_t cols, int64_t ld);
amls_status amls_hip_store_read_sum(amls_store* st, int64_t target, double* out, int64_t rows,
                                    int64_t cols, int64_t ld, int32_t consume, int32_t* found);
amls_status amls_hip_store_shape(amls_store* st, int64_t target, int64_t* rows, int64_t* cols,
                                 int64_t* records);
amls_status amls_hip_store_remove(amls_store* st, int64_t target);
amls_status amls_hip_store_reset(amls_store* st);
amls_status amls_hip_store_disk_bytes(amls_store* st, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* AMLS_HIP_H */
