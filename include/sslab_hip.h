/*
 * sslab_hip.h -- C ABI of libsslab_hip.so, the MI355X (gfx950) slab direct solver.
 *
 * Plain pointers and sizes only.  All matrix pointers are DEVICE pointers to
 * row-major (C-order) FP64 data, exactly the memory order of the numpy float64
 * C-order arrays the reference passes around.  `stream` is a hipStream_t (NULL
 * = the default stream); every call is asynchronous on it unless stated.
 * Every function returns 0 on success or a non-zero status; the message of the
 * last failure in the calling thread is returned by sslab_last_error().
 *
 * Reference citations are to SimonDirckx/SslabLU.
 */
#ifndef SSLAB_HIP_H
#define SSLAB_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SSLAB_OK 0
#define SSLAB_BAD_ARGUMENT 1
#define SSLAB_HIP_ERROR 2
#define SSLAB_WORKSPACE_TOO_SMALL 3
#define SSLAB_SINGULAR 4
#define SSLAB_NOT_SUPPORTED 5

/* ABI version (bumped on any signature change). */
int sslab_abi_version(void);
/* Message of the last failed call in this thread ("" if none). */
const char* sslab_last_error(void);

/* ---- dense primitives ------------------------------------------------------
 * C_b = alpha * op(A_b) op(B_b) + beta * C_b, b < batch, X_b = X + b*strideX.
 * Replaces the numpy `@` products on block lists in
 * direct_solve/omsdirectsolve.py:89-96,125-139 and multislab/omsdirectsolve.py:55-63.
 */
int sslab_dgemm_strided_batched(int transA, int transB, int M, int N, int K, double alpha,
                                const double* A, int lda, long long strideA, const double* B,
                                int ldb, long long strideB, double beta, double* C, int ldc,
                                long long strideC, int batch, void* stream);
/* Grouped products: one launch computes every descriptor
 *   C_b = beta * Cin_b + sum_t alpha_t * op(A_t,b) op(B_t,b),   b < batch,
 * with up to 3 summed terms per descriptor (terms share M and N; K may differ; X_b =
 * X + b*strideX, stride 0 broadcasts one matrix over the batch).  Cin may equal C (in-place
 * accumulation); Cin = NULL requires beta = 0.  The descriptors of one call run in any
 * order and concurrently, so none may read another's output; more than 12 descriptors
 * are split over several launches.  Replaces chains of the numpy products of one
 * tree level in matAssembly/HBS/ULVsparse.py:158-226 and HBSnew.py:306-416 (ULV solve and
 * HBS apply) with one launch each. */
typedef struct {
  const double* A;
  const double* B;
  long long strideA, strideB;
  int lda, ldb, K, transA, transB;
  double alpha;
} sslab_gemm_term;
typedef struct {
  sslab_gemm_term term[3];
  int nterms, M, N, batch;
  const double* Cin;
  long long strideCin;
  int ldcin;
  double* C;
  long long strideC;
  int ldc;
  double beta;
} sslab_gemm_desc;
int sslab_dgemm_grouped(const sslab_gemm_desc* descs, int ndesc, void* stream);
/* FP32 version (v_mfma_f32_16x16x4_f32, FP32 accumulation; K > 0): the sketch products
 * Y = S Omega, Z = S^T Psi of the HBS compression in FP32 (matAssembly/HBS/HBSnew.py:
 * 147-170 forms them in FP64 with numpy). */
int sslab_sgemm_strided_batched(int transA, int transB, int M, int N, int K, float alpha,
                                const float* A, int lda, long long strideA, const float* B,
                                int ldb, long long strideB, float beta, float* C, int ldc,
                                long long strideC, int batch, void* stream);

/* In-place LU with partial pivoting (scipy.linalg.lu_factor semantics: ipiv is
 * 0-based, info = i+1 for the first zero pivot).  Replaces lu_factor at
 * direct_solve/omsdirectsolve.py:84,96,210,333. */
int sslab_dgetrf_batched(int n, double* A, int lda, long long strideA, int* ipiv,
                         long long strideP, int* info, int batch, void* stream);

/* X = A^{-1} B from getrf factors; B and X distinct; perm_work: batch*n ints.
 * Replaces lu_solve at direct_solve/omsdirectsolve.py:89-91,132-139,264-277. */
int sslab_dgetrs_batched(int n, int nrhs, const double* LU, int lda, long long strideA,
                         const int* ipiv, long long strideP, const double* B, int ldb,
                         long long strideB, double* X, int ldx, long long strideX,
                         int* perm_work, int batch, void* stream);

/* Explicit inverses X_b = A_b^{-1} of `batch` n x n matrices (A read only, X distinct),
 * by 2 x 2 block elimination on MFMA GEMMs down to 64 x 64 blocks that one workgroup
 * inverts by Gauss-Jordan with partial pivoting; the normwise backward error ||G - A X G||_F / (||A||_F ||X G||_F)
 * is then sampled on four +-1 columns G (one stream synchronisation) and, when it exceeds
 * tol (tol <= 0: 1e-12) or is not finite for some matrix, the whole batch is inverted
 * again by pivoted LU against the identity.  info (device, batch ints, may be NULL): that LU's info, else zeros;
 * *used_lu (host, may be NULL): 1 when the LU ran.  work / iwork: sslab_inv_sizes.
 * Replaces the explicit block inverses of the red-black levels, lu_factor(T) followed by
 * lu_solve(T, I) (direct_solve/omsdirectsolve.py:31-36, 211, 308-333), and the local
 * solves of the HPS assembly (2-D leaves, merges, pair solves; the 3-D pair factors). */
int sslab_inv_sizes(int n, int batch, long long* work_doubles, long long* work_ints);
int sslab_dinv_batched(int n, const double* A, int lda, long long strideA, double* X, int ldx,
                       long long strideX, int batch, double* work, int* iwork, int* info,
                       double tol, int* used_lu, void* stream);

/* Solve A_b^T X_b = B_b with the factors of sslab_dgetrf_batched (P A = L U, so
 * U^T L^T (P X) = B).  B is read, X (distinct from B) written; work: batch x n x nrhs
 * doubles (stride n*nrhs), perm_work: batch x n ints.  Replaces the transposed sparse
 * solves behind the S maps' rmatvec / rmatmat (multislab/oms.py:111-129:
 * solver_ii.rmatvec = A_ii^{-T}, used by the rkHBS sketches Z = S^T Psi,
 * matAssembly/matAssembler.py:103-133). */
int sslab_dgetrs_trans_batched(int n, int nrhs, const double* LU, int lda, long long strideA,
                               const int* ipiv, long long strideP, const double* B, int ldb,
                               long long strideB, double* X, int ldx, long long strideX,
                               double* work, int* perm_work, int batch, void* stream);
/* Householder QR, optionally column-pivoted (m, n <= 1024), one matrix per
 * workgroup: A_b (m x n) is overwritten by R (zeros below the diagonal); Q (m x m,
 * optional) receives the complete orthogonal factor; perm (n ints, optional)
 * switches on column pivoting and receives the order with A_in[:, perm] = Q R.
 * Replaces numpy/torch qr in matAssembly/HBS/HBSnew.py:26 and
 * ULVsparse.py:18-25,28-61,79 (the HBS basis / ULV factor steps). */
int sslab_dgeqrf_batched(int m, int n, double* A, int lda, long long strideA, double* Q,
                         int ldq, long long strideQ, int* perm, long long stridePerm,
                         int batch, void* stream);

/* In place X_b <- U_b^{-1} X_b for upper-triangular U_b (n x n).  Replaces the
 * triangular np.linalg.solve of ULVsparse.block_solve (ULVsparse.py:213-226). */
int sslab_dtrsm_upper_batched(int n, int nrhs, const double* U, int ldu, long long strideU,
                              double* X, int ldx, long long strideX, int batch, void* stream);

/* Y_b[diag block blk] += X[b*nblk + blk] (bs x bs blocks), b < nparent: the
 * block_diag_add of ULVsparse.py:158-183. */
int sslab_add_diag_blocks(int bs, int nblk, int nparent, const double* X, double* Y, int ldy,
                          long long strideY, void* stream);

/* ---- red-black block-tridiagonal solver ------------------------------------
 * System (non-cyclic): T_i x_i + SiM_i x_{i-1} + SiP_i x_{i+1} = v_i, i < nb,
 * with T, SiM, SiP each nb contiguous m x m blocks (SiM_0, SiP_{nb-1} unused).
 * Any nb >= 1 (no power-of-two restriction).
 * Replaces RedBlackSolver.factorize / .solve / .build_block_RB_solver_level
 * (direct_solve/omsdirectsolve.py:180-219, 221-279, 283-333).
 */
int sslab_rb_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
                   long long* work_ints);
/* Number of ints written to `info` by sslab_rb_factor (one per eliminated block). */
int sslab_rb_info_count(int nb);
/* T may be NULL for identity diagonal blocks (the oms S-system); the first level then
 * needs no LU. */
int sslab_rb_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
                    double* factor, double* work, int* iwork, int* info, void* stream);
/* V (nb*m x nrhs, leading dim ldv) holds the right-hand sides and is overwritten;
 * X (nb*m x nrhs, ldx) receives the solution. */
int sslab_rb_solve(int nb, int m, int nrhs, const double* factor, double* V, int ldv,
                   double* X, int ldx, void* stream);

/* sslab_rb_solve with flags; bit 0 (SSLAB_RB_IDENTITY_DIAGONAL): the factor came from
 * sslab_rb_factor with T = NULL, so the first level's T^{-1} = I and its product is
 * skipped (y = v is read from V).  No reference counterpart (an optimisation of the oms
 * S-system, whose diagonal blocks are identities, multislab/oms.py:292). */
#define SSLAB_RB_IDENTITY_DIAGONAL 1
/* bit 1 (SSLAB_RB_LEVEL0_REDUCED, requires bit 0 and nb >= 2): V and X hold the entry
 * state written by sslab_oms_rhs_rb0 -- V's even block rows the first level's reduced
 * right-hand sides v_i - SiM_i v_{i-1} - SiP_i v_{i+1}, X's odd block rows v_j, X's even
 * block rows zero -- so the first level's forward sweep and the zeroing of X are
 * skipped and its backward sweep updates X in place. */
#define SSLAB_RB_LEVEL0_REDUCED 2
/* bit 2 (SSLAB_RB_FORWARD_ONLY): run the forward sweep and the coarsest level only,
 * leaving in X the backward sweep's entry state -- y_j = T_j^{-1} v_j at every level's
 * odd block rows, x at the coarsest level's block rows -- and V overwritten.  bit 3
 * (SSLAB_RB_BACKWARD_ONLY): X holds that entry state (e.g. from sslab_oms_fold_apply);
 * run the backward sweep only (V is not read and may be NULL).  No reference
 * counterpart (the reference solve is one piece, direct_solve/omsdirectsolve.py:221-279). */
#define SSLAB_RB_FORWARD_ONLY 4
#define SSLAB_RB_BACKWARD_ONLY 8
/* bit 4 (SSLAB_RB_CHAINED_BACKWARD): run the backward sweep as one dependency-ordered
 * launch (every level's tiles in one grid; a tile starts once the blocks it reads are
 * final, tracked by per-GPU device counters) when the system has at most 128 blocks,
 * else level by level as without the flag.  Chained solves on one GPU must not overlap
 * in time (they share the counters).  No reference counterpart (a launch-structure
 * choice for the sweep of direct_solve/omsdirectsolve.py:221-279). */
#define SSLAB_RB_CHAINED_BACKWARD 16
int sslab_rb_solve_ex(int nb, int m, int nrhs, const double* factor, double* V, int ldv,
                      double* X, int ldx, int flags, void* stream);

/* The forward sweep and coarsest level of sslab_rb_solve_ex(..., SSLAB_RB_IDENTITY_DIAGONAL
 * | SSLAB_RB_FORWARD_ONLY) for nb >= 2 and a V (nb*m x ncols, ldv) whose block row i is
 * zero outside the columns [lo[i], hi[i]) (host int arrays of length nb): every product
 * touches only the columns where its operands can be nonzero.  X (zeroed here) receives
 * the same entry state; V is overwritten.  Builds the boundary-map fold's maps from the rhs
 * maps of a column-permuted identity, whose block rows read one interval of boundary
 * columns each.  No reference counterpart (the reference forms rhs per solve,
 * multislab/oms.py:180-186, and sweeps every column, direct_solve/omsdirectsolve.py:
 * 221-279). */
int sslab_rb_forward_windowed(int nb, int m, int ncols, const double* factor, double* V, int ldv,
                              double* X, int ldx, const int* lo, const int* hi, void* stream);

/* Whether a chained backward sweep on the current GPU gave up waiting since the last
 * call (*aborted = 1: a tile's bounded wait ran out and that solve's result is not
 * valid).  Synchronises the device and clears the flag.  No reference counterpart. */
int sslab_rb_chain_status(int* aborted);

/* Several independent systems (any mix of nb, m, nrhs) solved together: each
 * elimination level of all systems with nrhs <= 32 shares one grouped GEMM launch;
 * systems with more right-hand sides keep their own launches.  Arrays of length
 * nsys (host memory) give each system's sizes, factor (from sslab_rb_factor), V and X
 * exactly as for sslab_rb_solve.  No reference counterpart (the reference solves one
 * system per call, direct_solve/omsdirectsolve.py:221-279); nsys <= 64. */
int sslab_rb_solve_multi(int nsys, const int* nb, const int* m, const int* nrhs,
                         const double* const* factor, double* const* V, const int* ldv,
                         double* const* X, const int* ldx, void* stream);

/* ---- GEMM tile policy (autotuning hook) ----------------------------------------
 * No reference counterpart (the reference calls BLAS).  Sets / reads the calling
 * thread's tile and split-K selection for every later GEMM launch made by this
 * library on that thread.  p = {target_wgs, split_wgs, max_ksplit, min_split_k,
 * cfg_mid, cfg_split, max_tile}, n = 7.  Tile ids: 0 128x128, 1 128x64, 2 64x64/BK32,
 * 3 64x64/BK16, 4 64x32, 5 32x32, 6 32x64, 7 64x128 (8 32x16 and 9 32x32/BK32 are
 * chosen automatically for launches with at most 16 / 32 columns), 10 / 11 128x64 with A
 * streamed to registers, 12 32x32/BK32, 13 32x32/BK64, 14 64x32/BK32.  Results do not
 * depend on the policy beyond floating-point summation order. */
int sslab_gemm_set_policy(const int* p, int n);
/* Per-step tile policies of the red-black backward sweep for the calling thread's later
 * solves: step u (level L-1-u) runs under p[7u .. 7u+6] (same layout as above); steps past
 * nsteps use the thread's policy.  nsteps = 0 clears the table; nsteps <= 16.  No
 * reference counterpart (autotuning hook, like sslab_gemm_set_policy). */
int sslab_rb_set_backward_policies(const int* p, int nsteps);
int sslab_gemm_get_policy(int* p, int n);

/* ---- mixed precision (FP32 red-black preconditioner + FP64 refinement) -------
 * No reference counterpart: the reference is FP64 throughout.  The FP32 factor is
 * the FP64 red-black factor rounded to float (same layout, factor_doubles floats). */
int sslab_cast_f64_to_f32(long long n, const double* x, float* y, void* stream);
/* Red-black factorization computed in FP32 (FP32 LU, FP32 MFMA Schur updates) from FP32
 * copies of the blocks: the low-precision factor used as the refinement preconditioner
 * by sslab_rb_solve_mixed.  Same layout and sizes as sslab_rb_factor (sizes counted in
 * floats); T may be NULL for identity diagonal blocks. */
int sslab_rb_factor_f32(int nb, int m, const float* T, const float* SiM, const float* SiP,
                        float* factor32, float* work32, int* iwork, int* info, void* stream);
/* Red-black solve with an FP32 factor; V (float, overwritten), X (float). */
int sslab_rb_solve_f32(int nb, int m, int nrhs, const float* factor32, float* V, int ldv,
                       float* X, int ldx, void* stream);
/* R = B - S X (FP64) for S = diag(T) + sub/super-diagonal SiM / SiP blocks; T may be
 * NULL for identity diagonal blocks (the oms S-system), then B, X, R must be
 * contiguous (ld = nrhs). */
int sslab_rb_residual(int nb, int m, int nrhs, const double* T, const double* SiM,
                      const double* SiP, const double* B, int ldb, const double* X, int ldx,
                      double* R, int ldr, void* stream);
/* x = M b, then `iters` steps x += M (b - S x), M = the FP32 red-black solve.
 * B, X: nb*m x nrhs contiguous doubles; work64: nb*m*nrhs doubles; work32:
 * 2*nb*m*nrhs floats. */
int sslab_rb_solve_mixed(int nb, int m, int nrhs, const float* factor32, const double* T,
                         const double* SiM, const double* SiP, const double* B, double* X,
                         double* work64, float* work32, int iters, void* stream);
/* sslab_rb_solve_mixed that allocates nothing and takes no workspace: no counterpart in
 * the reference.  It is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The FP64
 * residual and the two FP32 panels live in a static device array of 64 MiB that every GPU
 * the library runs on holds until the process ends, so 16 * nb*m*nrhs bytes (plus up to
 * 768 bytes of alignment) must fit in 64 MiB; larger calls fail with SSLAB_BAD_ARGUMENT
 * and a message.  All calls on one GPU share that array: calls that overlap on one GPU
 * (different streams or threads) race, so serialise them. */
int sslab_rb_solve_mixed_noalloc(int nb, int m, int nrhs, const float* factor32,
                                 const double* T, const double* SiM, const double* SiP,
                                 const double* B, double* X, int iters, void* stream);

/* ---- Thomas block-tridiagonal solver ---------------------------------------
 * Same system; sequential block elimination with explicit block inverses.
 * Replaces BlockTridiagonalSolver.factorize_helper / .solve_helper
 * (direct_solve/omsdirectsolve.py:50-100, 102-141) and the identity-diagonal
 * build_block_tridiagonal_solver / block_tridiagonal_solve
 * (multislab/omsdirectsolve.py:17-99, with T_i = I).
 * factor layout: A (nb-1 blocks) | B^{-1} (nb) | LU(B) (nb) | C (nb-1);
 * iwork starts with the nb*m pivots of LU(B). */
int sslab_thomas_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
                       long long* work_ints);
int sslab_thomas_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
                        double* factor, double* work, int* iwork, int* info, void* stream);
int sslab_thomas_solve(int nb, int m, int nrhs, const double* factor, double* V, int ldv,
                       double* X, int ldx, void* stream);

/* ---- HPS slab assembly (2-D Chebyshev leaves, q x q points, p = q - 2) -----
 * Replaces the per-slab discretization + S-map assembly of
 * multislab/oms.py:160-205 (solverWrapper.construct, solver/solver.py:79-163, and
 * compute_stmaps + the dense assembler, oms.py:99-129, matAssembler.py:93-98),
 * whose HPS leaf solver (hpsalt, solver/solver.py:121-144) is not in the reference.
 *
 * Leaf DtN maps T_l = N_b - N_i A_ii^{-1} A_ib (4p x 4p; edges left, right, bottom,
 * top) for nleaves leaves.  Dx/Dxx/Dy/Dyy: q x q derivative matrices of the leaf
 * grid; Nb (4p x 4p), Ni (4p x p^2): outward-normal derivative rows (shared).
 * c11..c0: nleaves*p^2 coefficient values at the interior nodes (row r =
 * (a-1)*p + (b-1), a along x), any may be NULL (= 0).  X (optional, nleaves x
 * p^2 x 4p) receives the leaf solution operators A_ii^{-1} A_ib.  info: nleaves
 * ints (LAPACK getrf info of A_ii). */
int sslab_hps_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints);
int sslab_hps_leaf_dtn(int nleaves, int q, const double* Dx, const double* Dxx,
                       const double* Dy, const double* Dyy, const double* Nb,
                       const double* Ni, const double* c11, const double* c22,
                       const double* c1, const double* c2, const double* c0, double* T,
                       double* X, double* work, int* iwork, int* info, void* stream);

/* ---- Spectral double slabs, 2-D, variable coefficients -----------------------
 * The reference's 'spectral' slab discretization (solver/spectral/spectralSolver.py:
 * 161-173 constructPDO2D, 198-235 spectralSolver; solver/solver.py:145-157): every double
 * slab one qx x qy tensor Chebyshev grid (qx = ordx + 1 odd, qy = ordy + 1).  For nslab
 * slabs: Y_s = (A_ii^{-1} A_ib)[centre line] (m x nbdry, m = qy - 2, nbdry = 2m + 2(qx-2),
 * columns [left | right | bottom | top], corners dropped), the S maps and boundary-data
 * maps of multislab/oms.py:99-129, 180-186.  Dx..Dyy: qx x qx / qy x qy derivative
 * matrices of the slab grid (shared); c11, c22, c1, c2, c0: nslab*(qx-2)*(qy-2) values at
 * the interior nodes (row (a-1)(qy-2) + (b-1), a along x), any may be NULL (= 0); info:
 * nslab ints. */
int sslab_spectral2d_sizes(int nslab, int qx, int qy, long long* work_doubles,
                           long long* work_ints);
int sslab_spectral2d_slab_maps(int nslab, int qx, int qy, const double* Dx, const double* Dxx,
                               const double* Dy, const double* Dyy, const double* c11,
                               const double* c22, const double* c1, const double* c2,
                               const double* c0, double* Y, double* work, int* iwork, int* info,
                               void* stream);

/* ---- HPS slab assembly, 3-D Chebyshev leaves (q^3 points, p = q - 2) ----------
 * The 3-D discretization of the reference's cube drivers (test/3D/cubeTest.py:114-126:
 * solverOptions('hpsalt', [p_disc]*3, a), hpsalt's Domain_Driver, solver/solver.py:
 * 120-144; absent upstream).  Leaf DtN maps T_l = N_b - N_i A_ii^{-1} A_ib (6p^2 x 6p^2;
 * faces x-lo, x-hi, y-lo, y-hi, z-lo, z-hi, each p^2 nodes in lexicographic order of
 * its two tangential grid indices, edges and corners excluded).  Dx..Dzz: q x q
 * derivative matrices; Nb (6p^2 x 6p^2), Ni (6p^2 x p^3): outward-normal derivative
 * rows (shared by all leaves); c11, c22, c33, c1, c2, c3, c0: nleaves*p^3 values at the
 * interior nodes (row ((a-1) p + (b-1)) p + (c-1), a along x), any may be NULL (= 0);
 * X (optional, nleaves x p^3 x 6p^2) receives A_ii^{-1} A_ib; info: nleaves ints.
 * Box merges and pair factors of 3-D slabs use the generic sslab_hps_merge,
 * sslab_gather2, sslab_dgetrf_batched and sslab_dgetrs(_trans)_batched. */
int sslab_hps3_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints);
int sslab_hps3_leaf_dtn(int nleaves, int q, const double* Dx, const double* Dxx,
                        const double* Dy, const double* Dyy, const double* Dz,
                        const double* Dzz, const double* Nb, const double* Ni,
                        const double* c11, const double* c22, const double* c33,
                        const double* c1, const double* c2, const double* c3,
                        const double* c0, double* T, double* X, double* work, int* iwork,
                        int* info, void* stream);

/* Merge nmerge box pairs (alpha_b = Ta + b*sTa, beta_b = Tb + b*sTb) that share c
 * edge nodes (positions ca in alpha, cb in beta) into boxes with n_new boundary
 * nodes; na[i] / nb_idx[i] give the position of merged node i in alpha / beta
 * (-1 when absent).  All index arrays are device int arrays. */
int sslab_hps_merge_sizes(int nmerge, int n_new, int c, long long* work_doubles,
                          long long* work_ints);
int sslab_hps_merge(int nmerge, int n_a, int n_b, int n_new, int c, const double* Ta,
                    long long sTa, const double* Tb, long long sTb, const int* ca,
                    const int* cb, const int* na, const int* nb_idx, double* Tnew,
                    long long sTnew, double* work, int* iwork, int* info, void* stream);
/* sslab_hps_merge that also keeps every merge's solution operator
 * Xsol_b = (T^a_cc + T^b_cc)^{-1} T_{c,new} (c x n_new, batch stride sXsol): the shared
 * edge's values are -Xsol u_new, which the downward pass (interior reconstruction,
 * solver/pde_solver.py:169-201 solve_dir) applies. */
int sslab_hps_merge_ex(int nmerge, int n_a, int n_b, int n_new, int c, const double* Ta,
                       long long sTa, const double* Tb, long long sTb, const int* ca,
                       const int* cb, const int* na, const int* nb_idx, double* Tnew,
                       long long sTnew, double* work, int* iwork, int* info, double* Xsol,
                       long long sXsol, void* stream);

/* Double-slab S maps: for pairs (slab i, slab i+1) of slab DtN maps Ts (n_s x n_s,
 * stride sTs), Y_i = (T^i_cc + T^{i+1}_cc)^{-1} [T^i(ca, cola) | T^{i+1}(cb, colb)]
 * (m x ncols, m = c), column j taken from slab i when cola[j] >= 0, else from
 * slab i+1.  Workspace from sslab_hps_merge_sizes(npair, ncols, c). */
int sslab_hps_pair_solve(int npair, int n_s, int c, int ncols, const double* Ts,
                         long long sTs, const int* ca, const int* cb, const int* cola,
                         const int* colb, double* Y, double* work, int* iwork, int* info,
                         void* stream);

/* S-system right-hand sides for r columns of global Dirichlet data (multislab/oms.py:
 * 180-186, rhs = -(A_ii^{-1} A_ib[:, Igb] f)[Ic], for every double slab at once).
 * Y: nb x m x ncols pair-solve output (ncols = 2m + 4 nxp, nxp = boundary nodes per
 * slab along x); F: (2 N nxp + 2m) x r boundary data ordered [bottom | top | left |
 * right] with leading dimension ldF; out: nb*m x r (ldo).  N = nb + 1. */
int sslab_oms_rhs(int nb, int m, int nxp, int ncols, const double* Y, const double* F,
                  int ldF, int r, int N, double* out, int ldo, void* stream);

/* sslab_oms_rhs followed by the first red-black elimination level of the S-system
 * (identity diagonal blocks): one zero kernel and one grouped launch that write the
 * entry state of sslab_rb_solve_ex with SSLAB_RB_IDENTITY_DIAGONAL |
 * SSLAB_RB_LEVEL0_REDUCED -- V's even block rows rhs_i - SiM_i rhs_{i-1} - SiP_i
 * rhs_{i+1}, X's odd block rows rhs_j, X's even block rows zero (V's odd rows are left
 * as they are).  Ycb/Yct (ceil(nb/2) x m x 4 nxp each) are the even rows' combined maps
 * over bottom/top rows [(i-1) nxp, (i+3) nxp) of F; Ycl (m x m) maps the left boundary
 * into row 0, Ycr (m x m) the right boundary into the last even row that sees it; they
 * are built once per operator (HPSSlabOperator.rb0_maps in the Python package).
 * nb >= 2, V and X distinct.  No reference counterpart: the reference forms rhs
 * (multislab/oms.py:180-186) and eliminates in separate steps
 * (direct_solve/omsdirectsolve.py:221-279). */
int sslab_oms_rhs_rb0(int nb, int m, int nxp, int ncols, const double* Y, const double* Ycb,
                      const double* Yct, const double* Ycl, const double* Ycr, const double* F,
                      int ldF, int r, int N, double* V, int ldv, double* X, int ldx,
                      void* stream);

/* Boundary-folded forward sweep of the oms solve.  maps (nb*m x nF, leading dim ldm) is
 * the backward sweep's entry state for F = I: sslab_oms_rhs of the nF x nF identity,
 * then sslab_rb_solve_ex(..., SSLAB_RB_IDENTITY_DIAGONAL | SSLAB_RB_FORWARD_ONLY) with
 * the factor of identity diagonal blocks.  sslab_oms_fold_create keeps the nonzero
 * column runs of its blocks (boundary windows) in a compact device copy and returns an
 * opaque plan; sslab_oms_fold_apply writes X = maps F (nb*m x nrhs) for boundary data F
 * (nF x nrhs) with one zero kernel and one grouped GEMM launch, after which
 * sslab_rb_solve_ex(..., SSLAB_RB_IDENTITY_DIAGONAL | SSLAB_RB_BACKWARD_ONLY) finishes
 * the solve.  No reference counterpart (the reference forms rhs, multislab/oms.py:180-186,
 * and runs the whole elimination per solve, direct_solve/omsdirectsolve.py:221-279). */
int sslab_oms_fold_create(int nb, int m, int nF, const double* maps, int ldm, void** plan,
                          void* stream);
int sslab_oms_fold_apply(const void* plan, const double* F, int ldF, int nrhs, double* X,
                         int ldx, void* stream);
/* Mixed-precision solve from boundary data (no reference counterpart: the reference is
 * FP64 throughout).  sslab_oms_fold_enable_f32 keeps an FP32 copy of the plan's maps
 * (once; freed with the plan).  sslab_oms_fold_solve_mixed then solves the identity-
 * diagonal S-system for boundary data F (nF x nrhs, contiguous FP64): X0 from the FP32
 * maps and the FP32 backward sweep (factor32: the FP64 red-black factor rounded to float,
 * as for sslab_rb_solve_mixed), then `iters` refinement steps x += M (B - S x) as there.
 * B = rhs(F) in FP64 (sslab_oms_rhs), X: nb*m x nrhs contiguous doubles; work64:
 * nb*m*nrhs doubles; work32: (2*nb*m + nF)*nrhs floats. */
int sslab_oms_fold_enable_f32(void* plan, void* stream);
int sslab_oms_fold_solve_mixed(const void* plan, const float* factor32, const double* SiM,
                               const double* SiP, const double* F, int ldF, int nrhs,
                               const double* B, double* X, double* work64, float* work32,
                               int iters, void* stream);
/* device bytes held by a plan (its FP32 maps included); number of descriptor groups of
 * its launch */
long long sslab_oms_fold_bytes(const void* plan);
int sslab_oms_fold_groups(const void* plan);
void sslab_oms_fold_destroy(void* plan);

/* Column-batched vector operations of the device GMRES (sslablu_amd/krylov.py), the
 * outer iteration the reference runs with scipy's gmres and the direct solver as
 * preconditioner (cubeTest.py:163-186).  A panel W (n x r, leading dim ldw) holds r
 * independent vectors; a basis V is k panels at stride sV (elements).
 *   sslab_col_dots:  H[j][c] = sum_i V_j[i][c] W[i][c]   (j < k; H is k x r, ldh)
 *   sslab_col_axpy:  W[i][c] += alpha sum_j V_j[i][c] H[j][c]
 *   sslab_col_scale: Out[i][c] = W[i][c] s[c]           (s: r device values) */
int sslab_col_dots(int n, int r, int k, const double* V, long long sV, int ldv,
                   const double* W, int ldw, double* H, int ldh, void* stream);
int sslab_col_axpy(int n, int r, int k, double alpha, const double* V, long long sV, int ldv,
                   const double* H, int ldh, double* W, int ldw, void* stream);
int sslab_col_scale(int n, int r, const double* W, int ldw, const double* s, double* Out,
                    int ldo, void* stream);

/* out[b][i][j] = sum over s in {0,1} of src_s[b][rid_s[i]][cid_s[j]] where both
 * indices are >= 0 (index arrays may be NULL = absent). */
int sslab_gather2(int batch, int nrows, int ncols, const double* s0, int ld0, long long st0,
                  const double* s1, int ld1, long long st1, const int* rid0, const int* cid0,
                  const int* rid1, const int* cid1, double* out, int ldo, long long sto,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SSLAB_HIP_H */
