// extern "C" entry points of libsslab_hip.so (declared in include/sslab_hip.h).
#include <cstdarg>

#include "../../include/sslab_hip.h"
#include "blocktri_f64.hpp"
#include "gemm.hpp"
#include "hps_f64.hpp"
#include "inv.hpp"
#include "spectral_f64.hpp"
#include "krylov_f64.hpp"
#include "lu_f64.hpp"
#include "oms_f64.hpp"
#include "oms_fold.hpp"
#include "qr_f64.hpp"

namespace sslab {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* get_error() { return g_err; }

}  // namespace sslab

using sslab::as_stream;

extern "C" {

int sslab_abi_version(void) { return 1; }

const char* sslab_last_error(void) { return sslab::get_error(); }

int sslab_dgemm_strided_batched(int transA, int transB, int M, int N, int K, double alpha,
                                const double* A, int lda, long long strideA, const double* B,
                                int ldb, long long strideB, double beta, double* C, int ldc,
                                long long strideC, int batch, void* stream) {
  SSLAB_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "dgemm: negative size");
  SSLAB_REQUIRE(C != nullptr, "dgemm: null C");
  if (M == 0 || N == 0 || batch == 0) return 0;
  SSLAB_REQUIRE(K == 0 || (A != nullptr && B != nullptr), "dgemm: null operand");
  SSLAB_REQUIRE(ldc >= N, "dgemm: ldc < N");
  SSLAB_REQUIRE(lda >= (transA ? M : K) && ldb >= (transB ? K : N), "dgemm: bad lda/ldb");
  if (K == 0) {
    // C = beta*C: express as a product with K = 0 and the input term only.
    sslab::GemmLaunch L{};
    L.nd = 1;
    L.d[0] = sslab::gemm_desc(M, N, batch, C, ldc, strideC, beta != 0.0 ? C : nullptr, ldc,
                              strideC, beta);
    sslab::gemm_add_term(L.d[0], sslab::gemm_term(C, ldc, 0, 0, C, ldc, 0, 0, 0, 0.0));
    return sslab::gemm_launch(L, as_stream(stream));
  }
  return sslab::gemm_strided(transA, transB, M, N, K, alpha, A, lda, strideA, B, ldb, strideB,
                             beta, C, ldc, strideC, batch, as_stream(stream));
}

int sslab_dgemm_grouped(const sslab_gemm_desc* descs, int ndesc, void* stream) {
  SSLAB_REQUIRE(ndesc >= 0 && (ndesc == 0 || descs != nullptr), "dgemm_grouped: bad table");
  sslab::GemmLaunch L{};
  L.nd = 0;
  for (int i = 0; i < ndesc; ++i) {
    const sslab_gemm_desc& e = descs[i];
    SSLAB_REQUIRE(e.nterms >= 1 && e.nterms <= 3, "dgemm_grouped: descriptor %d: 1..3 terms", i);
    SSLAB_REQUIRE(e.M >= 0 && e.N >= 0 && e.batch >= 0, "dgemm_grouped: descriptor %d: negative size", i);
    if (e.M == 0 || e.N == 0 || e.batch == 0) continue;
    SSLAB_REQUIRE(e.C != nullptr && e.ldc >= e.N, "dgemm_grouped: descriptor %d: bad C", i);
    SSLAB_REQUIRE(e.beta == 0.0 || (e.Cin != nullptr && e.ldcin >= e.N),
                  "dgemm_grouped: descriptor %d: beta != 0 needs Cin", i);
    sslab::GemmDesc d = sslab::gemm_desc(e.M, e.N, e.batch, e.C, e.ldc, e.strideC,
                                         e.beta != 0.0 ? e.Cin : nullptr, e.ldcin, e.strideCin,
                                         e.beta);
    for (int t = 0; t < e.nterms; ++t) {
      const sslab_gemm_term& m = e.term[t];
      SSLAB_REQUIRE(m.K >= 1 && m.A != nullptr && m.B != nullptr,
                    "dgemm_grouped: descriptor %d term %d: K >= 1 and operands required", i, t);
      SSLAB_REQUIRE(m.lda >= (m.transA ? e.M : m.K) && m.ldb >= (m.transB ? m.K : e.N),
                    "dgemm_grouped: descriptor %d term %d: bad lda/ldb", i, t);
      sslab::gemm_add_term(d, sslab::gemm_term(m.A, m.lda, m.strideA, m.transA ? 1 : 0, m.B,
                                               m.ldb, m.strideB, m.transB ? 1 : 0, m.K,
                                               m.alpha));
    }
    if (L.nd == sslab::kMaxDesc) {
      SSLAB_TRY(sslab::gemm_launch(L, as_stream(stream)));
      L.nd = 0;
    }
    L.d[L.nd++] = d;
  }
  if (L.nd > 0) SSLAB_TRY(sslab::gemm_launch(L, as_stream(stream)));
  return 0;
}

int sslab_sgemm_strided_batched(int transA, int transB, int M, int N, int K, float alpha,
                                const float* A, int lda, long long strideA, const float* B,
                                int ldb, long long strideB, float beta, float* C, int ldc,
                                long long strideC, int batch, void* stream) {
  SSLAB_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "sgemm: negative size");
  SSLAB_REQUIRE(C != nullptr, "sgemm: null C");
  if (M == 0 || N == 0 || batch == 0) return 0;
  SSLAB_REQUIRE(K > 0, "sgemm: K must be positive");
  SSLAB_REQUIRE(A != nullptr && B != nullptr, "sgemm: null operand");
  SSLAB_REQUIRE(ldc >= N, "sgemm: ldc < N");
  SSLAB_REQUIRE(lda >= (transA ? M : K) && ldb >= (transB ? K : N), "sgemm: bad lda/ldb");
  return sslab::gemm_strided<float>(transA, transB, M, N, K, alpha, A, lda, strideA, B, ldb,
                                    strideB, beta, C, ldc, strideC, batch, as_stream(stream));
}

int sslab_gemm_set_policy(const int* p, int n) {
  SSLAB_REQUIRE(p != nullptr && n == 7, "gemm_set_policy: expected 7 integers");
  sslab::GemmPolicy q;
  q.target_wgs = p[0]; q.split_wgs = p[1]; q.max_ksplit = p[2]; q.min_split_k = p[3];
  q.cfg_mid = p[4]; q.cfg_split = p[5]; q.max_tile = p[6];
  return sslab::gemm_set_policy(q);
}

int sslab_rb_set_backward_policies(const int* p, int nsteps) {
  SSLAB_REQUIRE(nsteps >= 0 && nsteps <= 16, "rb_set_backward_policies: 0..16 steps");
  SSLAB_REQUIRE(nsteps == 0 || p != nullptr, "rb_set_backward_policies: null table");
  sslab::GemmPolicy q[16];
  for (int u = 0; u < nsteps; ++u) {
    const int* e = p + 7 * u;
    q[u].target_wgs = e[0]; q[u].split_wgs = e[1]; q[u].max_ksplit = e[2];
    q[u].min_split_k = e[3]; q[u].cfg_mid = e[4]; q[u].cfg_split = e[5]; q[u].max_tile = e[6];
  }
  return sslab::rb_set_backward_policies(q, nsteps);
}

int sslab_gemm_get_policy(int* p, int n) {
  SSLAB_REQUIRE(p != nullptr && n == 7, "gemm_get_policy: expected 7 integers");
  const sslab::GemmPolicy q = sslab::gemm_get_policy();
  p[0] = q.target_wgs; p[1] = q.split_wgs; p[2] = q.max_ksplit; p[3] = q.min_split_k;
  p[4] = q.cfg_mid; p[5] = q.cfg_split; p[6] = q.max_tile;
  return 0;
}

int sslab_dgetrf_batched(int n, double* A, int lda, long long strideA, int* ipiv,
                         long long strideP, int* info, int batch, void* stream) {
  return sslab::getrf_batched(n, A, lda, strideA, ipiv, strideP, info, batch, as_stream(stream));
}

int sslab_inv_sizes(int n, int batch, long long* work_doubles, long long* work_ints) {
  if (n < 1 || batch < 0) {
    sslab::set_error("sslab_inv_sizes: bad sizes");
    return sslab::kBadArgument;
  }
  sslab::inv_sizes(n, batch, work_doubles, work_ints);
  return sslab::kOk;
}

int sslab_dinv_batched(int n, const double* A, int lda, long long strideA, double* X, int ldx,
                       long long strideX, int batch, double* work, int* iwork, int* info,
                       double tol, int* used_lu, void* stream) {
  return sslab::inv_batched<double>(n, A, lda, strideA, X, ldx, strideX, batch, work, iwork, info,
                                    tol, used_lu, as_stream(stream));
}

int sslab_dgetrs_batched(int n, int nrhs, const double* LU, int lda, long long strideA,
                         const int* ipiv, long long strideP, const double* B, int ldb,
                         long long strideB, double* X, int ldx, long long strideX,
                         int* perm_work, int batch, void* stream) {
  return sslab::getrs_batched(n, nrhs, LU, lda, strideA, ipiv, strideP, B, ldb, strideB, X, ldx,
                              strideX, perm_work, batch, as_stream(stream));
}

int sslab_dgetrs_trans_batched(int n, int nrhs, const double* LU, int lda, long long strideA,
                               const int* ipiv, long long strideP, const double* B, int ldb,
                               long long strideB, double* X, int ldx, long long strideX,
                               double* work, int* perm_work, int batch, void* stream) {
  return sslab::getrs_trans_batched(n, nrhs, LU, lda, strideA, ipiv, strideP, B, ldb, strideB, X,
                                    ldx, strideX, work, perm_work, batch, as_stream(stream));
}

int sslab_dgeqrf_batched(int m, int n, double* A, int lda, long long strideA, double* Q, int ldq,
                         long long strideQ, int* perm, long long stridePerm, int batch,
                         void* stream) {
  return sslab::geqrf_batched(m, n, A, lda, strideA, Q, ldq, strideQ, perm, stridePerm, batch,
                              as_stream(stream));
}

int sslab_dtrsm_upper_batched(int n, int nrhs, const double* U, int ldu, long long strideU,
                              double* X, int ldx, long long strideX, int batch, void* stream) {
  return sslab::trsm_upper_batched(n, nrhs, U, ldu, strideU, X, ldx, strideX, batch,
                                   as_stream(stream));
}

int sslab_add_diag_blocks(int bs, int nblk, int nparent, const double* X, double* Y, int ldy,
                          long long strideY, void* stream) {
  return sslab::add_diag_blocks(bs, nblk, nparent, X, Y, ldy, strideY, as_stream(stream));
}

int sslab_rb_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
                   long long* work_ints) {
  return sslab::rb_sizes(nb, m, factor_doubles, work_doubles, work_ints);
}

int sslab_rb_info_count(int nb) { return nb >= 1 ? sslab::rb_info_count(nb) : 0; }

int sslab_rb_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
                    double* factor, double* work, int* iwork, int* info, void* stream) {
  return sslab::rb_factor(nb, m, T, SiM, SiP, factor, work, iwork, info, as_stream(stream));
}

int sslab_rb_solve(int nb, int m, int nrhs, const double* factor, double* V, int ldv, double* X,
                   int ldx, void* stream) {
  return sslab::rb_solve(nb, m, nrhs, factor, V, ldv, X, ldx, as_stream(stream));
}

int sslab_rb_solve_ex(int nb, int m, int nrhs, const double* factor, double* V, int ldv,
                      double* X, int ldx, int flags, void* stream) {
  return sslab::rb_solve_ex(nb, m, nrhs, factor, V, ldv, X, ldx, flags, as_stream(stream));
}

int sslab_rb_forward_windowed(int nb, int m, int ncols, const double* factor, double* V, int ldv,
                              double* X, int ldx, const int* lo, const int* hi, void* stream) {
  return sslab::rb_forward_windowed(nb, m, ncols, factor, V, ldv, X, ldx, lo, hi,
                                    as_stream(stream));
}

int sslab_rb_chain_status(int* aborted) { return sslab::rb_chain_status(aborted); }

int sslab_rb_solve_multi(int nsys, const int* nb, const int* m, const int* nrhs,
                         const double* const* factor, double* const* V, const int* ldv,
                         double* const* X, const int* ldx, void* stream) {
  return sslab::rb_solve_multi(nsys, nb, m, nrhs, factor, V, ldv, X, ldx, as_stream(stream));
}

int sslab_rb_factor_f32(int nb, int m, const float* T, const float* SiM, const float* SiP,
                        float* factor32, float* work32, int* iwork, int* info, void* stream) {
  return sslab::rb_factor_f32(nb, m, T, SiM, SiP, factor32, work32, iwork, info, as_stream(stream));
}

int sslab_cast_f64_to_f32(long long n, const double* x, float* y, void* stream) {
  return sslab::cast_f64_to_f32(n, x, y, as_stream(stream));
}

int sslab_rb_solve_f32(int nb, int m, int nrhs, const float* factor32, float* V, int ldv,
                       float* X, int ldx, void* stream) {
  return sslab::rb_solve_f32(nb, m, nrhs, factor32, V, ldv, X, ldx, as_stream(stream));
}

int sslab_rb_residual(int nb, int m, int nrhs, const double* T, const double* SiM,
                      const double* SiP, const double* B, int ldb, const double* X, int ldx,
                      double* R, int ldr, void* stream) {
  return sslab::rb_residual(nb, m, nrhs, T, SiM, SiP, B, ldb, X, ldx, R, ldr, as_stream(stream));
}

int sslab_rb_solve_mixed(int nb, int m, int nrhs, const float* factor32, const double* T,
                         const double* SiM, const double* SiP, const double* B, double* X,
                         double* work64, float* work32, int iters, void* stream) {
  const long long n = (long long)nb * m * nrhs;
  return sslab::rb_solve_mixed(nb, m, nrhs, factor32, T, SiM, SiP, B, X, work64, work32,
                               work32 ? work32 + n : nullptr, iters, as_stream(stream));
}

int sslab_rb_solve_mixed_noalloc(int nb, int m, int nrhs, const float* factor32, const double* T,
                                 const double* SiM, const double* SiP, const double* B,
                                 double* X, int iters, void* stream) {
  return sslab::rb_solve_mixed_noalloc(nb, m, nrhs, factor32, T, SiM, SiP, B, X, iters,
                                       as_stream(stream));
}

int sslab_thomas_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
                       long long* work_ints) {
  return sslab::thomas_sizes(nb, m, factor_doubles, work_doubles, work_ints);
}

int sslab_thomas_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
                        double* factor, double* work, int* iwork, int* info, void* stream) {
  return sslab::thomas_factor(nb, m, T, SiM, SiP, factor, work, iwork, info, as_stream(stream));
}

int sslab_thomas_solve(int nb, int m, int nrhs, const double* factor, double* V, int ldv,
                       double* X, int ldx, void* stream) {
  return sslab::thomas_solve(nb, m, nrhs, factor, V, ldv, X, ldx, as_stream(stream));
}

int sslab_hps_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints) {
  return sslab::hps_leaf_sizes(nleaves, q, work_doubles, work_ints);
}

int sslab_hps_leaf_dtn(int nleaves, int q, const double* Dx, const double* Dxx, const double* Dy,
                       const double* Dyy, const double* Nb, const double* Ni, const double* c11,
                       const double* c22, const double* c1, const double* c2, const double* c0,
                       double* T, double* X, double* work, int* iwork, int* info, void* stream) {
  return sslab::hps_leaf_dtn(nleaves, q, Dx, Dxx, Dy, Dyy, Nb, Ni, c11, c22, c1, c2, c0, T, X,
                             work, iwork, info, as_stream(stream));
}

int sslab_hps3_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints) {
  return sslab::hps3_leaf_sizes(nleaves, q, work_doubles, work_ints);
}

int sslab_hps3_leaf_dtn(int nleaves, int q, const double* Dx, const double* Dxx, const double* Dy,
                        const double* Dyy, const double* Dz, const double* Dzz, const double* Nb,
                        const double* Ni, const double* c11, const double* c22, const double* c33,
                        const double* c1, const double* c2, const double* c3, const double* c0,
                        double* T, double* X, double* work, int* iwork, int* info, void* stream) {
  const double* D[6] = {Dx, Dxx, Dy, Dyy, Dz, Dzz};
  const double* coef[7] = {c11, c22, c33, c1, c2, c3, c0};
  return sslab::hps3_leaf_dtn(nleaves, q, D, Nb, Ni, coef, T, X, work, iwork, info,
                              as_stream(stream));
}

int sslab_spectral2d_sizes(int nslab, int qx, int qy, long long* work_doubles,
                           long long* work_ints) {
  return sslab::spectral2d_sizes(nslab, qx, qy, work_doubles, work_ints);
}

int sslab_spectral2d_slab_maps(int nslab, int qx, int qy, const double* Dx, const double* Dxx,
                               const double* Dy, const double* Dyy, const double* c11,
                               const double* c22, const double* c1, const double* c2,
                               const double* c0, double* Y, double* work, int* iwork, int* info,
                               void* stream) {
  return sslab::spectral2d_slab_maps(nslab, qx, qy, Dx, Dxx, Dy, Dyy, c11, c22, c1, c2, c0, Y,
                                     work, iwork, info, as_stream(stream));
}

int sslab_hps_merge_sizes(int nmerge, int n_new, int c, long long* work_doubles,
                          long long* work_ints) {
  return sslab::hps_merge_sizes(nmerge, n_new, c, work_doubles, work_ints);
}

int sslab_hps_merge(int nmerge, int n_a, int n_b, int n_new, int c, const double* Ta,
                    long long sTa, const double* Tb, long long sTb, const int* ca, const int* cb,
                    const int* na, const int* nb_idx, double* Tnew, long long sTnew, double* work,
                    int* iwork, int* info, void* stream) {
  return sslab::hps_merge(nmerge, n_a, n_b, n_new, c, Ta, sTa, Tb, sTb, ca, cb, na, nb_idx, Tnew,
                          sTnew, work, iwork, info, as_stream(stream));
}

int sslab_hps_merge_ex(int nmerge, int n_a, int n_b, int n_new, int c, const double* Ta,
                       long long sTa, const double* Tb, long long sTb, const int* ca,
                       const int* cb, const int* na, const int* nb_idx, double* Tnew,
                       long long sTnew, double* work, int* iwork, int* info, double* Xsol,
                       long long sXsol, void* stream) {
  return sslab::hps_merge(nmerge, n_a, n_b, n_new, c, Ta, sTa, Tb, sTb, ca, cb, na, nb_idx, Tnew,
                          sTnew, work, iwork, info, as_stream(stream), Xsol, sXsol);
}

int sslab_hps_pair_solve(int npair, int n_s, int c, int ncols, const double* Ts, long long sTs,
                         const int* ca, const int* cb, const int* cola, const int* colb, double* Y,
                         double* work, int* iwork, int* info, void* stream) {
  return sslab::hps_pair_solve(npair, n_s, c, ncols, Ts, sTs, ca, cb, cola, colb, Y, work, iwork,
                               info, as_stream(stream));
}

int sslab_oms_rhs(int nb, int m, int nxp, int ncols, const double* Y, const double* F, int ldF,
                  int r, int N, double* out, int ldo, void* stream) {
  return sslab::oms_rhs(nb, m, nxp, ncols, Y, F, ldF, r, N, out, ldo, as_stream(stream));
}

int sslab_oms_rhs_rb0(int nb, int m, int nxp, int ncols, const double* Y, const double* Ycb,
                      const double* Yct, const double* Ycl, const double* Ycr, const double* F,
                      int ldF, int r, int N, double* V, int ldv, double* X, int ldx,
                      void* stream) {
  return sslab::oms_rhs_rb0(nb, m, nxp, ncols, Y, Ycb, Yct, Ycl, Ycr, F, ldF, r, N, V, ldv, X,
                            ldx, as_stream(stream));
}

int sslab_oms_fold_create(int nb, int m, int nF, const double* maps, int ldm, void** plan,
                          void* stream) {
  sslab::FoldPlan* P = nullptr;
  const int rc = sslab::oms_fold_create(nb, m, nF, maps, ldm, &P, as_stream(stream));
  if (plan) *plan = P;
  return rc;
}

int sslab_oms_fold_apply(const void* plan, const double* F, int ldF, int nrhs, double* X,
                         int ldx, void* stream) {
  return sslab::oms_fold_apply(static_cast<const sslab::FoldPlan*>(plan), F, ldF, nrhs, X, ldx,
                               as_stream(stream));
}

int sslab_oms_fold_enable_f32(void* plan, void* stream) {
  return sslab::oms_fold_enable_f32(static_cast<sslab::FoldPlan*>(plan), as_stream(stream));
}

int sslab_oms_fold_solve_mixed(const void* plan, const float* factor32, const double* SiM,
                               const double* SiP, const double* F, int ldF, int nrhs,
                               const double* B, double* X, double* work64, float* work32,
                               int iters, void* stream) {
  return sslab::oms_fold_solve_mixed(static_cast<const sslab::FoldPlan*>(plan), factor32, SiM,
                                     SiP, F, ldF, nrhs, B, X, work64, work32, iters,
                                     as_stream(stream));
}

long long sslab_oms_fold_bytes(const void* plan) {
  return sslab::oms_fold_bytes(static_cast<const sslab::FoldPlan*>(plan));
}

int sslab_oms_fold_groups(const void* plan) {
  return sslab::oms_fold_groups(static_cast<const sslab::FoldPlan*>(plan));
}

void sslab_oms_fold_destroy(void* plan) { sslab::oms_fold_destroy(static_cast<sslab::FoldPlan*>(plan)); }

int sslab_col_dots(int n, int r, int k, const double* V, long long sV, int ldv,
                   const double* W, int ldw, double* H, int ldh, void* stream) {
  return sslab::col_dots(n, r, k, V, sV, ldv, W, ldw, H, ldh, as_stream(stream));
}

int sslab_col_axpy(int n, int r, int k, double alpha, const double* V, long long sV, int ldv,
                   const double* H, int ldh, double* W, int ldw, void* stream) {
  return sslab::col_axpy(n, r, k, alpha, V, sV, ldv, H, ldh, W, ldw, as_stream(stream));
}

int sslab_col_scale(int n, int r, const double* W, int ldw, const double* s, double* Out,
                    int ldo, void* stream) {
  return sslab::col_scale(n, r, W, ldw, s, Out, ldo, as_stream(stream));
}

int sslab_gather2(int batch, int nrows, int ncols, const double* s0, int ld0, long long st0,
                  const double* s1, int ld1, long long st1, const int* rid0, const int* cid0,
                  const int* rid1, const int* cid1, double* out, int ldo, long long sto,
                  void* stream) {
  return sslab::gather2(batch, nrows, ncols, s0, ld0, st0, s1, ld1, st1, rid0, cid0, rid1, cid1,
                        out, ldo, sto, as_stream(stream));
}

}  // extern "C"
