// Grouped / batched FP64 GEMM on gfx950 MFMA (v_mfma_f64_16x16x4_f64).
//
//   C_b = beta * Cin_b + sum_t alpha_t * op(A_t,b) * op(B_t,b)        (row-major)
//
// One launch executes up to kMaxDesc descriptors; each descriptor is a strided
// batch of equally-shaped products with up to kMaxTerms summed terms (the terms
// share M and N but may have different K).  Summed terms let one launch express
// the block-tridiagonal sweeps  v_i - W^L_i v_{i-1} - W^R_i v_{i+1}  without a
// second pass over the output.  Cin may alias C (in-place update).
#pragma once

#include "sslab_common.hpp"

namespace sslab {

constexpr int kMaxTerms = 3;
constexpr int kMaxDesc = 6;

struct GemmTerm {
  const double* A;   // op(A): M x K
  const double* B;   // op(B): K x N
  long long strideA;  // batch strides in elements
  long long strideB;
  int lda, ldb;
  int K;
  int transA, transB;  // 0: as stored (row-major), 1: transposed
  double alpha;
};

struct GemmDesc {
  GemmTerm term[kMaxTerms];
  int nterms;
  int M, N, batch;
  const double* Cin;  // may be null when beta == 0
  long long strideCin;
  int ldcin;
  double* C;
  long long strideC;
  int ldc;
  double beta;
  int tile_begin;  // filled in by gemm_launch
  int tiles_m, tiles_n;
};

struct GemmLaunch {
  GemmDesc d[kMaxDesc];
  int nd;
};

// Helpers to fill descriptors.
inline GemmTerm gemm_term(const double* A, int lda, long long sA, int transA,
                          const double* B, int ldb, long long sB, int transB,
                          int K, double alpha) {
  GemmTerm t;
  t.A = A; t.B = B; t.strideA = sA; t.strideB = sB; t.lda = lda; t.ldb = ldb;
  t.K = K; t.transA = transA; t.transB = transB; t.alpha = alpha;
  return t;
}

inline GemmDesc gemm_desc(int M, int N, int batch, double* C, int ldc, long long sC,
                          const double* Cin, int ldcin, long long sCin, double beta) {
  GemmDesc d{};
  d.M = M; d.N = N; d.batch = batch; d.C = C; d.ldc = ldc; d.strideC = sC;
  d.Cin = Cin; d.ldcin = ldcin; d.strideCin = sCin; d.beta = beta; d.nterms = 0;
  return d;
}

inline void gemm_add_term(GemmDesc& d, const GemmTerm& t) { d.term[d.nterms++] = t; }

// Launch all descriptors in `L` (L.nd <= kMaxDesc) on `stream`.  Descriptors with
// batch == 0, M == 0 or N == 0 are skipped.  Returns a Status.
int gemm_launch(GemmLaunch& L, hipStream_t stream);

// Convenience: single strided-batched product with one term.
int gemm_strided(int transA, int transB, int M, int N, int K, double alpha,
                 const double* A, int lda, long long sA, const double* B, int ldb,
                 long long sB, double beta, double* C, int ldc, long long sC, int batch,
                 hipStream_t stream);

}  // namespace sslab
