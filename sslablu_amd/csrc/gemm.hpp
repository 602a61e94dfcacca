// Grouped / batched GEMM on gfx950 MFMA, FP64 (v_mfma_f64_16x16x4_f64) and FP32
// (v_mfma_f32_16x16x4_f32, FP32 accumulation).
//
//   C_b = beta * Cin_b + sum_t alpha_t * op(A_t,b) * op(B_t,b)        (row-major)
//
// One launch executes up to kMaxDesc descriptors; each descriptor is a strided
// batch of equally-shaped products with up to kMaxTerms summed terms (the terms
// share M and N but may have different K).  Summed terms let one launch express
// the block-tridiagonal sweeps  v_i - SiM_i y_{i-1} - SiP_i y_{i+1}  without a
// second pass over the output.  Cin may alias C (in-place update).
#pragma once

#include "sslab_common.hpp"

namespace sslab {

constexpr int kMaxTerms = 3;
constexpr int kMaxDesc = 12;  // 12 descriptors = 3.5 KB of kernel arguments

template <class T>
struct same_t {
  using type = T;
};

template <typename T>
struct GemmTermT {
  const T* A;   // op(A): M x K
  const T* B;   // op(B): K x N
  long long strideA;  // batch strides in elements
  long long strideB;
  int lda, ldb;
  int K;
  int transA, transB;  // 0: as stored (row-major), 1: transposed
  T alpha;
};

template <typename T>
struct GemmDescT {
  GemmTermT<T> term[kMaxTerms];
  int nterms;
  int M, N, batch;
  const T* Cin;  // may be null when beta == 0
  long long strideCin;
  int ldcin;
  T* C;
  long long strideC;
  int ldc;
  T beta;
  int tile_begin;  // filled in by gemm_launch
  int tiles_m, tiles_n;
  int ksplit;  // workgroups sharing one output tile along K (filled in by gemm_launch)
  int kchunk;  // K range of one of them, a multiple of the tile's BK
  int atomic;  // 1: accumulate into C atomically (requires Cin == C, beta == 1), so several
               // descriptors (or launches) may add into the same output in any order
  int nosplit;  // 1: never split this descriptor's K range (its C is then written once,
                // by plain stores when beta == 0, with no initialisation pass)
};

template <typename T>
struct GemmLaunchT {
  GemmDescT<T> d[kMaxDesc];
  int nd;
  int xcd_chunk;  // logical tiles per XCD chunk (set by gemm_launch)
};

// A dependency-ordered launch (k_gemm_chain): up to kChainDesc descriptors whose tiles
// run in descriptor order inside one launch.  Batch element b of descriptor i writes the
// output block sig1[i] - 1 + b * sig_stride[i] (sig1 = 0: none) and reads, through term
// t's B operand, block dep1[i][t] - 1 + b * dep_stride[i][t] (dep1 = 0: none); a tile
// starts once every block it reads has been written by all of the launch's tiles that
// write it (need[], filled in by gemm_chain_launch).  Blocks may only be read by later
// descriptors than the ones that write them.  cnt (kChainBlocks ints) and ctl (3 ints:
// ticket, finished tiles, abort flag) are device memory that starts at zero; the kernel
// leaves the counters at zero again (the abort flag stays set until the host clears it).
// Launches that share cnt / ctl must not overlap.
constexpr int kChainDesc = 8;
constexpr int kChainBlocks = 128;

template <typename T>
struct GemmChainT {
  GemmDescT<T> d[kChainDesc];
  int nd;
  int sig1[kChainDesc];
  int sig_stride[kChainDesc];
  int dep1[kChainDesc][kMaxTerms];
  int dep_stride[kChainDesc][kMaxTerms];
  int need[kChainBlocks];
  int* cnt;
  int* ctl;
  int spin_limit;  // polls of a counter before a tile gives up and sets the abort flag
};

using GemmTerm = GemmTermT<double>;
using GemmDesc = GemmDescT<double>;
using GemmLaunch = GemmLaunchT<double>;

template <typename T>
inline GemmTermT<T> gemm_term(const T* A, int lda, long long sA, int transA, const T* B, int ldb,
                              long long sB, int transB, int K,
                              typename same_t<T>::type alpha) {
  GemmTermT<T> t;
  t.A = A; t.B = B; t.strideA = sA; t.strideB = sB; t.lda = lda; t.ldb = ldb;
  t.K = K; t.transA = transA; t.transB = transB; t.alpha = alpha;
  return t;
}

template <typename T>
inline GemmDescT<T> gemm_desc(int M, int N, int batch, T* C, int ldc, long long sC,
                              const typename same_t<T>::type* Cin, int ldcin, long long sCin,
                              typename same_t<T>::type beta) {
  GemmDescT<T> d{};
  d.M = M; d.N = N; d.batch = batch; d.C = C; d.ldc = ldc; d.strideC = sC;
  d.Cin = Cin; d.ldcin = ldcin; d.strideCin = sCin; d.beta = beta; d.nterms = 0;
  return d;
}

template <typename T>
inline void gemm_add_term(GemmDescT<T>& d, const GemmTermT<T>& t) { d.term[d.nterms++] = t; }

// Tile / split-K selection of gemm_launch, per calling thread (an autotuner sets it
// before the calls it has tuned).  Defaults reproduce the built-in heuristic.
struct GemmPolicy {
  int target_wgs = 1024;  // largest tile that still yields this many workgroups
  int split_wgs = 512;    // split-K launches aim at this many workgroups
  int max_ksplit = 64;    // at most this many K parts per output tile
  int min_split_k = 256;  // split K only when the longest summed K reaches this
  int cfg_mid = 3;        // tile id below 128x64 (2: 64x64 BK32, 3: 64x64 BK16, ...)
  int cfg_split = 3;      // tile id of split-K launches (both for more than 32 columns)
  int max_tile = 0;       // 0: up to 128x128, 1: up to 128x64, 2: no 128-row tiles
};
int gemm_set_policy(const GemmPolicy& p);
GemmPolicy gemm_get_policy();

// Launch all descriptors in `L` (L.nd <= kMaxDesc) on `stream`.  Descriptors with
// batch == 0, M == 0 or N == 0 are skipped.  Returns a Status.  (double, float)
template <typename T>
int gemm_launch(GemmLaunchT<T>& L, hipStream_t stream);

// Launch a dependency-ordered chain (FP64) on `stream` with the current policy's split
// tile (cfg_split: ids 2-5, 9, 12-14), split budget (split_wgs) and part cap (max_ksplit).
template <typename T>
int gemm_chain_launch(GemmChainT<T>& L, hipStream_t stream);

// Convenience: single strided-batched product with one term.
template <typename T>
int gemm_strided(int transA, int transB, int M, int N, int K, typename same_t<T>::type alpha,
                 const T* A, int lda, long long sA, const T* B, int ldb, long long sB,
                 typename same_t<T>::type beta, T* C, int ldc, long long sC, int batch,
                 hipStream_t stream);

// X[i][j] = 0 for i < rows, j < cols (leading dimension ld); a kernel, never a memset.
template <typename T>
int zero_2d(long long rows, int cols, T* X, int ld, hipStream_t st);

// Element-wise helpers for the mixed-precision paths (a kernel each, no memsets).
int cast_f64_to_f32(long long n, const double* x, float* y, hipStream_t st);
int cast_f32_to_f64(long long n, const float* x, double* y, hipStream_t st);
// y += x (x float, y double)
int acc_f32_into_f64(long long n, const float* x, double* y, hipStream_t st);
// z = a - b (double)
int sub_f64(long long n, const double* a, const double* b, double* z, hipStream_t st);

}  // namespace sslab
