// Batched explicit inverses by recursive 2 x 2 block elimination (gfx950), FP64 and FP32.
//
// For A = [A11 A12; A21 A22] (A11 h x h, h about n/2 rounded up to a multiple of 32):
//   X11 = A11^{-1}                        (recursively)
//   W = X11 A12,  Z = A21 X11             (one grouped GEMM launch)
//   S = A22 - A21 W                       (one launch)
//   Y = S^{-1}                            (recursively, written straight into X22)
//   X12 = -W Y,  X21 = -Y Z               (one grouped launch)
//   X11 = X11 - X12 Z                     (one launch, in place)
// so an n x n inverse is 2 n^3 flops of MFMA GEMMs plus the 64 x 64 diagonal blocks,
// which k_inv_gj inverts with one wave per matrix (Gauss-Jordan with partial pivoting
// inside the block, no barriers between waves).  The pivoted LU this replaces
// (getrf_batched + getrs_batched against I) spends about 2 us per column on a one-
// workgroup panel; here the sequential part is the 64-column blocks and four launches per
// halving, and every launch covers the whole batch.
//
// The elimination does not pivot between blocks, so its accuracy rests on the diagonal
// blocks and Schur complements being well conditioned (true of the red-black levels of
// the slab S-systems, whose blocks are identity-dominated Schur complements).  Every call
// therefore samples the residual G - A X G on four +-1 columns and falls back to the
// pivoted LU when it misses the tolerance or is not finite (an exactly singular block
// included), so the result is never worse than the LU path's.
#include "inv.hpp"

#include <vector>

#include "gemm.hpp"
#include "lu_f64.hpp"
#include "wave_ops.hpp"

namespace sslab {

namespace {

constexpr int kInvBase = 64;  // diagonal blocks inverted by one wave
constexpr int kInvCheckCols = 4;

inline int inv_split(int n) {
  const int h = ((n / 2 + 31) / 32) * 32;
  return h < n ? h : n - 1;
}

// Elements of E the recursion needs per matrix (W, Z and S of every halving on the
// X22 path; the X11 path reuses the same space before W is written).
long long inv_rec_elems(int n) {
  if (n <= kInvBase) return 0;
  const int h = inv_split(n), h2 = n - h;
  const long long here = 2LL * h * h2 + (long long)h2 * h2 + inv_rec_elems(h2);
  const long long left = inv_rec_elems(h);
  return here > left ? here : left;
}

// One wave inverts one n x n block (n <= 64) by in-place Gauss-Jordan with partial
// pivoting.  Lane i keeps row i in registers as s_i * a_i[.]: the pivot row R of step k
// (a row not pivoted yet, so s = 1) is broadcast unscaled through LDS, the other rows
// subtract (a_i[k] / R[k]) R and set a_i[k] = -a_i[k] / R[k] (no scale enters), and the
// pivot lane only sets its scale to 1 / R[k] and a_r[k] to 1 -- so no lane runs a
// divergent row update.  Pivot: the largest |a_i[k]| over the rows not pivoted yet
// (first lane on ties).  After step k the row pivoted at step k holds row k of the
// inverse with its columns permuted: X[k][p_j] = s_i a_i[j] for i = p_k, where p_j is
// the lane pivoted at step j; the rows are put in place through an LDS tile so that the
// stores are coalesced.  Rows / columns n..63 are the identity's and never pivot.
// The steps are instantiated one by one (GjStep<E, K>) so that every a[K] index is a
// compile-time constant and the row stays in registers.
template <typename E>
struct GjState {
  E a[kInvBase];
  E s;       // row scale
  int step;  // step at which this row was the pivot (-1: not yet)
};

template <typename E, int K>
struct GjStep {
  __device__ __forceinline__ static void run(GjState<E>& g, int lane, bool real, int n,
                                             E* prow, int* pl) {
    constexpr int NB = kInvBase;
    if (K < n) {
      // candidates: the open rows (|a| >= 0; NaN counts as +inf, so the maximum is never
      // NaN and always belongs to an open row -- a non-finite column then makes the
      // result non-finite, which the caller's residual check rejects); the rest -1
      const bool open = real && g.step < 0;
      const E x = fabs(g.a[K]);
      const E v = open ? (x == x ? x : E(INFINITY)) : E(-1);
      const E vmax = wave::wave_max(v);
      const int r = wave::wave_min(v == vmax ? lane : NB);
      if (lane == r) {
#pragma unroll
        for (int j = 0; j < NB; ++j) prow[j] = g.a[j];
        pl[K] = lane;
        g.step = K;
      }
      __syncthreads();
      // the whole row into registers first, so its LDS reads are in flight together
      E R[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) R[j] = prow[j];
      const E ip = E(1) / R[K];
      const bool me = lane == r;
      const E c = me ? E(0) : g.a[K] * ip;
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j != K) g.a[j] = fma(-c, R[j], g.a[j]);
      g.a[K] = me ? E(1) : -c;
      if (me) g.s = ip;
    }
    GjStep<E, K + 1>::run(g, lane, real, n, prow, pl);
  }
};

template <typename E>
struct GjStep<E, kInvBase> {
  __device__ __forceinline__ static void run(GjState<E>&, int, bool, int, E*, int*) {}
};

template <typename E>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_inv_gj(int n, const E* __restrict__ A, int lda, long long sA, E* __restrict__ X, int ldx,
              long long sX) {
  constexpr int NB = kInvBase;
  constexpr int LD = NB + 1;
  __shared__ E tile[NB * LD];
  __shared__ E prow[NB];
  __shared__ int pl[NB];
  const int lane = threadIdx.x;
  const long long b = blockIdx.x;
  A += b * sA;
  X += b * sX;
  // coalesced load of the block through the tile
  if (lane < n)
    for (int i = 0; i < n; ++i) tile[i * LD + lane] = A[(long long)i * lda + lane];
  __syncthreads();
  const bool real = lane < n;
  GjState<E> g;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const E v = tile[lane * LD + j];
    g.a[j] = (real && j < n) ? v : (lane == j ? E(1) : E(0));
  }
  g.s = E(1);
  g.step = -1;
  GjStep<E, 0>::run(g, lane, real, n, prow, pl);
  __syncthreads();
  // scatter X[step][p_j] = s a[j] into the tile, then store its rows
  if (real && g.step >= 0) {  // (every real row pivots exactly once)
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < n) tile[g.step * LD + pl[j]] = g.s * g.a[j];
  }
  __syncthreads();
  if (real)
    for (int i = 0; i < n; ++i) X[(long long)i * ldx + lane] = tile[i * LD + lane];
}

// G[i][c] = +-1 (a fixed hash of i and c), n x kInvCheckCols.
template <typename E>
__global__ void k_inv_fill_g(int n, E* __restrict__ G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * kInvCheckCols) return;
  unsigned h = (unsigned)e * 2654435761u + 0x9e3779b9u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  G[e] = (h & 1u) ? E(1) : E(-1);
}

// flag[b] = 1 unless sum(Q_b^2) <= thr2 (NaN / Inf fail); Q_b: n x kInvCheckCols.
template <typename E>
__global__ __launch_bounds__(256) void k_inv_check(int n, const E* __restrict__ Q, double thr2,
                                                   int* __restrict__ flag) {
  const long long b = blockIdx.x;
  const int cnt = n * kInvCheckCols;
  const E* q = Q + b * cnt;
  double ss = 0.0;
  for (int e = threadIdx.x; e < cnt; e += 256) {
    const double v = (double)q[e];
    ss += v * v;
  }
  __shared__ double red[4];
  ss = wave::wave_sum(ss);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    flag[b] = (t <= thr2) ? 0 : 1;
  }
}

__global__ void k_inv_zero_int(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename E>
int inv_rec(int n, const E* A, int lda, long long sA, E* X, int ldx, long long sX, E* W,
            long long sW, int batch, hipStream_t st) {
  if (n <= kInvBase) {
    hipLaunchKernelGGL(k_inv_gj<E>, dim3(batch), dim3(kInvBase), 0, st, n, A, lda, sA, X, ldx,
                       sX);
    SSLAB_HIP_CHECK(hipGetLastError());
    return kOk;
  }
  const int h = inv_split(n), h2 = n - h;
  const E* A12 = A + h;
  const E* A21 = A + (long long)h * lda;
  const E* A22 = A21 + h;
  E* X12 = X + h;
  E* X21 = X + (long long)h * ldx;
  E* X22 = X21 + h;
  E* Wm = W;                          // h x h2
  E* Zm = Wm + (long long)h * h2;     // h2 x h
  E* S = Zm + (long long)h2 * h;      // h2 x h2
  E* Wn = S + (long long)h2 * h2;     // the Schur complement's own recursion
  SSLAB_TRY(inv_rec<E>(h, A, lda, sA, X, ldx, sX, W, sW, batch, st));
  GemmLaunchT<E> G{};
  G.nd = 2;
  G.d[0] = gemm_desc<E>(h, h2, batch, Wm, h2, sW, nullptr, 0, 0, 0.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(X, ldx, sX, 0, A12, lda, sA, 0, h, 1.0));
  G.d[1] = gemm_desc<E>(h2, h, batch, Zm, h, sW, nullptr, 0, 0, 0.0);
  gemm_add_term<E>(G.d[1], gemm_term<E>(A21, lda, sA, 0, X, ldx, sX, 0, h, 1.0));
  SSLAB_TRY(gemm_launch<E>(G, st));
  G.nd = 1;
  G.d[0] = gemm_desc<E>(h2, h2, batch, S, h2, sW, A22, lda, sA, 1.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(A21, lda, sA, 0, Wm, h2, sW, 0, h, -1.0));
  SSLAB_TRY(gemm_launch<E>(G, st));
  SSLAB_TRY(inv_rec<E>(h2, S, h2, sW, X22, ldx, sX, Wn, sW, batch, st));
  G.nd = 2;
  G.d[0] = gemm_desc<E>(h, h2, batch, X12, ldx, sX, nullptr, 0, 0, 0.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(Wm, h2, sW, 0, X22, ldx, sX, 0, h2, -1.0));
  G.d[1] = gemm_desc<E>(h2, h, batch, X21, ldx, sX, nullptr, 0, 0, 0.0);
  gemm_add_term<E>(G.d[1], gemm_term<E>(X22, ldx, sX, 0, Zm, h, sW, 0, h2, -1.0));
  SSLAB_TRY(gemm_launch<E>(G, st));
  G.nd = 1;
  G.d[0] = gemm_desc<E>(h, h, batch, X, ldx, sX, X, ldx, sX, 1.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(X12, ldx, sX, 0, Zm, h, sW, 0, h2, -1.0));
  return gemm_launch<E>(G, st);
}

}  // namespace

void inv_sizes(int n, int batch, long long* work, long long* iwork) {
  const long long nn = (long long)n * n;
  long long w = inv_rec_elems(n) * batch;
  const long long chk = (long long)kInvCheckCols * n * (1 + 2LL * batch);
  const long long lu = 2 * nn * batch;
  if (chk > w) w = chk;
  if (lu > w) w = lu;
  if (work) *work = w;
  if (iwork) *iwork = 2LL * n * batch + batch;
}

template <typename E>
int inv_batched(int n, const E* A, int lda, long long sA, E* X, int ldx, long long sX, int batch,
                E* work, int* iwork, int* info, double tol, int* fallback, hipStream_t st) {
  if (fallback) *fallback = 0;
  SSLAB_REQUIRE(n >= 1 && batch >= 0 && lda >= n && ldx >= n, "inv_batched: bad sizes");
  if (batch == 0) return kOk;
  SSLAB_REQUIRE(A && X && work && iwork, "inv_batched: null pointer");
  if (tol <= 0) tol = sizeof(E) == 8 ? 1e-11 : 1e-3;
  const long long rec = inv_rec_elems(n);
  SSLAB_TRY(inv_rec<E>(n, A, lda, sA, X, ldx, sX, work, rec, batch, st));

  // sampled residual: Q = G - A (X G)
  E* Gm = work;
  E* P = Gm + (long long)kInvCheckCols * n;
  E* Q = P + (long long)kInvCheckCols * n * batch;
  const long long sPQ = (long long)kInvCheckCols * n;
  int* flag = iwork + 2LL * n * batch;
  hipLaunchKernelGGL(k_inv_fill_g<E>, dim3(ceil_div(n * kInvCheckCols, 256)), dim3(256), 0, st, n,
                     Gm);
  SSLAB_HIP_CHECK(hipGetLastError());
  GemmLaunchT<E> G{};
  G.nd = 1;
  G.d[0] = gemm_desc<E>(n, kInvCheckCols, batch, P, kInvCheckCols, sPQ, nullptr, 0, 0, 0.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(X, ldx, sX, 0, Gm, kInvCheckCols, 0, 0, n, 1.0));
  SSLAB_TRY(gemm_launch<E>(G, st));
  G.d[0] = gemm_desc<E>(n, kInvCheckCols, batch, Q, kInvCheckCols, sPQ, Gm, kInvCheckCols, 0, 1.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(A, lda, sA, 0, P, kInvCheckCols, sPQ, 0, n, -1.0));
  SSLAB_TRY(gemm_launch<E>(G, st));
  const double thr2 = tol * tol * (double)kInvCheckCols * n;
  hipLaunchKernelGGL(k_inv_check<E>, dim3(batch), dim3(256), 0, st, n, Q, thr2, flag);
  SSLAB_HIP_CHECK(hipGetLastError());
  std::vector<int> hf(batch);
  SSLAB_HIP_CHECK(hipMemcpyAsync(hf.data(), flag, sizeof(int) * batch, hipMemcpyDeviceToHost, st));
  SSLAB_HIP_CHECK(hipStreamSynchronize(st));
  bool bad = false;
  for (int b = 0; b < batch; ++b) bad = bad || hf[b] != 0;
  if (!bad) {
    if (info) {
      hipLaunchKernelGGL(k_inv_zero_int, dim3(ceil_div(batch, 256)), dim3(256), 0, st, info, batch);
      SSLAB_HIP_CHECK(hipGetLastError());
    }
    return kOk;
  }
  // pivoted LU against the identity for the whole batch
  if (fallback) *fallback = 1;
  const long long nn = (long long)n * n;
  E* LU = work;
  E* Id = work + nn * batch;
  int* ipiv = iwork;
  int* perm = iwork + (long long)n * batch;
  SSLAB_TRY(copy_block_batched<E>(n, n, A, lda, sA, LU, n, nn, batch, st));
  SSLAB_TRY(getrf_batched<E>(n, LU, n, nn, ipiv, n, info, batch, st));
  SSLAB_TRY(set_identity_batched<E>(n, Id, n, nn, batch, st));
  return getrs_batched<E>(n, n, LU, n, nn, ipiv, n, Id, n, nn, X, ldx, sX, perm, batch, st);
}

template int inv_batched<double>(int, const double*, int, long long, double*, int, long long, int,
                                 double*, int*, int*, double, int*, hipStream_t);
template int inv_batched<float>(int, const float*, int, long long, float*, int, long long, int,
                                float*, int*, int*, double, int*, hipStream_t);

}  // namespace sslab
