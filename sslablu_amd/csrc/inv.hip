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
// which k_inv_gj inverts with one workgroup of sixteen waves per matrix (Gauss-Jordan
// with partial pivoting inside the block, one barrier per panel of four pivot steps).  The
// pivoted LU this replaces (getrf_batched + getrs_batched against I) spends about 2 us per
// column on a one-workgroup panel; here the sequential part is the 64-column blocks and
// four launches per halving, and every launch covers the whole batch.
//
// The elimination does not pivot between blocks, so its accuracy rests on the diagonal
// blocks and Schur complements being well conditioned (true of the red-black levels of
// the slab S-systems, whose blocks are identity-dominated Schur complements, and of the
// HPS leaf and merge systems).  Every call therefore samples the normwise backward error
// ||G - A X G||_F / (||A||_F ||X G||_F) on four +-1 columns G and falls back to the
// pivoted LU when it misses the tolerance or is not finite (an exactly singular block
// included), so the result is never worse than the LU path's.  The ratio is scale-free:
// a matrix of large norm (a leaf's spectral second derivatives scale with p^4 / h^2)
// passes when its inverse is as good as a backward-stable one.
#include "inv.hpp"

#include <vector>

#include "gemm.hpp"
#include "lu_f64.hpp"
#include "wave_ops.hpp"

namespace sslab {

namespace {

constexpr int kInvBase = 64;  // diagonal blocks inverted by one workgroup (Gauss-Jordan)
constexpr int kGjWaves = 16;  // waves of that workgroup; each owns kInvBase / kGjWaves columns
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

// Pivot key of a column value: the high 32 bits of |x| (sign cleared; NaN above +inf)
// plus one, so every candidate row ranks above the rows that may not pivot (key 0).  For
// FP64 the key drops the low 32 mantissa bits: the chosen pivot is within a relative
// 2^-20 of the largest, which partial pivoting's stability does not notice.
__device__ __forceinline__ unsigned gj_key(double x) {
  return (unsigned)((unsigned long long)__double_as_longlong(fabs(x)) >> 32) + 1u;
}
__device__ __forceinline__ unsigned gj_key(float x) { return __float_as_uint(fabsf(x)) + 1u; }

// One workgroup of kGjWaves waves inverts one n x n block (n <= 64) by in-place
// Gauss-Jordan with partial pivoting.  Lane i of wave w keeps row i restricted to the
// wave's CW = 64 / kGjWaves columns CW w .. CW w + CW - 1, as s_i * (stored values).  The
// steps go by panels: the CW steps k of panel g pivot on the columns of wave g.  Step k
// picks the pivot row p (the largest |a_i[k]| over the rows not pivoted yet, first lane on
// ties), sets every row's multiplier c_i = a_i[k] / R[k] (c_p = 0) and, in every column j,
// subtracts c_i R[j] from row i (R: the pivot row, read from lane p with readlane);
// column k then becomes a_i[k] = -c_i (a_p[k] = 1) and lane p's scale 1 / R[k].  The
// pivot row is used unscaled and keeps its own values, so no lane runs a divergent row
// update.  Wave g runs its panel's steps on its own columns alone and then publishes the
// panel's pivots, 1 / R[k] and multipliers through LDS (one barrier per panel, the buffers
// alternating by panel parity so that the next panel's writes never meet this panel's
// reads); every other wave then applies the same steps, in the same order, to its columns.
// A column sees exactly the operations of step-by-step Gauss-Jordan, so the result does
// not depend on the panel width; only the barriers (one per CW steps) and the
// serialisation change: panel g + 1 starts once wave g + 1 has applied panel g.  After
// step k the row pivoted at step k holds row k of the inverse with its columns permuted:
// X[k][p_j] = s_i a_i[j] for i = p_k, where p_j is the row pivoted at step j; the rows are
// put in place through an LDS tile so that the stores are coalesced.  Rows / columns
// n..63 are the identity's and never pivot.
template <typename E, bool kFull>
__global__ __launch_bounds__(kGjWaves * 64) void k_inv_gj(int n_arg, const E* __restrict__ A,
                                                int lda, long long sA, E* __restrict__ X,
                                                int ldx, long long sX) {
  constexpr int NB = kInvBase;
  const int n = kFull ? NB : n_arg;  // full blocks: every step guard folds away
  constexpr int CW = NB / kGjWaves;  // columns per wave = steps per panel
  constexpr int LD = NB + 1;
  __shared__ E tile[NB * LD];
  __shared__ E cbuf[2][CW][NB];
  __shared__ E ibuf[2][CW];
  __shared__ int pbuf[2][CW];
  __shared__ int pl[NB];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const long long b = blockIdx.x;
  A += b * sA;
  X += b * sX;
  // coalesced load of the block through the tile
  if (lane < n)
    for (int i = w; i < n; i += kGjWaves) tile[i * LD + lane] = A[(long long)i * lda + lane];
  __syncthreads();
  const bool real = lane < n;
  E a[CW];
#pragma unroll
  for (int j = 0; j < CW; ++j) {
    const int c = w * CW + j;
    const E v = tile[lane * LD + c];  // read for every lane; replaced outside the block
    a[j] = (real && c < n) ? v : (lane == c ? E(1) : E(0));
  }
  E s = E(1);
  int step = -1;  // step at which this row was the pivot (-1: not yet)
  for (int g = 0; g * CW < n; ++g) {
    const int par = g & 1;
    if (w == g) {
#pragma unroll
      for (int kk = 0; kk < CW; ++kk) {
        const int k = g * CW + kk;
        if (k < n) {  // uniform over the wave
          const unsigned key = (real && step < 0) ? gj_key(a[kk]) : 0u;
          const unsigned kmax = wave::wave_max_u32(key);
          // the maximum is attained by some candidate row (n - k of them remain)
          const unsigned long long bal = __ballot(key == kmax);
          const int p = __builtin_ctzll(bal);
          const E piv = wave::readlane_e(a[kk], p);
          const E ip = E(1) / piv;
          const E c = (lane == p) ? E(0) : a[kk] * ip;
          cbuf[par][kk][lane] = c;
          if (lane == 0) {
            pbuf[par][kk] = p;
            ibuf[par][kk] = ip;
            pl[k] = p;
          }
#pragma unroll
          for (int j = 0; j < CW; ++j) {
            const E r = wave::readlane_e(a[j], p);
            a[j] = fma(-c, r, a[j]);
          }
          a[kk] = (lane == p) ? E(1) : -c;
          if (lane == p) {
            s = ip;
            step = k;
          }
        }
      }
    }
    __syncthreads();
    if (w != g) {
      E c[CW];
#pragma unroll
      for (int kk = 0; kk < CW; ++kk) c[kk] = cbuf[par][kk][lane];
#pragma unroll
      for (int kk = 0; kk < CW; ++kk) {
        const int k = g * CW + kk;
        if (k < n) {  // uniform over the workgroup
          const int p = __builtin_amdgcn_readfirstlane(pbuf[par][kk]);
#pragma unroll
          for (int j = 0; j < CW; ++j) {
            const E r = wave::readlane_e(a[j], p);
            a[j] = fma(-c[kk], r, a[j]);
          }
          if (lane == p) {
            s = ibuf[par][kk];
            step = k;
          }
        }
      }
    }
  }
  __syncthreads();
  // scatter X[step][p_j] = s a[j] into the tile, then store its rows
  if (real) {  // (every real row pivots exactly once)
#pragma unroll
    for (int j = 0; j < CW; ++j) {
      const int c = w * CW + j;
      if (c < n) tile[step * LD + pl[c]] = s * a[j];
    }
  }
  __syncthreads();
  if (lane < n)
    for (int i = w; i < n; i += kGjWaves) X[(long long)i * ldx + lane] = tile[i * LD + lane];
}

// G[i][c] = +-1 (a fixed hash of i and c), n x kInvCheckCols, written to Gs (shared by the
// batch) and to columns kInvCheckCols.. of every matrix's n x (2 kInvCheckCols) block B_b.
template <typename E>
__global__ void k_inv_fill_g(int n, int batch, E* __restrict__ Gs, E* __restrict__ B) {
  constexpr int C = kInvCheckCols;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)n * C * (batch + 1)) return;
  const int b = (int)(e / ((long long)n * C)) - 1;  // -1: the shared copy
  const int ic = (int)(e % ((long long)n * C));
  unsigned h = (unsigned)ic * 2654435761u + 0x9e3779b9u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  const E g = (h & 1u) ? E(1) : E(-1);
  if (b < 0) Gs[ic] = g;
  else B[(long long)b * n * 2 * C + (long long)(ic / C) * 2 * C + C + ic % C] = g;
}

// Per matrix b: B_b = [P | G] (P = X_b G) and Y_b = A_b B_b = [A P | A G], n x 2C each.
// Q = G - A P is the residual of the sample; E ||A g||^2 = ||A||_F^2 for a column g of
// independent +-1 entries, so ||A G||_F^2 / C estimates ||A||_F^2 without a pass of its
// own over A (an underestimate only makes the test stricter).  flag[b] = 0 when
// ||Q||_F^2 <= tol^2 (||A G||_F^2 / C) ||P||_F^2, else 1 (NaN / Inf fail).
template <typename E>
__global__ __launch_bounds__(256) void k_inv_check(int n, const E* __restrict__ B,
                                                   const E* __restrict__ Y, double tol2,
                                                   int* __restrict__ flag) {
  constexpr int C = kInvCheckCols;
  const long long b = blockIdx.x;
  const long long cnt = (long long)n * 2 * C;
  const E* bb = B + b * cnt;
  const E* yy = Y + b * cnt;
  double qq = 0.0, pp = 0.0, aa = 0.0;
  for (long long e = threadIdx.x; e < cnt; e += 256) {
    const int c = (int)(e % (2 * C));
    if (c < C) {
      const double q = (double)bb[e + C] - (double)yy[e];  // g - (A P)
      const double v = (double)bb[e];
      qq += q * q;
      pp += v * v;
    } else {
      const double v = (double)yy[e];
      aa += v * v;
    }
  }
  __shared__ double red[3][4];
  qq = wave::wave_sum(qq);
  pp = wave::wave_sum(pp);
  aa = wave::wave_sum(aa);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = qq;
    red[1][w] = pp;
    red[2][w] = aa;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double tq = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double tp = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    const double ta = (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]);
    flag[b] = (tq <= tol2 * (ta / C) * tp) ? 0 : 1;
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
    if (n == kInvBase)
      hipLaunchKernelGGL((k_inv_gj<E, true>), dim3(batch), dim3(kGjWaves * 64), 0, st, n, A, lda,
                         sA, X, ldx, sX);
    else
      hipLaunchKernelGGL((k_inv_gj<E, false>), dim3(batch), dim3(kGjWaves * 64), 0, st, n, A,
                         lda, sA, X, ldx, sX);
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

template <typename E>
int inv_elim_check(int n, const E* A, int lda, long long sA, E* X, int ldx, long long sX,
                   int batch, E* work, double tol, int* flag, hipStream_t st) {
  const long long rec = inv_rec_elems(n);
  SSLAB_TRY(inv_rec<E>(n, A, lda, sA, X, ldx, sX, work, rec, batch, st));
  // B_b = [X_b G | G], Y_b = A_b B_b
  constexpr int C = kInvCheckCols;
  E* Gs = work;
  E* Bm = Gs + (long long)C * n;
  E* Ym = Bm + 2LL * C * n * batch;
  const long long sB = 2LL * C * n;
  const long long nfill = (long long)n * C * (batch + 1);
  hipLaunchKernelGGL(k_inv_fill_g<E>, dim3((unsigned)((nfill + 255) / 256)), dim3(256), 0, st, n,
                     batch, Gs, Bm);
  SSLAB_HIP_CHECK(hipGetLastError());
  GemmLaunchT<E> G{};
  G.nd = 1;
  G.d[0] = gemm_desc<E>(n, C, batch, Bm, 2 * C, sB, nullptr, 0, 0, 0.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(X, ldx, sX, 0, Gs, C, 0, 0, n, 1.0));
  SSLAB_TRY(gemm_launch<E>(G, st));
  G.d[0] = gemm_desc<E>(n, 2 * C, batch, Ym, 2 * C, sB, nullptr, 0, 0, 0.0);
  gemm_add_term<E>(G.d[0], gemm_term<E>(A, lda, sA, 0, Bm, 2 * C, sB, 0, n, 1.0));
  SSLAB_TRY(gemm_launch<E>(G, st));
  hipLaunchKernelGGL(k_inv_check<E>, dim3(batch), dim3(256), 0, st, n, Bm, Ym, tol * tol, flag);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace

void inv_sizes(int n, int batch, long long* work, long long* iwork) {
  const long long nn = (long long)n * n;
  long long w = inv_rec_elems(n) * batch;
  const long long chk = (long long)kInvCheckCols * n * (1 + 4LL * batch);
  const long long lu = 2 * nn * batch;
  if (chk > w) w = chk;
  if (lu > w) w = lu;
  if (work) *work = w;
  if (iwork) *iwork = 2LL * n * batch + batch;
}

template <typename E>
int inv_batched(int n, const E* A, int lda, long long sA, E* X, int ldx, long long sX, int batch,
                E* work, int* iwork, int* info, double tol, int* fallback, hipStream_t st,
                int mode, int* flags) {
  if (fallback) *fallback = 0;
  SSLAB_REQUIRE(n >= 1 && batch >= 0 && lda >= n && ldx >= n, "inv_batched: bad sizes");
  SSLAB_REQUIRE(mode >= kInvChecked && mode <= kInvLU && (mode != kInvDeferred || flags),
                "inv_batched: bad mode or missing flags");
  if (batch == 0) return kOk;
  SSLAB_REQUIRE(A && X && work && iwork, "inv_batched: null pointer");
  if (tol <= 0) tol = sizeof(E) == 8 ? 1e-12 : 1e-3;  // normwise backward error
  if (mode != kInvLU) {
    int* flag = mode == kInvDeferred ? flags : iwork + 2LL * n * batch;
    SSLAB_TRY(inv_elim_check<E>(n, A, lda, sA, X, ldx, sX, batch, work, tol, flag, st));
    bool bad = false;
    if (mode == kInvChecked) {
      std::vector<int> hf(batch);
      SSLAB_HIP_CHECK(
          hipMemcpyAsync(hf.data(), flag, sizeof(int) * batch, hipMemcpyDeviceToHost, st));
      SSLAB_HIP_CHECK(hipStreamSynchronize(st));
      for (int b = 0; b < batch; ++b) bad = bad || hf[b] != 0;
    }
    if (!bad) {
      if (info) {
        hipLaunchKernelGGL(k_inv_zero_int, dim3(ceil_div(batch, 256)), dim3(256), 0, st, info,
                           batch);
        SSLAB_HIP_CHECK(hipGetLastError());
      }
      return kOk;
    }
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
                                 double*, int*, int*, double, int*, hipStream_t, int, int*);
template int inv_batched<float>(int, const float*, int, long long, float*, int, long long, int,
                                float*, int*, int*, double, int*, hipStream_t, int, int*);

}  // namespace sslab
