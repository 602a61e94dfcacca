// Batched FP64 LU (partial pivoting) and triangular solves for gfx950.
//
// getrf: right-looking blocked LU with 256-column outer blocks, so the trailing
// Schur update A22 -= L21 U12 is an MFMA GEMM with K = 256.  Each 256-column panel
// is factored recursively (halves: left half, its interchanges on the right half,
// U12 solve, panel-local GEMM update, right half, its interchanges on the left
// half) down to a base panel that one workgroup per matrix factors with the panel
// held column-major in LDS (pivot search and rank-1 updates on consecutive LDS
// words).
//
// Triangular solves (getrs, the U12 block rows): two-level blocking -- 32-row
// diagonal solves, one column per lane with the triangle in LDS, inside 256-row
// blocks whose off-block updates are MFMA GEMMs with K = 256.
#include "lu_f64.hpp"

#include <cstdlib>

#include "gemm.hpp"
#include "wave_ops.hpp"

namespace sslab {

namespace {

constexpr int kPanelThreads = 256;
constexpr int kTrsmNB = 32;  // diagonal block of the getrs substitutions

template <typename E>
__device__ __forceinline__ E abs_e(E v) { return v < E(0) ? -v : v; }

__global__ void k_zero_int(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Factor panel columns [j0, j0+w) over rows [j0, n) of every batch matrix.
// kLds: the panel is staged column-major in dynamic LDS (rows x w elements);
// otherwise it is factored in place in global memory.
template <typename E, bool kLds, int NT = kPanelThreads>
__global__ void __launch_bounds__(NT)
k_getrf_panel(int n, int j0, int w, E* __restrict__ A, int lda, long long sA,
              int* __restrict__ ipiv, long long sP, int* __restrict__ info) {
  constexpr int kPanelThreads = NT;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  E* smem = reinterpret_cast<E*>(smem_raw);
  __shared__ E red_v[kPanelThreads / 64];
  __shared__ int red_i[kPanelThreads / 64];
  __shared__ int s_piv;

  const int b = blockIdx.x;
  E* Ab = A + (long long)b * sA;
  int* pv = ipiv + (long long)b * sP;
  const int rows = n - j0;
  const int tid = threadIdx.x;

  E* P;
  long long rs, cs;
  if (kLds) {
    P = smem;
    rs = 1;
    cs = rows;
    for (int e = tid; e < rows * w; e += kPanelThreads) {
      const int i = e / w, j = e - (e / w) * w;
      smem[(long long)j * rows + i] = Ab[(long long)(j0 + i) * lda + j0 + j];
    }
    __syncthreads();
  } else {
    P = Ab + (long long)j0 * lda + j0;
    rs = lda;
    cs = 1;
  }

  for (int c = 0; c < w; ++c) {
    // Pivot search: largest |P(i,c)| for i >= c, first index on ties (idamax).
    E best = -1.0;
    int bi = rows;
    for (int i = c + tid; i < rows; i += kPanelThreads) {
      const E v = abs_e(P[i * rs + c * cs]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const E ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if ((tid & 63) == 0) {
      red_v[tid >> 6] = best;
      red_i[tid >> 6] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      E bv = red_v[0];
      int bx = red_i[0];
      for (int q = 1; q < kPanelThreads / 64; ++q) {
        if (red_v[q] > bv || (red_v[q] == bv && red_i[q] < bx)) {
          bv = red_v[q];
          bx = red_i[q];
        }
      }
      if (bx >= rows) bx = c;  // all-NaN column: keep the diagonal
      s_piv = bx;
      pv[j0 + c] = j0 + bx;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != c) {
      for (int j = tid; j < w; j += kPanelThreads) {
        const E t = P[c * rs + j * cs];
        P[c * rs + j * cs] = P[p * rs + j * cs];
        P[p * rs + j * cs] = t;
      }
    }
    __syncthreads();
    const E piv = P[c * rs + c * cs];
    if (piv == 0.0 && tid == 0 && info != nullptr && info[b] == 0) info[b] = j0 + c + 1;
    for (int i = c + 1 + tid; i < rows; i += kPanelThreads) {
      E l = P[i * rs + c * cs];
      if (piv != 0.0) {
        l = l / piv;
        P[i * rs + c * cs] = l;
      }
      for (int j = c + 1; j < w; ++j) P[i * rs + j * cs] -= l * P[c * rs + j * cs];
    }
    __syncthreads();
  }

  if (kLds) {
    for (int e = tid; e < rows * w; e += kPanelThreads) {
      const int i = e / w, j = e - (e / w) * w;
      Ab[(long long)(j0 + i) * lda + j0 + j] = smem[(long long)j * rows + i];
    }
  }
}

using wave::wave_max;
using wave::wave_min;

// Base panel with its rows in registers: thread t holds panel rows t + k NT (k < R;
// rows <= R NT) over the w <= W panel columns, so the rank-1 updates are register FMAs.  The
// column loop is fully unrolled so that every register index is static.  Rows never move
// between lanes during the factorization: each row carries its current position (`pos`),
// an interchange swaps two positions, and every row is stored at its final position at the
// end.
//
// Per column c: each lane's candidate is its best row (largest |a| among positions >= c,
// smallest position on ties), each wave finds its best candidate with DPP reductions and
// publishes (value, position) and the candidate's U row (columns c..W-1) in LDS; after the
// column's single barrier every thread picks the winning wave and reads the pivot row from
// its slot.  The LDS slots alternate between two buffers, so a column's writes never meet
// the previous column's reads.  Same pivot choice as the LDS panel (first index on ties;
// NaN entries lose to every number, so an all-NaN column keeps its diagonal).  (A rolled
// loop over a row that shifts one column per step measured slower: it does the full row
// width of work every column.)
template <typename E, int NT, int W, int R = 1>
__global__ void __launch_bounds__(NT)
k_getrf_panel_reg(int n, int j0, int w, E* __restrict__ A, int lda, long long sA,
                  int* __restrict__ ipiv, long long sP, int* __restrict__ info) {
  static_assert(W <= NT, "panel width <= threads");
  constexpr int NW = NT / 64;
  __shared__ __attribute__((aligned(16))) E s_row[2][NW][W];
  __shared__ __attribute__((aligned(16))) E s_val[2][NW];
  __shared__ __attribute__((aligned(16))) int s_pos[2][NW];
  const int b = blockIdx.x;
  E* Ab = A + (long long)b * sA + (long long)j0 * lda + j0;
  int* pv = ipiv + (long long)b * sP;
  const int rows = n - j0;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  bool live[R];
  int pos[R];
  E P[R][W];
#pragma clang loop unroll(full)
  for (int k = 0; k < R; ++k) {
    const int i = tid + k * NT;
    live[k] = i < rows;
    pos[k] = i;
#pragma clang loop unroll(full)
    for (int j = 0; j < W; ++j) P[k][j] = (live[k] && j < w) ? Ab[(long long)i * lda + j] : E(0);
  }
  int my_piv = 0;
#pragma clang loop unroll(full)
  for (int c = 0; c < W; ++c) {
    if (c >= w) continue;  // uniform; a full unroll keeps every P index static
    const int buf = c & 1;
    // candidates: live rows at positions >= c; NaN ranks below every number (-0.5),
    // non-candidates below that (-1)
    E v = E(-1);
    int vp = INT_MAX;
#pragma clang loop unroll(full)
    for (int k = 0; k < R; ++k) {
      const E a = abs_e(P[k][c]);
      const E vk = (live[k] && pos[k] >= c) ? (a == a ? a : E(-0.5)) : E(-1);
      if (vk > v || (vk == v && pos[k] < vp)) {
        v = vk;
        vp = pos[k];
      }
    }
    const E m = wave_max(v);
    const int lp = wave_min(v == m ? vp : INT_MAX);
    if (lane == 0) {
      s_val[buf][wv] = m;
      s_pos[buf][wv] = m < E(-0.75) ? INT_MAX : lp;
    }
    if (m >= E(-0.75)) {
#pragma clang loop unroll(full)
      for (int k = 0; k < R; ++k) {
        if (pos[k] == lp) {
          // the U row (columns c..W-1), from an even column so that the stores pair up
#pragma clang loop unroll(full)
          for (int j = c & ~1; j < W; ++j) s_row[buf][wv][j] = P[k][j];
        }
      }
    }
    __syncthreads();
    // winning wave: lanes 0..NW-1 read the published candidates and the same DPP reductions
    // pick the largest value, then the smallest position; a ballot names its wave
    const E cv = lane < NW ? s_val[buf][lane] : E(-2);
    const int cp = lane < NW ? s_pos[buf][lane] : INT_MAX;
    const E cm = wave_max(cv);
    const int p = wave_min(cv == cm ? cp : INT_MAX);
    const int q = __ffsll((unsigned long long)__ballot(lane < NW && cv == cm && cp == p)) - 1;
    // the row at position c is always a candidate, so some wave published a row
    const E* prow = s_row[buf][q];
#pragma clang loop unroll(full)
    for (int k = 0; k < R; ++k) {
      if (pos[k] == p) pos[k] = c;
      else if (pos[k] == c) pos[k] = p;
    }
    if (tid == c) my_piv = p;
    const E piv = prow[c];
    if (piv == E(0) && tid == 0 && info != nullptr && info[b] == 0) info[b] = j0 + c + 1;
#pragma clang loop unroll(full)
    for (int k = 0; k < R; ++k) {
      if (live[k] && pos[k] > c) {
        E l = P[k][c];
        if (piv != E(0)) {
          l = l / piv;
          P[k][c] = l;
        }
#pragma clang loop unroll(full)
        for (int j = c + 1; j < W; ++j) P[k][j] -= l * prow[j];
      }
    }
  }
  if (tid < w) pv[j0 + tid] = j0 + my_piv;
#pragma clang loop unroll(full)
  for (int k = 0; k < R; ++k) {
    if (live[k]) {
#pragma clang loop unroll(full)
      for (int j = 0; j < W; ++j)
        if (j < w) Ab[(long long)pos[k] * lda + j] = P[k][j];
    }
  }
}

// Apply the interchanges of rows r0..r0+w-1 (in order) to the columns
// [a0, a1) U [b0, b1).
template <typename E>
__global__ void k_laswp(int r0, int w, int a0, int a1, int b0, int b1, E* __restrict__ A,
                        int lda, long long sA, const int* __restrict__ ipiv, long long sP) {
  const int na = a1 - a0;
  const int ncols = na + (b1 - b0);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncols) return;
  const int col = t < na ? a0 + t : b0 + (t - na);
  const int j0 = r0;
  E* Ab = A + (long long)blockIdx.y * sA;
  const int* pv = ipiv + (long long)blockIdx.y * sP;
  for (int c = 0; c < w; ++c) {
    const int r = j0 + c;
    const int p = pv[r];
    if (p != r) {
      const E t0 = Ab[(long long)r * lda + col];
      Ab[(long long)r * lda + col] = Ab[(long long)p * lda + col];
      Ab[(long long)p * lda + col] = t0;
    }
  }
}

// B[0:nr, col] <- T^{-1} B[0:nr, col] for the nr x nr triangle T (nr <= NB), one
// column per lane.  kLower: unit lower triangle; else upper with its diagonal.
// Both run the same backward (upper) substitution: the lower solve works in reversed
// coordinates (J L J is upper triangular for the reversal J), because the compiler
// schedules the unrolled backward loop without hoisting the whole triangle into
// registers, while the unrolled forward loop spills.
// The solved matrix M is T (kTrans = false) or T^T read from the same storage: kRev
// selects the reversed coordinates (M lower triangular), kUnit a unit diagonal (the
// packed LU's L, whose stored diagonal is U's).
template <typename E, int NB, bool kRev, bool kUnit, bool kTrans>
__global__ void __launch_bounds__(256)
k_trsm_cols(int nr, int ncols, const E* __restrict__ T, int ldt, long long sT,
            E* __restrict__ B, int ldb, long long sB) {
  __shared__ E Ts[NB * NB];  // the triangle in solve coordinates
  const E* Tb = T + (long long)blockIdx.y * sT;
  E* Bb = B + (long long)blockIdx.y * sB;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, j = e % NB;
    const int ri = kRev ? NB - 1 - i : i, rj = kRev ? NB - 1 - j : j;
    const long long off = kTrans ? (long long)rj * ldt + ri : (long long)ri * ldt + rj;
    Ts[e] = (ri < nr && rj < nr && !(kUnit && ri == rj)) ? Tb[off] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  const int col0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (col0 >= ncols) return;
  // x[i] is row ri(i) of the column; rows past nr re-read row nr-1 and are zeroed.
  const E* bcol = Bb + col0;
  E x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int ri = kRev ? NB - 1 - i : i;
    const E v = bcol[(long long)(ri < nr ? ri : nr - 1) * ldb];
    x[i] = (ri < nr) ? v : 0.0;
  }
#pragma unroll
  for (int i = NB - 1; i >= 0; --i) {
    E s = x[i];
#pragma unroll
    for (int j = i + 1; j < NB; ++j) s -= Ts[i * NB + j] * x[j];
    x[i] = s / Ts[i * NB + i];
  }
  E* ocol = Bb + col0;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int ri = kRev ? NB - 1 - i : i;
    if (ri < nr) ocol[(long long)ri * ldb] = x[i];
  }
}

// X[perm[i], :] = W[i, :]  (the inverse of k_gather_rows)
template <typename E>
__global__ void k_scatter_rows(int n, int ncols, const int* __restrict__ perm,
                               const E* __restrict__ W, int ldw, long long sW,
                               E* __restrict__ X, int ldx, long long sX) {
  const int b = blockIdx.z;
  const int i = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int dst = perm[(long long)b * n + i];
  X[(long long)b * sX + (long long)dst * ldx + c] = W[(long long)b * sW + (long long)i * ldw + c];
}

// perm = the row order produced by applying ipiv[0..n) to the identity.
__global__ void __launch_bounds__(256)
k_ipiv_to_perm(int n, const int* __restrict__ ipiv, long long sP, int* __restrict__ perm) {
  extern __shared__ int sperm[];
  const int* pv = ipiv + (long long)blockIdx.x * sP;
  int* out = perm + (long long)blockIdx.x * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) sperm[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) {
      const int p = pv[i];
      if (p != i && p >= 0 && p < n) {
        const int t = sperm[i];
        sperm[i] = sperm[p];
        sperm[p] = t;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = sperm[i];
}

// X[i, :] = B[perm[i], :]
template <typename E>
__global__ void k_gather_rows(int n, int ncols, const int* __restrict__ perm,
                              const E* __restrict__ B, int ldb, long long sB,
                              E* __restrict__ X, int ldx, long long sX) {
  const int b = blockIdx.z;
  const int i = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int src = perm[(long long)b * n + i];
  X[(long long)b * sX + (long long)i * ldx + c] = B[(long long)b * sB + (long long)src * ldb + c];
}

template <typename E>
__global__ void k_set_identity(int n, E* __restrict__ X, int ldx, long long sX) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  X[(long long)blockIdx.z * sX + (long long)i * ldx + j] = (i == j) ? 1.0 : 0.0;
}

template <typename E>
__global__ void k_copy_block(int n, const E* __restrict__ X, int ldx, long long sX,
                             E* __restrict__ Y, int ldy, long long sY) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  Y[(long long)blockIdx.z * sY + (long long)i * ldy + j] =
      X[(long long)blockIdx.z * sX + (long long)i * ldx + j];
}

template <typename E>
__global__ void k_zero_block(int m, int n, E* __restrict__ Y, int ldy, long long sY) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int i = blockIdx.y; i < m; i += gridDim.y)
    Y[(long long)blockIdx.z * sY + (long long)i * ldy + j] = 0.0;
}

int max_lds_bytes() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
      cached = v;
    else
      cached = 64 * 1024;
  }
  return cached;
}

}  // namespace

namespace {

constexpr int kOuterNB = 256;  // outer block of getrf and of the triangular solves

template <typename E>
int laswp(int r0, int w, int a0, int a1, int b0, int b1, E* A, int lda, long long sA,
          const int* ipiv, long long sP, int batch, hipStream_t st) {
  const int ncols = (a1 - a0) + (b1 - b0);
  if (w <= 0 || ncols <= 0) return kOk;
  hipLaunchKernelGGL(k_laswp<E>, dim3(ceil_div(ncols, 256), batch), dim3(256), 0, st, r0, w, a0, a1,
                     b0, b1, A, lda, sA, ipiv, sP);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

// X (n x nrhs) <- T^{-1} X for the unit-lower (lower = true) or upper n x n triangle T.
template <typename E>
int trsm_2level(bool lower, int n, int nrhs, const E* T, int ldt, long long sT, E* X,
                int ldx, long long sX, int batch, hipStream_t st) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return kOk;
  constexpr int NB = kTrsmNB;
  const dim3 grid(ceil_div(nrhs, 256), batch);
  auto diag = [&](int ib, int ni) -> int {
    if (lower)
      hipLaunchKernelGGL((k_trsm_cols<E, NB, true, true, false>), grid, dim3(256), 0, st, ni,
                         nrhs, T + (long long)ib * ldt + ib, ldt, sT, X + (long long)ib * ldx, ldx,
                         sX);
    else
      hipLaunchKernelGGL((k_trsm_cols<E, NB, false, false, false>), grid, dim3(256), 0, st, ni,
                         nrhs, T + (long long)ib * ldt + ib, ldt, sT, X + (long long)ib * ldx, ldx,
                         sX);
    SSLAB_HIP_CHECK(hipGetLastError());
    return kOk;
  };
  // X[r0:r0+nr) -= T[r0:r0+nr, c0:c0+nc) X[c0:c0+nc)
  auto upd = [&](int r0, int nr, int c0, int nc) -> int {
    if (nr <= 0 || nc <= 0) return kOk;
    return gemm_strided<E>(0, 0, nr, nrhs, nc, -1.0, T + (long long)r0 * ldt + c0, ldt, sT,
                                X + (long long)c0 * ldx, ldx, sX, 1.0, X + (long long)r0 * ldx,
                                ldx, sX, batch, st);
  };
  if (lower) {
    for (int kb = 0; kb < n; kb += kOuterNB) {
      const int nr = n - kb < kOuterNB ? n - kb : kOuterNB;
      for (int ib = kb; ib < kb + nr; ib += NB) {
        const int ni = kb + nr - ib < NB ? kb + nr - ib : NB;
        SSLAB_TRY(diag(ib, ni));
        SSLAB_TRY(upd(ib + ni, kb + nr - ib - ni, ib, ni));
      }
      SSLAB_TRY(upd(kb + nr, n - kb - nr, kb, nr));
    }
  } else {
    const int nob = ceil_div(n, kOuterNB);
    for (int ob = nob - 1; ob >= 0; --ob) {
      const int kb = ob * kOuterNB;
      const int nr = n - kb < kOuterNB ? n - kb : kOuterNB;
      const int nib = ceil_div(nr, NB);
      for (int q = nib - 1; q >= 0; --q) {
        const int ib = kb + q * NB;
        const int ni = kb + nr - ib < NB ? kb + nr - ib : NB;
        SSLAB_TRY(diag(ib, ni));
        SSLAB_TRY(upd(kb, ib - kb, ib, ni));
      }
      SSLAB_TRY(upd(0, kb, kb, nr));
    }
  }
  return kOk;
}

// X (n x nrhs) <- (T^T)^{-1} X, T read from the packed LU: U^T (lower, non-unit: a
// forward substitution, ut = true) or L^T (unit upper: a backward substitution).
// Off-diagonal blocks are transposed-operand GEMMs on the same storage.
template <typename E>
int trsm_2level_trans(bool ut, int n, int nrhs, const E* T, int ldt, long long sT, E* X,
                      int ldx, long long sX, int batch, hipStream_t st) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return kOk;
  constexpr int NB = kTrsmNB;
  const dim3 grid(ceil_div(nrhs, 256), batch);
  auto diag = [&](int ib, int ni) -> int {
    if (ut)
      hipLaunchKernelGGL((k_trsm_cols<E, NB, true, false, true>), grid, dim3(256), 0, st, ni,
                         nrhs, T + (long long)ib * ldt + ib, ldt, sT, X + (long long)ib * ldx, ldx,
                         sX);
    else
      hipLaunchKernelGGL((k_trsm_cols<E, NB, false, true, true>), grid, dim3(256), 0, st, ni,
                         nrhs, T + (long long)ib * ldt + ib, ldt, sT, X + (long long)ib * ldx, ldx,
                         sX);
    SSLAB_HIP_CHECK(hipGetLastError());
    return kOk;
  };
  // X[r0:r0+nr) -= (T^T)[r0:r0+nr, c0:c0+nc) X[c0:c0+nc) = T[c0:c0+nc, r0:r0+nr)^T X[c0:..)
  auto upd = [&](int r0, int nr, int c0, int nc) -> int {
    if (nr <= 0 || nc <= 0) return kOk;
    return gemm_strided<E>(1, 0, nr, nrhs, nc, -1.0, T + (long long)c0 * ldt + r0, ldt, sT,
                           X + (long long)c0 * ldx, ldx, sX, 1.0, X + (long long)r0 * ldx, ldx,
                           sX, batch, st);
  };
  if (ut) {
    for (int kb = 0; kb < n; kb += kOuterNB) {
      const int nr = n - kb < kOuterNB ? n - kb : kOuterNB;
      for (int ib = kb; ib < kb + nr; ib += NB) {
        const int ni = kb + nr - ib < NB ? kb + nr - ib : NB;
        SSLAB_TRY(diag(ib, ni));
        SSLAB_TRY(upd(ib + ni, kb + nr - ib - ni, ib, ni));
      }
      SSLAB_TRY(upd(kb + nr, n - kb - nr, kb, nr));
    }
  } else {
    const int nob = ceil_div(n, kOuterNB);
    for (int ob = nob - 1; ob >= 0; --ob) {
      const int kb = ob * kOuterNB;
      const int nr = n - kb < kOuterNB ? n - kb : kOuterNB;
      const int nib = ceil_div(nr, NB);
      for (int q = nib - 1; q >= 0; --q) {
        const int ib = kb + q * NB;
        const int ni = kb + nr - ib < NB ? kb + nr - ib : NB;
        SSLAB_TRY(diag(ib, ni));
        SSLAB_TRY(upd(kb, ib - kb, ib, ni));
      }
      SSLAB_TRY(upd(0, kb, kb, nr));
    }
  }
  return kOk;
}

template <typename E>
struct PanelCtx {
  int n, lda, batch, lds, base_lds, base_glb;
  long long sA, sP;
  E* A;
  int* ipiv;
  int* info;
  hipStream_t st;
};

// Widest base panel (8, 16 or 32 columns) whose rows fit in LDS; 0 = none does.
template <typename E>
int base_width(const PanelCtx<E>& c, int rows) {
  for (int w = 32; w >= 8; w >>= 1)
    if ((long long)rows * w * (long long)sizeof(E) <= c.lds) return w;
  return 0;
}

// Register base panel with the fewest waves that hold one row per lane.
template <typename E, int W>
int launch_panel_reg(const PanelCtx<E>& c, int j0, int w, int rows) {
#define SSLAB_PANEL_REG_LAUNCH(NT)                                                          \
  hipLaunchKernelGGL((k_getrf_panel_reg<E, NT, W>), dim3(c.batch), dim3(NT), 0, c.st, c.n, j0, \
                     w, c.A, c.lda, c.sA, c.ipiv, c.sP, c.info)
  if (rows <= 64) SSLAB_PANEL_REG_LAUNCH(64);
  else if (rows <= 128) SSLAB_PANEL_REG_LAUNCH(128);
  else if (rows <= 256) SSLAB_PANEL_REG_LAUNCH(256);
  else if constexpr (W > 32) SSLAB_PANEL_REG_LAUNCH(512);  // callers: rows <= 512
  else if (rows <= 512) SSLAB_PANEL_REG_LAUNCH(512);
  else SSLAB_PANEL_REG_LAUNCH(1024);
#undef SSLAB_PANEL_REG_LAUNCH
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

// Factor rows [j0, n) of columns [j0, j0 + w).
template <typename E>
int panel_rec(const PanelCtx<E>& c, int j0, int w) {
  const int rows = c.n - j0;
  const int bw = base_width<E>(c, rows);
  static const bool reg64 = [] {
    const char* e = getenv("SSLAB_PANEL_REG");
    return !(e && e[0] == '0');
  }();
  const int base = (reg64 && rows <= 512) ? 64
                   : (reg64 && rows > 1024 && rows <= 2048) ? 16
                                                            : (bw > 0 ? bw : 32);
  if (w <= base) {
    static const bool reg = [] {
      const char* e = getenv("SSLAB_PANEL_REG");
      return !(e && e[0] == '0');
    }();
    if (reg && w > 32 && w <= 64 && rows <= 512) {  // 64-column panel: half the recursion
      SSLAB_TRY((launch_panel_reg<E, 64>(c, j0, w, rows)));
      return kOk;
    }
    if (reg && w <= 32 && rows <= 1024) {
      SSLAB_TRY((launch_panel_reg<E, 32>(c, j0, w, rows)));
      return kOk;
    }
    if (reg && w <= 16 && rows <= 2048) {  // two rows per lane, 1024 threads
      hipLaunchKernelGGL((k_getrf_panel_reg<E, 1024, 16, 2>), dim3(c.batch), dim3(1024), 0, c.st,
                         c.n, j0, w, c.A, c.lda, c.sA, c.ipiv, c.sP, c.info);
      SSLAB_HIP_CHECK(hipGetLastError());
      return kOk;
    }
    if (bw > 0) {
      const size_t shm = (size_t)rows * w * sizeof(E);
      static const int pnt = [] {
        const char* e = getenv("SSLAB_PANEL_NT");
        return e ? atoi(e) : kPanelThreads;
      }();
      if (pnt == 64)
        hipLaunchKernelGGL((k_getrf_panel<E, true, 64>), dim3(c.batch), dim3(64), shm, c.st, c.n,
                           j0, w, c.A, c.lda, c.sA, c.ipiv, c.sP, c.info);
      else if (pnt == 128)
        hipLaunchKernelGGL((k_getrf_panel<E, true, 128>), dim3(c.batch), dim3(128), shm, c.st,
                           c.n, j0, w, c.A, c.lda, c.sA, c.ipiv, c.sP, c.info);
      else if (pnt == 512)
        hipLaunchKernelGGL((k_getrf_panel<E, true, 512>), dim3(c.batch), dim3(512), shm, c.st,
                           c.n, j0, w, c.A, c.lda, c.sA, c.ipiv, c.sP, c.info);
      else
        hipLaunchKernelGGL((k_getrf_panel<E, true>), dim3(c.batch), dim3(kPanelThreads), shm, c.st, c.n,
                           j0, w, c.A, c.lda, c.sA, c.ipiv, c.sP, c.info);
    } else {
      hipLaunchKernelGGL((k_getrf_panel<E, false>), dim3(c.batch), dim3(kPanelThreads), 0, c.st, c.n,
                         j0, w, c.A, c.lda, c.sA, c.ipiv, c.sP, c.info);
    }
    SSLAB_HIP_CHECK(hipGetLastError());
    return kOk;
  }
  const int w1 = (w / 2 + 7) / 8 * 8;
  const int w2 = w - w1;
  E* A11 = c.A + (long long)j0 * c.lda + j0;
  SSLAB_TRY(panel_rec<E>(c, j0, w1));
  SSLAB_TRY(laswp<E>(j0, w1, j0 + w1, j0 + w, 0, 0, c.A, c.lda, c.sA, c.ipiv, c.sP, c.batch, c.st));
  SSLAB_TRY(trsm_2level<E>(true, w1, w2, A11, c.lda, c.sA, A11 + w1, c.lda, c.sA, c.batch, c.st));
  SSLAB_TRY(gemm_strided<E>(0, 0, rows - w1, w2, w1, -1.0, A11 + (long long)w1 * c.lda,
                                 c.lda, c.sA, A11 + w1, c.lda, c.sA, 1.0,
                                 A11 + (long long)w1 * c.lda + w1, c.lda, c.sA, c.batch, c.st));
  SSLAB_TRY(panel_rec<E>(c, j0 + w1, w2));
  return laswp<E>(j0 + w1, w2, j0, j0 + w1, 0, 0, c.A, c.lda, c.sA, c.ipiv, c.sP, c.batch, c.st);
}

}  // namespace

template <typename E>
int getrf_batched(int n, E* A, int lda, long long sA, int* ipiv, long long sP, int* info,
                  int batch, hipStream_t stream) {
  SSLAB_REQUIRE(n >= 0 && batch >= 0, "getrf: negative size");
  SSLAB_REQUIRE(lda >= n, "getrf: lda < n");
  SSLAB_REQUIRE(A != nullptr && ipiv != nullptr, "getrf: null pointer");
  if (n == 0 || batch == 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535, "getrf: batch too large (%d)", batch);
  if (info) {
    hipLaunchKernelGGL(k_zero_int, dim3(ceil_div(batch, 256)), dim3(256), 0, stream, info, batch);
    SSLAB_HIP_CHECK(hipGetLastError());
  }
  PanelCtx<E> c{};
  c.n = n;
  c.lda = lda;
  c.batch = batch;
  c.sA = sA;
  c.sP = sP;
  c.A = A;
  c.ipiv = ipiv;
  c.info = info;
  c.st = stream;
  c.lds = max_lds_bytes() - 512;  // room for the static reduction scratch
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getrf_panel<E, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getrf_panel<E, true, 64>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getrf_panel<E, true, 128>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getrf_panel<E, true, 512>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, c.lds);
  for (int j0 = 0; j0 < n; j0 += kOuterNB) {
    const int w = (n - j0) < kOuterNB ? (n - j0) : kOuterNB;
    SSLAB_TRY(panel_rec<E>(c, j0, w));
    SSLAB_TRY(laswp<E>(j0, w, 0, j0, j0 + w, n, A, lda, sA, ipiv, sP, batch, stream));
    const int rest = n - j0 - w;
    if (rest > 0) {
      E* A11 = A + (long long)j0 * lda + j0;
      SSLAB_TRY(trsm_2level<E>(true, w, rest, A11, lda, sA, A11 + w, lda, sA, batch, stream));
      SSLAB_TRY(gemm_strided<E>(0, 0, rest, rest, w, -1.0, A11 + (long long)w * lda, lda, sA,
                                     A11 + w, lda, sA, 1.0, A11 + (long long)w * lda + w, lda, sA,
                                     batch, stream));
    }
  }
  return kOk;
}

template <typename E>
int getrs_batched(int n, int nrhs, const E* LU, int lda, long long sA, const int* ipiv,
                  long long sP, const E* B, int ldb, long long sB, E* X, int ldx,
                  long long sX, int* perm, int batch, hipStream_t stream) {
  SSLAB_REQUIRE(n >= 0 && nrhs >= 0 && batch >= 0, "getrs: negative size");
  SSLAB_REQUIRE(lda >= n && ldb >= nrhs && ldx >= nrhs, "getrs: bad leading dimension");
  SSLAB_REQUIRE(B != X, "getrs: B and X must be distinct buffers");
  if (n == 0 || nrhs == 0 || batch == 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535 && n <= 65535, "getrs: batch or n too large");
  SSLAB_REQUIRE((size_t)n * sizeof(int) <= (size_t)max_lds_bytes(), "getrs: n too large");
  hipLaunchKernelGGL(k_ipiv_to_perm, dim3(batch), dim3(256), (size_t)n * sizeof(int), stream, n,
                     ipiv, sP, perm);
  SSLAB_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gather_rows<E>, dim3(ceil_div(nrhs, 256), n, batch), dim3(256), 0, stream, n,
                     nrhs, perm, B, ldb, sB, X, ldx, sX);
  SSLAB_HIP_CHECK(hipGetLastError());
  SSLAB_TRY(trsm_2level<E>(true, n, nrhs, LU, lda, sA, X, ldx, sX, batch, stream));   // L Y = P B
  return trsm_2level<E>(false, n, nrhs, LU, lda, sA, X, ldx, sX, batch, stream);      // U X = Y
}

template <typename E>
int getrs_trans_batched(int n, int nrhs, const E* LU, int lda, long long sA, const int* ipiv,
                        long long sP, const E* B, int ldb, long long sB, E* X, int ldx,
                        long long sX, E* W, int* perm, int batch, hipStream_t stream) {
  SSLAB_REQUIRE(n >= 0 && nrhs >= 0 && batch >= 0, "getrs_trans: negative size");
  SSLAB_REQUIRE(lda >= n && ldb >= nrhs && ldx >= nrhs, "getrs_trans: bad leading dimension");
  SSLAB_REQUIRE(W != nullptr && W != B && W != X, "getrs_trans: W must be a distinct workspace");
  if (n == 0 || nrhs == 0 || batch == 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535 && n <= 65535, "getrs_trans: batch or n too large");
  SSLAB_REQUIRE((size_t)n * sizeof(int) <= (size_t)max_lds_bytes(), "getrs_trans: n too large");
  // P A = L U, so A^T X = B reads U^T L^T (P X) = B: W = B, W <- U^-T W, W <- L^-T W,
  // X[perm[i]] = W[i]
  const long long sW = (long long)n * nrhs;
  hipLaunchKernelGGL(k_ipiv_to_perm, dim3(batch), dim3(256), (size_t)n * sizeof(int), stream, n,
                     ipiv, sP, perm);
  SSLAB_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_copy_block<E>, dim3(ceil_div(nrhs, 256), n, batch), dim3(256), 0, stream,
                     nrhs, B, ldb, sB, W, nrhs, sW);
  SSLAB_HIP_CHECK(hipGetLastError());
  SSLAB_TRY(trsm_2level_trans<E>(true, n, nrhs, LU, lda, sA, W, nrhs, sW, batch, stream));
  SSLAB_TRY(trsm_2level_trans<E>(false, n, nrhs, LU, lda, sA, W, nrhs, sW, batch, stream));
  hipLaunchKernelGGL(k_scatter_rows<E>, dim3(ceil_div(nrhs, 256), n, batch), dim3(256), 0, stream,
                     n, nrhs, perm, W, nrhs, sW, X, ldx, sX);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

template <typename E>
int trsm_upper_batched(int n, int nrhs, const E* U, int ldu, long long sU, E* X, int ldx,
                       long long sX, int batch, hipStream_t stream) {
  SSLAB_REQUIRE(n >= 0 && nrhs >= 0 && batch >= 0, "trsm: negative size");
  SSLAB_REQUIRE(ldu >= n && ldx >= nrhs, "trsm: bad leading dimension");
  if (n == 0 || nrhs == 0 || batch == 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535, "trsm: batch too large");
  return trsm_2level<E>(false, n, nrhs, U, ldu, sU, X, ldx, sX, batch, stream);
}

template <typename E>
int set_identity_batched(int n, E* X, int ldx, long long sX, int batch, hipStream_t stream) {
  if (n <= 0 || batch <= 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535 && n <= 65535, "set_identity: size too large");
  hipLaunchKernelGGL(k_set_identity<E>, dim3(ceil_div(n, 256), n, batch), dim3(256), 0, stream, n, X,
                     ldx, sX);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

template <typename E>
int copy_block_batched(int m, int n, const E* X, int ldx, long long sX, E* Y, int ldy,
                       long long sY, int batch, hipStream_t stream) {
  if (m <= 0 || n <= 0 || batch <= 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535 && m <= 65535, "copy_block: size too large");
  hipLaunchKernelGGL(k_copy_block<E>, dim3(ceil_div(n, 256), m, batch), dim3(256), 0, stream, n, X,
                     ldx, sX, Y, ldy, sY);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

template <typename E>
int zero_block_batched(int m, int n, E* Y, int ldy, long long sY, int batch,
                       hipStream_t stream) {
  if (m <= 0 || n <= 0 || batch <= 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535, "zero_block: batch too large");
  const int gy = m < 4096 ? m : 4096;
  hipLaunchKernelGGL(k_zero_block<E>, dim3(ceil_div(n, 256), gy, batch), dim3(256), 0, stream, m, n,
                     Y, ldy, sY);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

#define SSLAB_LU_INSTANTIATE(E)                                                                \
  template int getrf_batched<E>(int, E*, int, long long, int*, long long, int*, int, hipStream_t); \
  template int getrs_batched<E>(int, int, const E*, int, long long, const int*, long long,       \
                                const E*, int, long long, E*, int, long long, int*, int,       \
                                hipStream_t);                                                  \
  template int getrs_trans_batched<E>(int, int, const E*, int, long long, const int*, long long, \
                                      const E*, int, long long, E*, int, long long, E*, int*,  \
                                      int, hipStream_t);                                       \
  template int trsm_upper_batched<E>(int, int, const E*, int, long long, E*, int, long long,    \
                                     int, hipStream_t);                                        \
  template int set_identity_batched<E>(int, E*, int, long long, int, hipStream_t);             \
  template int copy_block_batched<E>(int, int, const E*, int, long long, E*, int, long long,    \
                                     int, hipStream_t);                                        \
  template int zero_block_batched<E>(int, int, E*, int, long long, int, hipStream_t);
SSLAB_LU_INSTANTIATE(double)
SSLAB_LU_INSTANTIATE(float)
#undef SSLAB_LU_INSTANTIATE

}  // namespace sslab
