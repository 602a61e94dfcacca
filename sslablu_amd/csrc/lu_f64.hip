// Batched FP64 LU (partial pivoting) and triangular solves for gfx950.
//
// getrf: right-looking blocked LU.  Each panel step is four launches over the
// whole batch: the panel factorization (one workgroup per matrix, the panel held
// column-major in LDS so the pivot search and the rank-1 updates read
// consecutive LDS words), the row interchanges on the remaining columns, a
// unit-lower triangular solve for the U12 block row, and the trailing Schur
// update A22 -= L21 U12 on the MFMA GEMM.
//
// getrs: gather rows by the pivot permutation, then block forward / backward
// substitution: the diagonal-block solves run one column per lane with the
// triangle in LDS, the off-diagonal updates are MFMA GEMMs.
#include "lu_f64.hpp"

#include "gemm.hpp"

namespace sslab {

namespace {

constexpr int kPanelThreads = 256;
constexpr int kTrsmNB = 32;  // diagonal block of the getrs substitutions

__global__ void k_zero_int(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Factor panel columns [j0, j0+w) over rows [j0, n) of every batch matrix.
// kLds: the panel is staged column-major in dynamic LDS (rows x w doubles);
// otherwise it is factored in place in global memory.
template <bool kLds>
__global__ void __launch_bounds__(kPanelThreads)
k_getrf_panel(int n, int j0, int w, double* __restrict__ A, int lda, long long sA,
              int* __restrict__ ipiv, long long sP, int* __restrict__ info) {
  extern __shared__ double smem[];
  __shared__ double red_v[kPanelThreads / 64];
  __shared__ int red_i[kPanelThreads / 64];
  __shared__ int s_piv;

  const int b = blockIdx.x;
  double* Ab = A + (long long)b * sA;
  int* pv = ipiv + (long long)b * sP;
  const int rows = n - j0;
  const int tid = threadIdx.x;

  double* P;
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
    double best = -1.0;
    int bi = rows;
    for (int i = c + tid; i < rows; i += kPanelThreads) {
      const double v = fabs(P[i * rs + c * cs]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
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
      double bv = red_v[0];
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
        const double t = P[c * rs + j * cs];
        P[c * rs + j * cs] = P[p * rs + j * cs];
        P[p * rs + j * cs] = t;
      }
    }
    __syncthreads();
    const double piv = P[c * rs + c * cs];
    if (piv == 0.0 && tid == 0 && info != nullptr && info[b] == 0) info[b] = j0 + c + 1;
    for (int i = c + 1 + tid; i < rows; i += kPanelThreads) {
      double l = P[i * rs + c * cs];
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

// Apply the panel's interchanges (rows j0..j0+w-1) to columns [0,j0) U [j0+w,n).
__global__ void k_laswp(int n, int j0, int w, double* __restrict__ A, int lda, long long sA,
                        const int* __restrict__ ipiv, long long sP) {
  const int ncols = n - w;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncols) return;
  const int col = t < j0 ? t : t + w;
  double* Ab = A + (long long)blockIdx.y * sA;
  const int* pv = ipiv + (long long)blockIdx.y * sP;
  for (int c = 0; c < w; ++c) {
    const int r = j0 + c;
    const int p = pv[r];
    if (p != r) {
      const double t0 = Ab[(long long)r * lda + col];
      Ab[(long long)r * lda + col] = Ab[(long long)p * lda + col];
      Ab[(long long)p * lda + col] = t0;
    }
  }
}

// B[0:nr, col] <- T^{-1} B[0:nr, col] for the nr x nr triangle T (nr <= NB), one
// column per lane.  kLower: unit lower triangle; else upper with its diagonal.
template <int NB, bool kLower>
__global__ void __launch_bounds__(256)
k_trsm_cols(int nr, int ncols, const double* __restrict__ T, int ldt, long long sT,
            double* __restrict__ B, int ldb, long long sB) {
  __shared__ double Ts[NB * NB];
  const double* Tb = T + (long long)blockIdx.y * sT;
  double* Bb = B + (long long)blockIdx.y * sB;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, j = e % NB;
    Ts[e] = (i < nr && j < nr) ? Tb[(long long)i * ldt + j] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  double x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = (i < nr) ? Bb[(long long)i * ldb + col] : 0.0;
  if (kLower) {
#pragma unroll
    for (int i = 1; i < NB; ++i) {
      double s = x[i];
#pragma unroll
      for (int j = 0; j < i; ++j) s -= Ts[i * NB + j] * x[j];
      x[i] = s;
    }
  } else {
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      double s = x[i];
#pragma unroll
      for (int j = i + 1; j < NB; ++j) s -= Ts[i * NB + j] * x[j];
      x[i] = s / Ts[i * NB + i];
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < nr) Bb[(long long)i * ldb + col] = x[i];
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
__global__ void k_gather_rows(int n, int ncols, const int* __restrict__ perm,
                              const double* __restrict__ B, int ldb, long long sB,
                              double* __restrict__ X, int ldx, long long sX) {
  const int b = blockIdx.z;
  const int i = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int src = perm[(long long)b * n + i];
  X[(long long)b * sX + (long long)i * ldx + c] = B[(long long)b * sB + (long long)src * ldb + c];
}

__global__ void k_set_identity(int n, double* __restrict__ X, int ldx, long long sX) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  X[(long long)blockIdx.z * sX + (long long)i * ldx + j] = (i == j) ? 1.0 : 0.0;
}

__global__ void k_copy_block(int n, const double* __restrict__ X, int ldx, long long sX,
                             double* __restrict__ Y, int ldy, long long sY) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  Y[(long long)blockIdx.z * sY + (long long)i * ldy + j] =
      X[(long long)blockIdx.z * sX + (long long)i * ldx + j];
}

__global__ void k_zero_block(int m, int n, double* __restrict__ Y, int ldy, long long sY) {
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

int getrf_batched(int n, double* A, int lda, long long sA, int* ipiv, long long sP, int* info,
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
  // Widest panel (<= 32 columns) whose column-major copy fits in LDS; reserve a
  // little room for the static reduction scratch.
  const int lds = max_lds_bytes() - 512;
  int nb = 32;
  while (nb > 8 && (long long)n * nb * 8 > lds) nb >>= 1;
  const bool use_lds = (long long)n * nb * 8 <= lds;
  if (!use_lds) nb = 32;
  if (use_lds) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_getrf_panel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  }
  for (int j0 = 0; j0 < n; j0 += nb) {
    const int w = (n - j0) < nb ? (n - j0) : nb;
    const size_t shm = use_lds ? (size_t)(n - j0) * w * sizeof(double) : 0;
    if (use_lds)
      hipLaunchKernelGGL(k_getrf_panel<true>, dim3(batch), dim3(kPanelThreads), shm, stream, n, j0,
                         w, A, lda, sA, ipiv, sP, info);
    else
      hipLaunchKernelGGL(k_getrf_panel<false>, dim3(batch), dim3(kPanelThreads), 0, stream, n, j0,
                         w, A, lda, sA, ipiv, sP, info);
    SSLAB_HIP_CHECK(hipGetLastError());
    if (n - w > 0) {
      hipLaunchKernelGGL(k_laswp, dim3(ceil_div(n - w, 256), batch), dim3(256), 0, stream, n, j0,
                         w, A, lda, sA, ipiv, sP);
      SSLAB_HIP_CHECK(hipGetLastError());
    }
    const int rest = n - j0 - w;
    if (rest > 0) {
      double* L11 = A + (long long)j0 * lda + j0;
      double* A12 = A + (long long)j0 * lda + j0 + w;
      hipLaunchKernelGGL((k_trsm_cols<32, true>), dim3(ceil_div(rest, 256), batch), dim3(256), 0,
                         stream, w, rest, L11, lda, sA, A12, lda, sA);
      SSLAB_HIP_CHECK(hipGetLastError());
      double* A21 = A + (long long)(j0 + w) * lda + j0;
      double* A22 = A + (long long)(j0 + w) * lda + j0 + w;
      SSLAB_TRY(gemm_strided(0, 0, rest, rest, w, -1.0, A21, lda, sA, A12, lda, sA, 1.0, A22, lda,
                             sA, batch, stream));
    }
  }
  return kOk;
}

int getrs_batched(int n, int nrhs, const double* LU, int lda, long long sA, const int* ipiv,
                  long long sP, const double* B, int ldb, long long sB, double* X, int ldx,
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
  hipLaunchKernelGGL(k_gather_rows, dim3(ceil_div(nrhs, 256), n, batch), dim3(256), 0, stream, n,
                     nrhs, perm, B, ldb, sB, X, ldx, sX);
  SSLAB_HIP_CHECK(hipGetLastError());
  constexpr int NB = kTrsmNB;
  // Forward: L Y = P B.
  for (int kb = 0; kb < n; kb += NB) {
    const int nr = (n - kb) < NB ? (n - kb) : NB;
    hipLaunchKernelGGL((k_trsm_cols<NB, true>), dim3(ceil_div(nrhs, 256), batch), dim3(256), 0,
                       stream, nr, nrhs, LU + (long long)kb * lda + kb, lda, sA,
                       X + (long long)kb * ldx, ldx, sX);
    SSLAB_HIP_CHECK(hipGetLastError());
    const int below = n - kb - nr;
    if (below > 0) {
      SSLAB_TRY(gemm_strided(0, 0, below, nrhs, nr, -1.0, LU + (long long)(kb + nr) * lda + kb, lda,
                             sA, X + (long long)kb * ldx, ldx, sX, 1.0,
                             X + (long long)(kb + nr) * ldx, ldx, sX, batch, stream));
    }
  }
  // Backward: U X = Y.
  const int nblk = ceil_div(n, NB);
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int kb = bi * NB;
    const int nr = (n - kb) < NB ? (n - kb) : NB;
    hipLaunchKernelGGL((k_trsm_cols<NB, false>), dim3(ceil_div(nrhs, 256), batch), dim3(256), 0,
                       stream, nr, nrhs, LU + (long long)kb * lda + kb, lda, sA,
                       X + (long long)kb * ldx, ldx, sX);
    SSLAB_HIP_CHECK(hipGetLastError());
    if (kb > 0) {
      SSLAB_TRY(gemm_strided(0, 0, kb, nrhs, nr, -1.0, LU + kb, lda, sA, X + (long long)kb * ldx,
                             ldx, sX, 1.0, X, ldx, sX, batch, stream));
    }
  }
  return kOk;
}

int trsm_upper_batched(int n, int nrhs, const double* U, int ldu, long long sU, double* X, int ldx,
                       long long sX, int batch, hipStream_t stream) {
  SSLAB_REQUIRE(n >= 0 && nrhs >= 0 && batch >= 0, "trsm: negative size");
  SSLAB_REQUIRE(ldu >= n && ldx >= nrhs, "trsm: bad leading dimension");
  if (n == 0 || nrhs == 0 || batch == 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535, "trsm: batch too large");
  constexpr int NB = kTrsmNB;
  const int nblk = ceil_div(n, NB);
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int kb = bi * NB;
    const int nr = (n - kb) < NB ? (n - kb) : NB;
    hipLaunchKernelGGL((k_trsm_cols<NB, false>), dim3(ceil_div(nrhs, 256), batch), dim3(256), 0,
                       stream, nr, nrhs, U + (long long)kb * ldu + kb, ldu, sU,
                       X + (long long)kb * ldx, ldx, sX);
    SSLAB_HIP_CHECK(hipGetLastError());
    if (kb > 0) {
      SSLAB_TRY(gemm_strided(0, 0, kb, nrhs, nr, -1.0, U + kb, ldu, sU, X + (long long)kb * ldx,
                             ldx, sX, 1.0, X, ldx, sX, batch, stream));
    }
  }
  return kOk;
}

int set_identity_batched(int n, double* X, int ldx, long long sX, int batch, hipStream_t stream) {
  if (n <= 0 || batch <= 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535 && n <= 65535, "set_identity: size too large");
  hipLaunchKernelGGL(k_set_identity, dim3(ceil_div(n, 256), n, batch), dim3(256), 0, stream, n, X,
                     ldx, sX);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int copy_block_batched(int m, int n, const double* X, int ldx, long long sX, double* Y, int ldy,
                       long long sY, int batch, hipStream_t stream) {
  if (m <= 0 || n <= 0 || batch <= 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535 && m <= 65535, "copy_block: size too large");
  hipLaunchKernelGGL(k_copy_block, dim3(ceil_div(n, 256), m, batch), dim3(256), 0, stream, n, X,
                     ldx, sX, Y, ldy, sY);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int zero_block_batched(int m, int n, double* Y, int ldy, long long sY, int batch,
                       hipStream_t stream) {
  if (m <= 0 || n <= 0 || batch <= 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535, "zero_block: batch too large");
  const int gy = m < 4096 ? m : 4096;
  hipLaunchKernelGGL(k_zero_block, dim3(ceil_div(n, 256), gy, batch), dim3(256), 0, stream, m, n,
                     Y, ldy, sY);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace sslab
