// Batched FP64 Householder QR (optionally column-pivoted) for gfx950.
//
// One workgroup (16 wavefronts) factors one small matrix held in global memory (the
// HBS boxes are at most a few hundred rows, so a matrix stays L2-resident): per
// column j, an optional pivot search over the remaining column norms, the
// reflector H_j = I - tau v v^T (LAPACK dlarfg convention, v_0 = 1) built in LDS,
// its application to the trailing columns (one wavefront per column, wave
// reductions for the dot products) and, when requested, to the explicit complete
// Q (m x m) from the right, Q <- Q H_j, one wavefront per row.
// On exit A holds R (zeros below the diagonal), Q the orthogonal factor with
// A_in[:, perm] = Q R.
#include "qr_f64.hpp"

#include <cstdlib>

#include "wave_ops.hpp"

namespace sslab {

namespace {

constexpr int kQrThreads = 1024;  // 16 wavefronts share the trailing columns
constexpr int kQrWaves = kQrThreads / 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ void __launch_bounds__(kQrThreads)
k_geqrf(int m, int n, double* __restrict__ A, int lda, long long sA, double* __restrict__ Q,
        int ldq, long long sQ, int* __restrict__ perm, long long sPerm) {
  __shared__ double v[kQrMaxRows];
  __shared__ double cn[kQrMaxCols];
  __shared__ double s_tau, s_beta, s_scale;
  __shared__ int s_piv;
  const int b = blockIdx.x;
  double* Ab = A + (long long)b * sA;
  double* Qb = Q ? Q + (long long)b * sQ : nullptr;
  int* pb = perm ? perm + (long long)b * sPerm : nullptr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  if (Qb)
    for (long long e = tid; e < (long long)m * m; e += kQrThreads) {
      const int i = (int)(e / m), j = (int)(e % m);
      Qb[(long long)i * ldq + j] = (i == j) ? 1.0 : 0.0;
    }
  if (pb)
    for (int c = tid; c < n; c += kQrThreads) pb[c] = c;
  __syncthreads();

  const int kmax = m < n ? m : n;
  for (int j = 0; j < kmax; ++j) {
    if (pb) {
      for (int c = j + wave; c < n; c += kQrWaves) {
        double s = 0.0;
        for (int i = j + lane; i < m; i += 64) {
          const double a = Ab[(long long)i * lda + c];
          s += a * a;
        }
        s = wave_sum(s);
        if (lane == 0) cn[c] = s;
      }
      __syncthreads();
      if (wave == 0) {
        // first column of largest norm (the serial scan's choice): each lane scans its
        // columns in order, then the wave takes the maximum and the smallest index at it
        double best = -1.0;
        int bi = 0x7fffffff;
        for (int c = j + lane; c < n; c += 64) {
          const double x = cn[c];
          if (x > best) {
            best = x;
            bi = c;
          }
        }
        const double mx = wave::wave_max(best);
        int p = wave::wave_min(best == mx ? bi : 0x7fffffff);
        if (p >= n) p = j;  // no finite norm: keep the column
        if (lane == 0) {
          s_piv = p;
          if (p != j) {
            const int t = pb[j];
            pb[j] = pb[p];
            pb[p] = t;
          }
        }
      }
      __syncthreads();
      const int p = s_piv;
      if (p != j)
        for (int i = tid; i < m; i += kQrThreads) {
          const double t = Ab[(long long)i * lda + j];
          Ab[(long long)i * lda + j] = Ab[(long long)i * lda + p];
          Ab[(long long)i * lda + p] = t;
        }
      __syncthreads();
    }
    // Reflector for x = A[j:, j]  (dlarfg): beta = -sign(alpha) ||x||, tau = (beta-alpha)/beta.
    if (wave == 0) {
      double s = 0.0;
      for (int i = j + 1 + lane; i < m; i += 64) {
        const double a = Ab[(long long)i * lda + j];
        s += a * a;
      }
      s = wave_sum(s);
      if (lane == 0) {
        const double alpha = Ab[(long long)j * lda + j];
        if (s == 0.0) {
          s_tau = 0.0;
          s_beta = alpha;
          s_scale = 0.0;
        } else {
          const double nrm = sqrt(alpha * alpha + s);
          const double beta = alpha >= 0.0 ? -nrm : nrm;
          s_tau = (beta - alpha) / beta;
          s_beta = beta;
          s_scale = 1.0 / (alpha - beta);
        }
      }
    }
    __syncthreads();
    const double tau = s_tau;
    for (int i = j + tid; i < m; i += kQrThreads)
      v[i] = (i == j) ? 1.0 : Ab[(long long)i * lda + j] * s_scale;
    __syncthreads();
    for (int i = j + tid; i < m; i += kQrThreads)
      Ab[(long long)i * lda + j] = (i == j) ? s_beta : 0.0;
    if (tau != 0.0) {
      for (int c = j + 1 + wave; c < n; c += kQrWaves) {
        double w = 0.0;
        for (int i = j + lane; i < m; i += 64) w += v[i] * Ab[(long long)i * lda + c];
        w = wave_sum(w) * tau;
        for (int i = j + lane; i < m; i += 64) Ab[(long long)i * lda + c] -= w * v[i];
      }
      if (Qb)
        for (int r = wave; r < m; r += kQrWaves) {
          double w = 0.0;
          for (int i = j + lane; i < m; i += 64) w += Qb[(long long)r * ldq + i] * v[i];
          w = wave_sum(w) * tau;
          for (int i = j + lane; i < m; i += 64) Qb[(long long)r * ldq + i] -= w * v[i];
        }
    }
    __syncthreads();
  }
}

// The same factorization with the matrix staged in LDS, column-major (every column walk
// -- norms, reflector products, updates -- reads consecutive LDS words instead of global
// rows lda apart), for matrices whose m*n + m + n doubles fit; 16 wavefronts share the
// trailing columns (and the rows of Q).  The complete Q stays in global memory, updated
// row by row (coalesced).
constexpr int kQrLdsThreads = 1024;
constexpr int kQrLdsWaves = kQrLdsThreads / 64;

__global__ void __launch_bounds__(kQrLdsThreads)
k_geqrf_lds(int m, int n, double* __restrict__ A, int lda, long long sA, double* __restrict__ Q,
            int ldq, long long sQ, int* __restrict__ perm, long long sPerm) {
  extern __shared__ __attribute__((aligned(16))) unsigned char qr_smem[];
  double* L = reinterpret_cast<double*>(qr_smem);  // L[c * m + i] = A[i][c]
  double* v = L + (long long)m * n;
  double* cn = v + m;
  __shared__ double s_tau, s_beta, s_scale;
  __shared__ int s_piv;
  const int b = blockIdx.x;
  double* Ab = A + (long long)b * sA;
  double* Qb = Q ? Q + (long long)b * sQ : nullptr;
  int* pb = perm ? perm + (long long)b * sPerm : nullptr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int e = tid; e < m * n; e += kQrLdsThreads) {
    const int i = e / n, c = e - (e / n) * n;
    L[(long long)c * m + i] = Ab[(long long)i * lda + c];
  }
  if (Qb)
    for (long long e = tid; e < (long long)m * m; e += kQrLdsThreads) {
      const int i = (int)(e / m), j = (int)(e % m);
      Qb[(long long)i * ldq + j] = (i == j) ? 1.0 : 0.0;
    }
  if (pb)
    for (int c = tid; c < n; c += kQrLdsThreads) pb[c] = c;
  __syncthreads();

  const int kmax = m < n ? m : n;
  for (int j = 0; j < kmax; ++j) {
    if (pb) {
      for (int c = j + wave; c < n; c += kQrLdsWaves) {
        const double* col = L + (long long)c * m;
        double s = 0.0;
        for (int i = j + lane; i < m; i += 64) s += col[i] * col[i];
        s = wave_sum(s);
        if (lane == 0) cn[c] = s;
      }
      __syncthreads();
      if (wave == 0) {
        // first column of largest norm (the serial scan's choice): each lane scans its
        // columns in order, then the wave takes the maximum and the smallest index at it
        double best = -1.0;
        int bi = 0x7fffffff;
        for (int c = j + lane; c < n; c += 64) {
          const double x = cn[c];
          if (x > best) {
            best = x;
            bi = c;
          }
        }
        const double mx = wave::wave_max(best);
        int p = wave::wave_min(best == mx ? bi : 0x7fffffff);
        if (p >= n) p = j;  // no finite norm: keep the column
        if (lane == 0) {
          s_piv = p;
          if (p != j) {
            const int t = pb[j];
            pb[j] = pb[p];
            pb[p] = t;
          }
        }
      }
      __syncthreads();
      const int p = s_piv;
      if (p != j) {
        double* cj = L + (long long)j * m;
        double* cp = L + (long long)p * m;
        for (int i = tid; i < m; i += kQrLdsThreads) {
          const double t = cj[i];
          cj[i] = cp[i];
          cp[i] = t;
        }
      }
      __syncthreads();
    }
    double* colj = L + (long long)j * m;
    if (wave == 0) {
      double s = 0.0;
      for (int i = j + 1 + lane; i < m; i += 64) s += colj[i] * colj[i];
      s = wave_sum(s);
      if (lane == 0) {
        const double alpha = colj[j];
        if (s == 0.0) {
          s_tau = 0.0;
          s_beta = alpha;
          s_scale = 0.0;
        } else {
          const double nrm = sqrt(alpha * alpha + s);
          const double beta = alpha >= 0.0 ? -nrm : nrm;
          s_tau = (beta - alpha) / beta;
          s_beta = beta;
          s_scale = 1.0 / (alpha - beta);
        }
      }
    }
    __syncthreads();
    const double tau = s_tau;
    for (int i = j + tid; i < m; i += kQrLdsThreads) v[i] = (i == j) ? 1.0 : colj[i] * s_scale;
    __syncthreads();
    for (int i = j + tid; i < m; i += kQrLdsThreads) colj[i] = (i == j) ? s_beta : 0.0;
    if (tau != 0.0) {
      for (int c = j + 1 + wave; c < n; c += kQrLdsWaves) {
        double* col = L + (long long)c * m;
        double w = 0.0;
        for (int i = j + lane; i < m; i += 64) w += v[i] * col[i];
        w = wave_sum(w) * tau;
        for (int i = j + lane; i < m; i += 64) col[i] -= w * v[i];
      }
      if (Qb)
        for (int r = wave; r < m; r += kQrLdsWaves) {
          double w = 0.0;
          for (int i = j + lane; i < m; i += 64) w += Qb[(long long)r * ldq + i] * v[i];
          w = wave_sum(w) * tau;
          for (int i = j + lane; i < m; i += 64) Qb[(long long)r * ldq + i] -= w * v[i];
        }
    }
    __syncthreads();
  }
  for (int e = tid; e < m * n; e += kQrLdsThreads) {
    const int i = e / n, c = e - (e / n) * n;
    Ab[(long long)i * lda + c] = L[(long long)c * m + i];
  }
}

// The LDS-staged factorization with the complete Q held in registers instead of global
// memory, so applying a reflector to Q needs no memory traffic and Q is stored once at the
// end.  The Q tile is grouped by 16-lane rows: a wave holds four Q rows side by side (one
// per DPP row of 16 lanes), so the row dot products of the per-column update
// Q <- Q (I - tau v v^T) end in one 16-lane reduction each.  For m <= min(16 RW, 64 CI)
// and matrices whose LDS panel fits; same results as k_geqrf_lds up to the summation order.
// RW must be a multiple of 4.  Up to 128 columns a lane keeps its 4 CI values of v in
// registers; beyond, it reads them from LDS twice per column (the 128-register budget of a
// 1024-thread workgroup holds the Q tile but not that slice as well).
template <int RW, int CI>
__global__ void __launch_bounds__(kQrLdsThreads)
k_geqrf_lds_qrow(int m, int n, double* __restrict__ A, int lda, long long sA,
                 double* __restrict__ Q, int ldq, long long sQ, int* __restrict__ perm,
                 long long sPerm) {
  extern __shared__ __attribute__((aligned(16))) unsigned char qr_smem[];
  double* L = reinterpret_cast<double*>(qr_smem);  // L[c * m + i] = A[i][c]
  double* v = L + (long long)m * n;
  double* cn = v + m;
  __shared__ double s_tau, s_beta, s_scale;
  __shared__ int s_piv;
  const int b = blockIdx.x;
  double* Ab = A + (long long)b * sA;
  double* Qb = Q + (long long)b * sQ;
  int* pb = perm ? perm + (long long)b * sPerm : nullptr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int e = tid; e < m * n; e += kQrLdsThreads) {
    const int i = e / n, c = e - (e / n) * n;
    L[(long long)c * m + i] = Ab[(long long)i * lda + c];
  }
  if (pb)
    for (int c = tid; c < n; c += kQrLdsThreads) pb[c] = c;
  // Q tile: lane row s = lane / 16 of wave w holds Q rows w + 16 (4 g + s) (g < RW / 4) at
  // columns (lane % 16) + 16 t (t < 4 CI)
  constexpr int G = RW / 4, C4 = 4 * CI;
  const int s16 = lane >> 4, l16 = lane & 15;
  double q[G][C4];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int t = 0; t < C4; ++t) q[g][t] = (wave + 16 * (4 * g + s16) == l16 + 16 * t) ? 1.0 : 0.0;
  __syncthreads();

  const int kmax = m < n ? m : n;
  for (int j = 0; j < kmax; ++j) {
    if (pb) {
      // squared norms of the remaining columns, four per wave at a time (16 lanes each)
      for (int c0 = j + 4 * wave; c0 < n; c0 += 4 * kQrLdsWaves) {  // wave-uniform
        const int c = c0 + s16;
        const double* col = L + (long long)(c < n ? c : j) * m;
        double s = 0.0;
        if (c < n)
          for (int i = j + l16; i < m; i += 16) s += col[i] * col[i];
        s = wave::row16_sum(s);
        if (c < n && l16 == 0) cn[c] = s;
      }
      __syncthreads();
      if (wave == 0) {
        // first column of largest norm (the serial scan's choice): each lane scans its
        // columns in order, then the wave takes the maximum and the smallest index at it
        double best = -1.0;
        int bi = 0x7fffffff;
        for (int c = j + lane; c < n; c += 64) {
          const double x = cn[c];
          if (x > best) {
            best = x;
            bi = c;
          }
        }
        const double mx = wave::wave_max(best);
        int p = wave::wave_min(best == mx ? bi : 0x7fffffff);
        if (p >= n) p = j;  // no finite norm: keep the column
        if (lane == 0) {
          s_piv = p;
          if (p != j) {
            const int t = pb[j];
            pb[j] = pb[p];
            pb[p] = t;
          }
        }
      }
      __syncthreads();
      const int p = s_piv;
      if (p != j) {
        double* cj = L + (long long)j * m;
        double* cp = L + (long long)p * m;
        for (int i = tid; i < m; i += kQrLdsThreads) {
          const double t = cj[i];
          cj[i] = cp[i];
          cp[i] = t;
        }
      }
      __syncthreads();
    }
    double* colj = L + (long long)j * m;
    if (wave == 0) {
      double s = 0.0;
      for (int i = j + 1 + lane; i < m; i += 64) s += colj[i] * colj[i];
      s = wave::wave_sum(s);
      if (lane == 0) {
        const double alpha = colj[j];
        if (s == 0.0) {
          s_tau = 0.0;
          s_beta = alpha;
          s_scale = 0.0;
        } else {
          const double nrm = sqrt(alpha * alpha + s);
          const double beta = alpha >= 0.0 ? -nrm : nrm;
          s_tau = (beta - alpha) / beta;
          s_beta = beta;
          s_scale = 1.0 / (alpha - beta);
        }
      }
    }
    __syncthreads();
    const double tau = s_tau;
    for (int i = j + tid; i < m; i += kQrLdsThreads) v[i] = (i == j) ? 1.0 : colj[i] * s_scale;
    __syncthreads();
    for (int i = j + tid; i < m; i += kQrLdsThreads) colj[i] = (i == j) ? s_beta : 0.0;
    if (tau != 0.0) {
      // four trailing columns per wave at a time, 16 lanes per column: the dot products
      // end in one 16-lane DPP row reduction each (no cross-row readlanes)
      const int sub = lane >> 4, sl = lane & 15;
      for (int c0 = j + 1 + 4 * wave; c0 < n; c0 += 4 * kQrLdsWaves) {  // wave-uniform
        const int c = c0 + sub;
        double* col = L + (long long)(c < n ? c : j) * m;
        double w = 0.0;
        if (c < n)
          for (int i = j + sl; i < m; i += 16) w += v[i] * col[i];
        w = wave::row16_sum(w) * tau;
        if (c < n)
          for (int i = j + sl; i < m; i += 16) col[i] -= w * v[i];
      }
      // Q <- Q (I - tau v v^T) on the register tile (v is zero above row j)
      if constexpr (C4 <= 8) {  // the lane's v slice stays in registers
        double vv[C4];
#pragma unroll
        for (int t = 0; t < C4; ++t) {
          const int i = l16 + 16 * t;
          vv[t] = (i >= j && i < m) ? v[i] : 0.0;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          double part = 0.0;
#pragma unroll
          for (int t = 0; t < C4; ++t) part += q[g][t] * vv[t];
          const double w = wave::row16_sum(part) * tau;  // one Q row per 16-lane DPP row
#pragma unroll
          for (int t = 0; t < C4; ++t) q[g][t] -= w * vv[t];
        }
      } else {  // v read from LDS once for the dots and once for the update (register budget)
        double part[G];
#pragma unroll
        for (int g = 0; g < G; ++g) part[g] = 0.0;
#pragma unroll
        for (int t = 0; t < C4; ++t) {
          const int i = l16 + 16 * t;
          const double vt = (i >= j && i < m) ? v[i] : 0.0;
#pragma unroll
          for (int g = 0; g < G; ++g) part[g] += q[g][t] * vt;
        }
        double w[G];
#pragma unroll
        for (int g = 0; g < G; ++g) w[g] = wave::row16_sum(part[g]) * tau;
        __asm__ volatile("" ::: "memory");  // reload v below instead of keeping 4 CI values live
#pragma unroll
        for (int t = 0; t < C4; ++t) {
          const int i = l16 + 16 * t;
          const double vt = (i >= j && i < m) ? v[i] : 0.0;
#pragma unroll
          for (int g = 0; g < G; ++g) q[g][t] -= w[g] * vt;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int r = wave + 16 * (4 * g + s16);
#pragma unroll
    for (int t = 0; t < C4; ++t) {
      const int i = l16 + 16 * t;
      if (r < m && i < m) Qb[(long long)r * ldq + i] = q[g][t];
    }
  }
  for (int e = tid; e < m * n; e += kQrLdsThreads) {
    const int i = e / n, c = e - (e / n) * n;
    Ab[(long long)i * lda + c] = L[(long long)c * m + i];
  }
}

// Y[b][blk*bs + i][blk*bs + j] += X[b*nblk + blk][i][j]: add nblk diagonal bs x bs blocks.
__global__ void k_add_diag_blocks(int bs, int nblk, const double* __restrict__ X,
                                  double* __restrict__ Y, int ldy, long long sY) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  const long long c = blockIdx.z;  // child index
  if (j >= bs) return;
  const long long b = c / nblk, blk = c % nblk;
  Y[b * sY + (blk * bs + i) * ldy + blk * bs + j] += X[c * bs * bs + (long long)i * bs + j];
}

}  // namespace

int geqrf_batched(int m, int n, double* A, int lda, long long sA, double* Q, int ldq, long long sQ,
                  int* perm, long long sPerm, int batch, hipStream_t st) {
  SSLAB_REQUIRE(m >= 0 && n >= 0 && batch >= 0, "geqrf: negative size");
  if (m == 0 || n == 0 || batch == 0) return kOk;
  SSLAB_REQUIRE(m <= kQrMaxRows && n <= kQrMaxCols,
                "geqrf: at most %d rows and %d columns (got %d x %d)", kQrMaxRows, kQrMaxCols, m,
                n);
  SSLAB_REQUIRE(lda >= n && (Q == nullptr || ldq >= m), "geqrf: bad leading dimension");
  SSLAB_REQUIRE(batch <= 1 << 30, "geqrf: batch too large");
  const size_t lds = sizeof(double) * ((size_t)m * n + m + n);
  static const bool use_lds = [] {
    const char* e = getenv("SSLAB_QR_LDS");
    return !(e && e[0] == '0');
  }();
  if (use_lds && lds <= 150 * 1024) {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_geqrf_lds),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
      attr = true;
    }
    static const bool qreg = [] {
      const char* e = getenv("SSLAB_QR_QREG");
      return !(e && e[0] == '0');
    }();
    if (qreg && Q != nullptr && m <= 192) {
      // the complete Q in registers: ceil(m / 16) rows per wave, ceil(m / 64) columns per lane
#define SSLAB_QR_QREG_LAUNCH(KERNEL, RW, CI)                                                  \
  do {                                                                                          \
    static bool attr_q = false;                                                                 \
    if (!attr_q) {                                                                              \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&KERNEL<RW, CI>),                \
                                hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);       \
      attr_q = true;                                                                            \
    }                                                                                           \
    hipLaunchKernelGGL((KERNEL<RW, CI>), dim3(batch), dim3(kQrLdsThreads), lds, st, m, n, A,    \
                       lda, sA, Q, ldq, sQ, perm, sPerm);                                       \
  } while (0)
      if (m <= 64) SSLAB_QR_QREG_LAUNCH(k_geqrf_lds_qrow, 4, 1);
      else if (m <= 128) SSLAB_QR_QREG_LAUNCH(k_geqrf_lds_qrow, 8, 2);
      else SSLAB_QR_QREG_LAUNCH(k_geqrf_lds_qrow, 12, 3);
#undef SSLAB_QR_QREG_LAUNCH
      SSLAB_HIP_CHECK(hipGetLastError());
      return kOk;
    }
    hipLaunchKernelGGL(k_geqrf_lds, dim3(batch), dim3(kQrLdsThreads), lds, st, m, n, A, lda, sA,
                       Q, ldq, sQ, perm, sPerm);
    SSLAB_HIP_CHECK(hipGetLastError());
    return kOk;
  }
  hipLaunchKernelGGL(k_geqrf, dim3(batch), dim3(kQrThreads), 0, st, m, n, A, lda, sA, Q, ldq, sQ,
                     perm, sPerm);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int add_diag_blocks(int bs, int nblk, int nparent, const double* X, double* Y, int ldy,
                    long long sY, hipStream_t st) {
  if (bs <= 0 || nblk <= 0 || nparent <= 0) return kOk;
  SSLAB_REQUIRE((long long)nblk * nparent <= 65535 && bs <= 65535,
                "add_diag_blocks: too many blocks");
  hipLaunchKernelGGL(k_add_diag_blocks, dim3(ceil_div(bs, 256), bs, nblk * nparent), dim3(256), 0,
                     st, bs, nblk, X, Y, ldy, sY);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace sslab
