// Column-batched vector kernels of the device GMRES (FP64).  A panel W (n x r, leading
// dimension ldw) holds r independent vectors in its columns; a basis V is k such panels
// at stride sV (elements).  Every column c has its own Krylov space, so the products are
// column-wise:
//   col_dots:  H[j][c] = sum_i V_j[i][c] W[i][c]              (j < k)
//   col_axpy:  W[i][c] += alpha * sum_j V_j[i][c] H[j][c]
//   col_scale: Out[i][c] = W[i][c] * s[c]
// All are HBM-bound (8 bytes per element read per basis vector).
#include "krylov_f64.hpp"

namespace sslab {

namespace {

constexpr int kColThreads = 256;
constexpr int kRowsPerBlock = 256;

__global__ void k_fill_f64(long long n, double* __restrict__ x, double v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// Block (x: row tile, y: basis vector j, z: column chunk of 64): thread t sums column
// c = z*64 + t%64 over the rows t/64, t/64 + 4, ... of its tile; the four partial sums of
// a column meet in LDS and one atomic add per column goes to H[j][c].
__global__ void __launch_bounds__(kColThreads)
k_col_dots(int n, int r, const double* __restrict__ V, long long sV, int ldv,
           const double* __restrict__ W, int ldw, double* __restrict__ H, int ldh) {
  __shared__ double part[kColThreads];
  const int t = threadIdx.x;
  const int c = blockIdx.z * 64 + (t & 63);
  const int g = t >> 6;
  const int j = blockIdx.y;
  const int i0 = blockIdx.x * kRowsPerBlock;
  const int i1 = i0 + kRowsPerBlock < n ? i0 + kRowsPerBlock : n;
  double s = 0.0;
  if (c < r) {
    const double* v = V + (long long)j * sV + c;
    const double* w = W + c;
    for (int i = i0 + g; i < i1; i += 4) s += v[(long long)i * ldv] * w[(long long)i * ldw];
  }
  part[t] = s;
  __syncthreads();
  if (g == 0 && c < r) {
    const double tot = part[t] + part[t + 64] + part[t + 128] + part[t + 192];
    atomicAdd(&H[(long long)j * ldh + c], tot);
  }
}

__global__ void k_col_axpy(int n, int r, int k, double alpha, const double* __restrict__ V,
                           long long sV, int ldv, const double* __restrict__ H, int ldh,
                           double* __restrict__ W, int ldw) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)n * r) return;
  const int i = (int)(e / r);
  const int c = (int)(e - (long long)i * r);
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += V[(long long)j * sV + (long long)i * ldv + c] * H[(long long)j * ldh + c];
  W[(long long)i * ldw + c] += alpha * s;
}

__global__ void k_col_scale(int n, int r, const double* __restrict__ W, int ldw,
                            const double* __restrict__ sc, double* __restrict__ Out, int ldo) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)n * r) return;
  const int i = (int)(e / r);
  const int c = (int)(e - (long long)i * r);
  Out[(long long)i * ldo + c] = W[(long long)i * ldw + c] * sc[c];
}

unsigned grid1(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int col_dots(int n, int r, int k, const double* V, long long sV, int ldv, const double* W,
             int ldw, double* H, int ldh, hipStream_t st) {
  SSLAB_REQUIRE(n >= 0 && r >= 0 && k >= 0, "col_dots: negative size");
  if (r == 0 || k == 0) return kOk;
  SSLAB_REQUIRE(V && W && H && ldv >= r && ldw >= r && ldh >= r && k <= 65535,
                "col_dots: bad pointers / strides");
  hipLaunchKernelGGL(k_fill_f64, dim3(grid1((long long)k * ldh)), dim3(256), 0, st,
                     (long long)k * ldh, H, 0.0);
  SSLAB_HIP_CHECK(hipGetLastError());
  if (n == 0) return kOk;
  const dim3 grid((unsigned)((n + kRowsPerBlock - 1) / kRowsPerBlock), (unsigned)k,
                  (unsigned)((r + 63) / 64));
  hipLaunchKernelGGL(k_col_dots, grid, dim3(kColThreads), 0, st, n, r, V, sV, ldv, W, ldw, H,
                     ldh);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int col_axpy(int n, int r, int k, double alpha, const double* V, long long sV, int ldv,
             const double* H, int ldh, double* W, int ldw, hipStream_t st) {
  SSLAB_REQUIRE(n >= 0 && r >= 0 && k >= 0, "col_axpy: negative size");
  if (n == 0 || r == 0 || k == 0) return kOk;
  SSLAB_REQUIRE(V && H && W && ldv >= r && ldw >= r && ldh >= r, "col_axpy: bad pointers / strides");
  hipLaunchKernelGGL(k_col_axpy, dim3(grid1((long long)n * r)), dim3(256), 0, st, n, r, k, alpha,
                     V, sV, ldv, H, ldh, W, ldw);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int col_scale(int n, int r, const double* W, int ldw, const double* s, double* Out, int ldo,
              hipStream_t st) {
  SSLAB_REQUIRE(n >= 0 && r >= 0, "col_scale: negative size");
  if (n == 0 || r == 0) return kOk;
  SSLAB_REQUIRE(W && s && Out && ldw >= r && ldo >= r, "col_scale: bad pointers / strides");
  hipLaunchKernelGGL(k_col_scale, dim3(grid1((long long)n * r)), dim3(256), 0, st, n, r, W, ldw,
                     s, Out, ldo);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

}  // namespace sslab
