// Spectral double-slab discretization (2-D, variable coefficients) on gfx950 (FP64).
//
// Every double slab is one qx x qy tensor Chebyshev grid (the reference's spectralSolver,
// solver/spectral/spectralSolver.py:161-173, 198-235: ord + 1 points per axis).  The
// collocation operator at interior node (a, b) (1 <= a <= qx-2 along x, 1 <= b <= qy-2)
//   A[u] = -c11 u_xx - c22 u_yy + c1 u_x + c2 u_y + c0 u          (pdo/pdo.py:19-39)
// touches only its grid line in x and in y, so A_ii ((qx-2)(qy-2) square, row
// r = (a-1)(qy-2) + (b-1)) and A_ib (boundary columns [left (qy-2) | right (qy-2) |
// bottom (qx-2) | top (qx-2)], corners dropped: no interior equation reads them) are
// assembled directly.  The slab's maps Y = (A_ii^{-1} A_ib)[centre line] are the rows of
// interior column ac = (qx-1)/2 (qx odd): one contiguous m x nbdry block, m = qy-2.
#include "spectral_f64.hpp"

#include "gemm.hpp"
#include "lu_f64.hpp"

namespace sslab {

namespace {

__global__ void k_rect_assemble(int qx, int qy, const double* __restrict__ Dx,
                                const double* __restrict__ Dxx, const double* __restrict__ Dy,
                                const double* __restrict__ Dyy, const double* __restrict__ c11,
                                const double* __restrict__ c22, const double* __restrict__ c1,
                                const double* __restrict__ c2, const double* __restrict__ c0,
                                double* __restrict__ Aii, double* __restrict__ Aib) {
  const int px = qx - 2, py = qy - 2;
  const int ni = px * py, nbd = 2 * py + 2 * px;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  const long long s = blockIdx.z;
  if (col >= ni + nbd) return;
  const int a = r / py + 1, b = r % py + 1;
  const long long ci = s * ni + r;
  const double k11 = c11 ? c11[ci] : 0.0;
  const double k22 = c22 ? c22[ci] : 0.0;
  const double k1 = c1 ? c1[ci] : 0.0;
  const double k2 = c2 ? c2[ci] : 0.0;
  const double k0 = c0 ? c0[ci] : 0.0;
  int a2, b2;
  if (col < ni) {
    a2 = col / py + 1;
    b2 = col % py + 1;
  } else {
    const int e = col - ni;
    if (e < py) { a2 = 0; b2 = e + 1; }
    else if (e < 2 * py) { a2 = qx - 1; b2 = e - py + 1; }
    else if (e < 2 * py + px) { a2 = e - 2 * py + 1; b2 = 0; }
    else { a2 = e - 2 * py - px + 1; b2 = qy - 1; }
  }
  double v = 0.0;
  if (b2 == b) v += -k11 * Dxx[a * qx + a2] + k1 * Dx[a * qx + a2];
  if (a2 == a) v += -k22 * Dyy[b * qy + b2] + k2 * Dy[b * qy + b2];
  if (a2 == a && b2 == b) v += k0;
  if (col < ni)
    Aii[s * ni * ni + (long long)r * ni + col] = v;
  else
    Aib[s * ni * nbd + (long long)r * nbd + (col - ni)] = v;
}

}  // namespace

int spectral2d_sizes(int nslab, int qx, int qy, long long* work_doubles, long long* work_ints) {
  SSLAB_REQUIRE(nslab >= 1 && qx >= 3 && qy >= 3 && (qx & 1) == 1,
                "spectral2d_sizes: nslab >= 1, qy >= 3 and an odd qx >= 3 required");
  const long long ni = (long long)(qx - 2) * (qy - 2), nbd = 2LL * (qy - 2) + 2LL * (qx - 2);
  if (work_doubles) *work_doubles = (long long)nslab * (ni * ni + 2 * ni * nbd);
  if (work_ints) *work_ints = (long long)nslab * 2 * ni + nslab;
  return kOk;
}

int spectral2d_slab_maps(int nslab, int qx, int qy, const double* Dx, const double* Dxx,
                         const double* Dy, const double* Dyy, const double* c11,
                         const double* c22, const double* c1, const double* c2, const double* c0,
                         double* Y, double* work, int* iwork, int* info, hipStream_t st) {
  SSLAB_REQUIRE(nslab >= 1 && qx >= 3 && qy >= 3 && (qx & 1) == 1,
                "spectral2d_slab_maps: nslab >= 1, qy >= 3 and an odd qx >= 3 required");
  SSLAB_REQUIRE(nslab <= 65535, "spectral2d_slab_maps: at most 65535 slabs per call");
  SSLAB_REQUIRE(Dx && Dxx && Dy && Dyy && Y && work && iwork, "spectral2d_slab_maps: null pointer");
  const int px = qx - 2, py = qy - 2;
  const int ni = px * py, nbd = 2 * py + 2 * px;
  SSLAB_REQUIRE(ni <= 65535, "spectral2d_slab_maps: at most 65535 interior nodes per slab");
  const long long ni2 = (long long)ni * ni, nib = (long long)ni * nbd;
  double* Aii = work;
  double* Aib = Aii + (long long)nslab * ni2;
  double* X = Aib + (long long)nslab * nib;
  int* ipiv = iwork;
  int* perm = iwork + (long long)nslab * ni;
  int* linfo = info ? info : perm + (long long)nslab * ni;
  hipLaunchKernelGGL(k_rect_assemble, dim3(ceil_div(ni + nbd, 256), ni, nslab), dim3(256), 0, st,
                     qx, qy, Dx, Dxx, Dy, Dyy, c11, c22, c1, c2, c0, Aii, Aib);
  SSLAB_HIP_CHECK(hipGetLastError());
  SSLAB_TRY(getrf_batched(ni, Aii, ni, ni2, ipiv, ni, linfo, nslab, st));
  SSLAB_TRY(getrs_batched(ni, nbd, Aii, ni, ni2, ipiv, ni, Aib, nbd, nib, X, nbd, nib, perm,
                          nslab, st));
  // the centre line's rows: interior column ac = (qx-1)/2, rows (ac-1) py .. ac py - 1
  const int ac = (qx - 1) / 2;
  return copy_block_batched(py, nbd, X + (long long)(ac - 1) * py * nbd, nbd, nib, Y, nbd,
                            (long long)py * nbd, nslab, st);
}

}  // namespace sslab
