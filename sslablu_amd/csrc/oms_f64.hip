// S-system right-hand sides from global Dirichlet data (multislab/oms.py:180-186):
//   rhs_i = -(A_ii^{-1} A_ib[:, Igb] f)[Ic]
// with the boundary maps stored in the double-slab pair solutions
//   Y_i = [ L | R | bot_a bot_b | top_a top_b ]   (m x ncols, ncols = 2m + 4 nxp)
// and the data F ordered [bottom (N nxp) | top (N nxp) | left (m) | right (m)], so
// double slab i reads rows [i nxp, (i+2) nxp) of the bottom and top segments: one
// grouped GEMM launch with strided terms, no gather.
#include "oms_f64.hpp"

#include "gemm.hpp"
#include "lu_f64.hpp"

namespace sslab {

int oms_rhs(int nb, int m, int nxp, int ncols, const double* Y, const double* F, int ldF, int r,
            int N, double* out, int ldo, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 1 && m >= 1 && nxp >= 1 && r >= 0, "oms_rhs: bad sizes");
  SSLAB_REQUIRE(N == nb + 1, "oms_rhs: N must be nb + 1");
  SSLAB_REQUIRE(ncols == 2 * m + 4 * nxp, "oms_rhs: ncols must be 2m + 4 nx p");
  SSLAB_REQUIRE(Y && F && out && ldF >= r && ldo >= r, "oms_rhs: bad pointers / strides");
  if (r == 0) return kOk;
  const long long sY = (long long)m * ncols;
  const long long sO = (long long)m * ldo;
  const long long sF = (long long)nxp * ldF;
  const double* Fbot = F;
  const double* Ftop = F + (long long)N * nxp * ldF;
  const double* Fleft = F + 2LL * N * nxp * ldF;
  const double* Fright = Fleft + (long long)m * ldF;
  auto bottom = [&](long long i) {
    return gemm_term(Y + i * sY + 2 * m, ncols, sY, 0, Fbot + i * sF, ldF, sF, 0, 2 * nxp, -1.0);
  };
  auto top = [&](long long i) {
    return gemm_term(Y + i * sY + 2 * m + 2 * nxp, ncols, sY, 0, Ftop + i * sF, ldF, sF, 0,
                     2 * nxp, -1.0);
  };
  // Accumulate into a zeroed output so split-K launches need no init pass.
  SSLAB_TRY(zero_block_batched((int)((long long)nb * m), r, out, ldo, 0, 1, st));
  GemmLaunch G{};
  int k = 0;
  if (nb > 2) {
    GemmDesc d = gemm_desc(m, r, nb - 2, out + sO, ldo, sO, out + sO, ldo, sO, 1.0);
    gemm_add_term(d, bottom(1));
    gemm_add_term(d, top(1));
    G.d[k++] = d;
  }
  {
    GemmDesc d = gemm_desc(m, r, 1, out, ldo, sO, out, ldo, sO, 1.0);
    gemm_add_term(d, bottom(0));
    gemm_add_term(d, top(0));
    gemm_add_term(d, gemm_term(Y, ncols, sY, 0, Fleft, ldF, 0, 0, m, -1.0));
    G.d[k++] = d;
  }
  if (nb >= 2) {
    const long long i = nb - 1;
    GemmDesc d = gemm_desc(m, r, 1, out + i * sO, ldo, sO, out + i * sO, ldo, sO, 1.0);
    gemm_add_term(d, bottom(i));
    gemm_add_term(d, top(i));
    gemm_add_term(d, gemm_term(Y + i * sY + m, ncols, sY, 0, Fright, ldF, 0, 0, m, -1.0));
    G.d[k++] = d;
  }
  G.nd = k;
  SSLAB_TRY(gemm_launch(G, st));
  if (nb == 1) {  // the single double slab touches both the left and the right boundary
    SSLAB_TRY(gemm_strided(0, 0, m, r, m, -1.0, Y + m, ncols, 0, Fright, ldF, 0, 1.0, out, ldo, 0,
                           1, st));
  }
  return kOk;
}

namespace {

int zero_grid(long long n) {
  const long long g = (n + 255) / 256;
  return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

// Entry state of a red-black solve from boundary data: V's even block rows and all of X
// zeroed (grid-stride over both regions, one launch).
__global__ void k_zero_rb0(int nE, int m, long long rowsX, int r, double* __restrict__ V,
                           int ldv, double* __restrict__ X, int ldx) {
  const long long rowsV = (long long)nE * m;
  const long long n = (rowsV + rowsX) * r;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long q = e / r;
    const int c = (int)(e - q * r);
    if (q < rowsV) {
      const long long b = q / m;
      V[(2 * b * m + (q - b * m)) * ldv + c] = 0.0;
    } else {
      X[(q - rowsV) * ldx + c] = 0.0;
    }
  }
}

}  // namespace

// Right-hand sides with the first red-black elimination level applied, straight from
// the boundary data (identity diagonal blocks, so W^L_i = SiM_i, W^R_i = SiP_i):
//   odd j :  X_j = rhs_j = -(Y_j F_j)            (y_j of the first level, T_j^{-1} = I)
//   even i:  V_i = rhs_i - SiM_i rhs_{i-1} - SiP_i rhs_{i+1}
//          = Ycb_i Fbot[(i-1) nxp, (i+3) nxp) + Yct_i Ftop[...] (+ Ycl F_left for i = 0)
//            (+ Ycr F_right for the even node nearest the right boundary)
//   X's even rows = 0 (the red-black solve accumulates into them).
// Ycb_i / Yct_i (m x 4 nxp, one per even node) combine the three double slabs' boundary
// maps over the four strip windows they touch; windows outside the strip are zero
// columns and are skipped here (the first and possibly the last even node use three).
// One zero kernel, then one grouped GEMM launch whose products all accumulate in place,
// so split-K needs no initialisation pass.
int oms_rhs_rb0(int nb, int m, int nxp, int ncols, const double* Y, const double* Ycb,
                const double* Yct, const double* Ycl, const double* Ycr, const double* F, int ldF,
                int r, int N, double* V, int ldv, double* X, int ldx, hipStream_t st) {
  SSLAB_REQUIRE(nb >= 2 && m >= 1 && nxp >= 1 && r >= 0, "oms_rhs_rb0: needs >= 2 interfaces");
  SSLAB_REQUIRE(N == nb + 1 && ncols == 2 * m + 4 * nxp, "oms_rhs_rb0: bad sizes");
  if (r == 0) return kOk;  // nothing to do (empty buffers may be null)
  SSLAB_REQUIRE(Y && Ycb && Yct && Ycl && Ycr && F && V && X && (void*)V != (void*)X &&
                    ldF >= r && ldv >= r && ldx >= r,
                "oms_rhs_rb0: bad pointers / strides");
  const long long sY = (long long)m * ncols;
  const long long sV = (long long)m * ldv;
  const long long sX = (long long)m * ldx;
  const long long sF = (long long)nxp * ldF;
  const long long w4 = 4LL * nxp;
  const long long sC = (long long)m * w4;
  const double* Fbot = F;
  const double* Ftop = F + (long long)N * nxp * ldF;
  const double* Fleft = F + 2LL * N * nxp * ldF;
  const double* Fright = Fleft + (long long)m * ldF;
  const int nE = (nb + 1) / 2, nO = nb / 2;
  const bool last_odd = (nb - 1) % 2 == 1;  // the right boundary's slab is an odd node
  {
    const long long n = ((long long)nE * m + (long long)nb * m) * r;
    hipLaunchKernelGGL(k_zero_rb0, dim3(zero_grid(n)), dim3(256), 0, st, nE, m,
                       (long long)nb * m, r, V, ldv, X, ldx);
    SSLAB_HIP_CHECK(hipGetLastError());
  }
  GemmLaunch G{};
  int k = 0;
  // odd nodes j = 2c+1 without the right boundary: -(Y_j[:, bot] Fb_j + Y_j[:, top] Ft_j)
  const int nOi = last_odd ? nO - 1 : nO;
  if (nOi > 0) {
    GemmDesc d = gemm_desc(m, r, nOi, X + sX, ldx, 2 * sX, X + sX, ldx, 2 * sX, 1.0);
    gemm_add_term(d, gemm_term(Y + sY + 2 * m, ncols, 2 * sY, 0, Fbot + sF, ldF, 2 * sF, 0,
                               2 * nxp, -1.0));
    gemm_add_term(d, gemm_term(Y + sY + 2 * m + 2 * nxp, ncols, 2 * sY, 0, Ftop + sF, ldF, 2 * sF,
                               0, 2 * nxp, -1.0));
    G.d[k++] = d;
  }
  if (last_odd) {  // j = nb-1 also maps the right boundary data
    const long long j = nb - 1;
    GemmDesc d = gemm_desc(m, r, 1, X + j * sX, ldx, sX, X + j * sX, ldx, sX, 1.0);
    gemm_add_term(d, gemm_term(Y + j * sY + 2 * m, ncols, sY, 0, Fbot + j * sF, ldF, sF, 0,
                               2 * nxp, -1.0));
    gemm_add_term(d, gemm_term(Y + j * sY + 2 * m + 2 * nxp, ncols, sY, 0, Ftop + j * sF, ldF, sF,
                               0, 2 * nxp, -1.0));
    gemm_add_term(d, gemm_term(Y + j * sY + m, ncols, sY, 0, Fright, ldF, 0, 0, m, -1.0));
    G.d[k++] = d;
  }
  // even node 0: windows 1..3 (rows [0, 3 nxp)) and the left boundary
  {
    GemmDesc d = gemm_desc(m, r, 1, V, ldv, sV, V, ldv, sV, 1.0);
    gemm_add_term(d, gemm_term(Ycb + nxp, (int)w4, sC, 0, Fbot, ldF, 0, 0, 3 * nxp, 1.0));
    gemm_add_term(d, gemm_term(Yct + nxp, (int)w4, sC, 0, Ftop, ldF, 0, 0, 3 * nxp, 1.0));
    gemm_add_term(d, gemm_term(Ycl, m, 0, 0, Fleft, ldF, 0, 0, m, 1.0));
    G.d[k++] = d;
  }
  // the even node that sees the right boundary: node nb-1 when it is even, else node
  // nb-2 (through SiP_{nb-2} rhs_{nb-1}); when that is node 0 (nb = 2) its right term is a
  // separate product below
  const int br = last_odd ? (nb - 2) / 2 : nE - 1;
  if (br - 1 > 0) {  // interior even nodes 2b, 1 <= b < br: all four windows in the strip
    GemmDesc d = gemm_desc(m, r, br - 1, V + 2 * sV, ldv, 2 * sV, V + 2 * sV, ldv, 2 * sV, 1.0);
    gemm_add_term(d, gemm_term(Ycb + sC, (int)w4, sC, 0, Fbot + sF, ldF, 2 * sF, 0, (int)w4, 1.0));
    gemm_add_term(d, gemm_term(Yct + sC, (int)w4, sC, 0, Ftop + sF, ldF, 2 * sF, 0, (int)w4, 1.0));
    G.d[k++] = d;
  }
  if (br >= 1) {
    const long long i = 2LL * br;
    const int nwin = last_odd ? 4 : 3;  // node nb-1 has no window past the strip end
    GemmDesc d = gemm_desc(m, r, 1, V + i * sV, ldv, sV, V + i * sV, ldv, sV, 1.0);
    gemm_add_term(d, gemm_term(Ycb + br * sC, (int)w4, sC, 0, Fbot + (i - 1) * sF, ldF, sF, 0,
                               nwin * nxp, 1.0));
    gemm_add_term(d, gemm_term(Yct + br * sC, (int)w4, sC, 0, Ftop + (i - 1) * sF, ldF, sF, 0,
                               nwin * nxp, 1.0));
    gemm_add_term(d, gemm_term(Ycr, m, 0, 0, Fright, ldF, 0, 0, m, 1.0));
    G.d[k++] = d;
  }
  G.nd = k;
  SSLAB_TRY(gemm_launch(G, st));
  if (br == 0) {  // nb = 2: node 0 is also the right boundary's even node
    SSLAB_TRY(gemm_strided(0, 0, m, r, m, 1.0, Ycr, m, 0, Fright, ldF, 0, 1.0, V, ldv, 0, 1, st));
  }
  return kOk;
}

}  // namespace sslab
