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

}  // namespace sslab
