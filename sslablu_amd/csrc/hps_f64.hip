// HPS slab assembly on gfx950 (FP64): leaf Dirichlet-to-Neumann maps (2-D and 3-D leaves)
// and box merges (any dimension: the caller's index maps describe the geometry).
//
// Leaf (q x q Chebyshev grid, p = q-2):  interior collocation rows
//   A[u] = -c11 u_xx - c22 u_yy + c1 u_x + c2 u_y + c0 u      (pdo/pdo.py:19-39)
// give A_ii (p^2 x p^2) and A_ib (p^2 x 4p, edges ordered left, right, bottom, top,
// each by ascending coordinate, corners excluded).  The DtN map is
//   T = N_b - N_i A_ii^{-1} A_ib            (4p x 4p)
// with N the outward-normal derivative rows at the edge nodes (shared by all leaves).
//
// Merge of boxes alpha, beta sharing the edge nodes c (outward fluxes cancel on c):
//   u_c   = -(T^a_cc + T^b_cc)^{-1} [T^a_ce T^b_ce] u_e
//   T_new = blkdiag(T^a_ee, T^b_ee) - [T^a_ec; T^b_ec] (T^a_cc + T^b_cc)^{-1} [T^a_ce T^b_ce]
// The caller supplies index maps that place every node of the merged box in its
// canonical order, so one gather kernel builds all four operands.
//
// Every local solve (A_ii in the leaves, T^a_cc + T^b_cc in the merges and the double-slab
// pair solves) is an explicit batched inverse followed by an MFMA product: inv_batched
// works on the whole batch at once (block elimination on GEMMs, 64 x 64 Gauss-Jordan
// blocks, a sampled backward-error check with a pivoted-LU fallback), where a batched LU
// would spend most of its time in one-workgroup-per-matrix panels and triangular solves.
#include "hps_f64.hpp"

#include "gemm.hpp"
#include "inv.hpp"
#include "lu_f64.hpp"

namespace sslab {

namespace {

// One thread per (row r, column col) of [A_ii | A_ib] for one leaf (blockIdx.z).
// Rows r = (a-1)*p + (b-1) for interior grid node (a, b), a along x.
__global__ void k_leaf_assemble(int q, const double* __restrict__ Dx, const double* __restrict__ Dxx,
                                const double* __restrict__ Dy, const double* __restrict__ Dyy,
                                const double* __restrict__ c11, const double* __restrict__ c22,
                                const double* __restrict__ c1, const double* __restrict__ c2,
                                const double* __restrict__ c0, double* __restrict__ Aii,
                                double* __restrict__ Aib) {
  const int p = q - 2;
  const int p2 = p * p;
  const int ncol = p2 + 4 * p;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  const long long leaf = blockIdx.z;
  if (col >= ncol) return;
  const int a = r / p + 1, b = r % p + 1;
  const long long ci = leaf * p2 + r;
  const double k11 = c11 ? c11[ci] : 0.0;
  const double k22 = c22 ? c22[ci] : 0.0;
  const double k1 = c1 ? c1[ci] : 0.0;
  const double k2 = c2 ? c2[ci] : 0.0;
  const double k0 = c0 ? c0[ci] : 0.0;
  // grid node (a2, b2) of this column
  int a2, b2;
  if (col < p2) {
    a2 = col / p + 1;
    b2 = col % p + 1;
  } else {
    const int e = col - p2;
    const int side = e / p, t = e % p + 1;
    if (side == 0) { a2 = 0; b2 = t; }
    else if (side == 1) { a2 = q - 1; b2 = t; }
    else if (side == 2) { a2 = t; b2 = 0; }
    else { a2 = t; b2 = q - 1; }
  }
  double v = 0.0;
  if (b2 == b) v += -k11 * Dxx[a * q + a2] + k1 * Dx[a * q + a2];
  if (a2 == a) v += -k22 * Dyy[b * q + b2] + k2 * Dy[b * q + b2];
  if (a2 == a && b2 == b) v += k0;
  if (col < p2)
    Aii[leaf * p2 * p2 + (long long)r * p2 + col] = v;
  else
    Aib[leaf * p2 * 4 * p + (long long)r * 4 * p + (col - p2)] = v;
}

// 3-D leaf (q x q x q Chebyshev grid, p = q-2): one thread per (row r, column col) of
// [A_ii | A_ib] for one leaf (blockIdx.z).  Interior node (a, b, c), 1-based grid
// indices with a along x, is row r = ((a-1) p + (b-1)) p + (c-1).  Boundary columns are
// the six faces [x-lo, x-hi, y-lo, y-hi, z-lo, z-hi], p^2 nodes each (edges and corners
// excluded), each face in lexicographic order of its two tangential grid indices.
//   A[u] = -c11 u_xx - c22 u_yy - c33 u_zz + c1 u_x + c2 u_y + c3 u_z + c0 u
__global__ void k_leaf_assemble3(int q, const double* __restrict__ Dx,
                                 const double* __restrict__ Dxx, const double* __restrict__ Dy,
                                 const double* __restrict__ Dyy, const double* __restrict__ Dz,
                                 const double* __restrict__ Dzz, const double* __restrict__ c11,
                                 const double* __restrict__ c22, const double* __restrict__ c33,
                                 const double* __restrict__ c1, const double* __restrict__ c2,
                                 const double* __restrict__ c3, const double* __restrict__ c0,
                                 double* __restrict__ Aii, double* __restrict__ Aib) {
  const int p = q - 2;
  const int p2 = p * p, p3 = p2 * p, nb6 = 6 * p2;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  const long long leaf = blockIdx.z;
  if (col >= p3 + nb6) return;
  const int a = r / p2 + 1, b = (r / p) % p + 1, c = r % p + 1;
  const long long ci = leaf * p3 + r;
  const double k11 = c11 ? c11[ci] : 0.0;
  const double k22 = c22 ? c22[ci] : 0.0;
  const double k33 = c33 ? c33[ci] : 0.0;
  const double k1 = c1 ? c1[ci] : 0.0;
  const double k2 = c2 ? c2[ci] : 0.0;
  const double k3 = c3 ? c3[ci] : 0.0;
  const double k0 = c0 ? c0[ci] : 0.0;
  int a2, b2, c2i;
  if (col < p3) {
    a2 = col / p2 + 1;
    b2 = (col / p) % p + 1;
    c2i = col % p + 1;
  } else {
    const int e = col - p3;
    const int face = e / p2, u = (e % p2) / p + 1, v = e % p + 1;
    const int end = (face & 1) ? q - 1 : 0;
    if (face < 2) { a2 = end; b2 = u; c2i = v; }
    else if (face < 4) { a2 = u; b2 = end; c2i = v; }
    else { a2 = u; b2 = v; c2i = end; }
  }
  double val = 0.0;
  if (b2 == b && c2i == c) val += -k11 * Dxx[a * q + a2] + k1 * Dx[a * q + a2];
  if (a2 == a && c2i == c) val += -k22 * Dyy[b * q + b2] + k2 * Dy[b * q + b2];
  if (a2 == a && b2 == b) val += -k33 * Dzz[c * q + c2i] + k3 * Dz[c * q + c2i];
  if (a2 == a && b2 == b && c2i == c) val += k0;
  if (col < p3)
    Aii[leaf * p3 * p3 + (long long)r * p3 + col] = val;
  else
    Aib[leaf * p3 * nb6 + (long long)r * nb6 + (col - p3)] = val;
}

// out[b][i][j] = sum_s [rid_s[i] >= 0 && cid_s[j] >= 0] * src_s[b][rid_s[i]][cid_s[j]]
__global__ void k_gather2(int nrows, int ncols, const double* __restrict__ s0, int ld0, long long st0,
                          const double* __restrict__ s1, int ld1, long long st1,
                          const int* __restrict__ rid0, const int* __restrict__ cid0,
                          const int* __restrict__ rid1, const int* __restrict__ cid1,
                          double* __restrict__ out, int ldo, long long sto) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  const long long b = blockIdx.z;
  if (j >= ncols) return;
  double v = 0.0;
  const int r0 = rid0 ? rid0[i] : -1, c0 = cid0 ? cid0[j] : -1;
  if (r0 >= 0 && c0 >= 0) v += s0[b * st0 + (long long)r0 * ld0 + c0];
  const int r1 = rid1 ? rid1[i] : -1, c1 = cid1 ? cid1[j] : -1;
  if (r1 >= 0 && c1 >= 0) v += s1[b * st1 + (long long)r1 * ld1 + c1];
  out[b * sto + (long long)i * ldo + j] = v;
}

}  // namespace

int gather2(int batch, int nrows, int ncols, const double* s0, int ld0, long long st0,
            const double* s1, int ld1, long long st1, const int* rid0, const int* cid0,
            const int* rid1, const int* cid1, double* out, int ldo, long long sto,
            hipStream_t st) {
  SSLAB_REQUIRE(batch >= 0 && nrows >= 0 && ncols >= 0, "gather2: negative size");
  if (batch == 0 || nrows == 0 || ncols == 0) return kOk;
  SSLAB_REQUIRE(batch <= 65535 && nrows <= 65535, "gather2: batch or rows too large");
  SSLAB_REQUIRE(out != nullptr && ldo >= ncols, "gather2: bad output");
  hipLaunchKernelGGL(k_gather2, dim3(ceil_div(ncols, 256), nrows, batch), dim3(256), 0, st, nrows,
                     ncols, s0, ld0, st0, s1, ld1, st1, rid0, cid0, rid1, cid1, out, ldo, sto);
  SSLAB_HIP_CHECK(hipGetLastError());
  return kOk;
}

int hps_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints) {
  SSLAB_REQUIRE(nleaves >= 1 && q >= 3, "hps_leaf_sizes: bad sizes");
  const long long p = q - 2, p2 = p * p;
  long long iw = 0, ii = 0;
  inv_sizes((int)p2, nleaves, &iw, &ii);
  // A_ii, A_ii^{-1}, A_ib, X, then the inverse's workspace
  if (work_doubles) *work_doubles = (long long)nleaves * (2 * p2 * p2 + 2 * p2 * 4 * p) + iw;
  if (work_ints) *work_ints = ii + nleaves;
  return kOk;
}

int hps_leaf_dtn(int nleaves, int q, const double* Dx, const double* Dxx, const double* Dy,
                 const double* Dyy, const double* Nb, const double* Ni, const double* c11,
                 const double* c22, const double* c1, const double* c2, const double* c0,
                 double* T, double* X, double* work, int* iwork, int* info, hipStream_t st) {
  SSLAB_REQUIRE(nleaves >= 1 && q >= 3, "hps_leaf_dtn: bad sizes");
  SSLAB_REQUIRE(nleaves <= 65535, "hps_leaf_dtn: at most 65535 leaves per call");
  SSLAB_REQUIRE(Dx && Dxx && Dy && Dyy && Nb && Ni && T && work && iwork,
                "hps_leaf_dtn: null pointer");
  const int p = q - 2, p2 = p * p, nb4 = 4 * p;
  const long long s2 = (long long)p2 * p2, sb = (long long)p2 * nb4;
  double* Aii = work;
  double* Ainv = Aii + nleaves * s2;
  double* Aib = Ainv + nleaves * s2;
  double* Xs = X ? X : Aib + nleaves * sb;
  double* iw = Aib + 2 * nleaves * sb;
  long long iwn = 0;
  inv_sizes(p2, nleaves, nullptr, &iwn);
  int* linfo = info ? info : iwork + iwn;
  hipLaunchKernelGGL(k_leaf_assemble, dim3(ceil_div(p2 + nb4, 256), p2, nleaves), dim3(256), 0,
                     st, q, Dx, Dxx, Dy, Dyy, c11, c22, c1, c2, c0, Aii, Aib);
  SSLAB_HIP_CHECK(hipGetLastError());
  // X = A_ii^{-1} A_ib, then T = N_b - N_i X   (N shared by all leaves: batch stride 0)
  SSLAB_TRY(inv_batched<double>(p2, Aii, p2, s2, Ainv, p2, s2, nleaves, iw, iwork, linfo, 0.0,
                                nullptr, st));
  SSLAB_TRY(gemm_strided(0, 0, p2, nb4, p2, 1.0, Ainv, p2, s2, Aib, nb4, sb, 0.0, Xs, nb4, sb,
                         nleaves, st));
  GemmLaunch G{};
  G.nd = 1;
  G.d[0] = gemm_desc(nb4, nb4, nleaves, T, nb4, (long long)nb4 * nb4, Nb, nb4, 0, 1.0);
  gemm_add_term(G.d[0], gemm_term(Ni, p2, 0, 0, Xs, nb4, sb, 0, p2, -1.0));
  SSLAB_TRY(gemm_launch(G, st));
  return kOk;
}

int hps3_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints) {
  SSLAB_REQUIRE(nleaves >= 1 && q >= 3, "hps3_leaf_sizes: bad sizes");
  const long long p = q - 2, p3 = p * p * p, nb6 = 6 * p * p;
  long long iw = 0, ii = 0;
  inv_sizes((int)p3, nleaves, &iw, &ii);
  // A_ii, A_ii^{-1}, A_ib, X, then the inverse's workspace
  if (work_doubles) *work_doubles = (long long)nleaves * (2 * p3 * p3 + 2 * p3 * nb6) + iw;
  if (work_ints) *work_ints = ii + nleaves;
  return kOk;
}

int hps3_leaf_dtn(int nleaves, int q, const double* const* D, const double* Nb, const double* Ni,
                  const double* const* coef, double* T, double* X, double* work, int* iwork,
                  int* info, hipStream_t st) {
  SSLAB_REQUIRE(nleaves >= 1 && q >= 3, "hps3_leaf_dtn: bad sizes");
  SSLAB_REQUIRE(nleaves <= 65535, "hps3_leaf_dtn: at most 65535 leaves per call");
  SSLAB_REQUIRE(D && coef && Nb && Ni && T && work && iwork, "hps3_leaf_dtn: null pointer");
  for (int i = 0; i < 6; ++i) SSLAB_REQUIRE(D[i] != nullptr, "hps3_leaf_dtn: null derivative");
  const int p = q - 2, p3 = p * p * p, nb6 = 6 * p * p;
  SSLAB_REQUIRE(p3 <= 65535, "hps3_leaf_dtn: leaf order too large");
  const long long s3 = (long long)p3 * p3, sb = (long long)p3 * nb6;
  double* Aii = work;
  double* Ainv = Aii + nleaves * s3;
  double* Aib = Ainv + nleaves * s3;
  double* Xs = X ? X : Aib + nleaves * sb;
  double* iw = Aib + 2 * nleaves * sb;
  long long iwn = 0;
  inv_sizes(p3, nleaves, nullptr, &iwn);
  int* linfo = info ? info : iwork + iwn;
  hipLaunchKernelGGL(k_leaf_assemble3, dim3(ceil_div(p3 + nb6, 256), p3, nleaves), dim3(256), 0,
                     st, q, D[0], D[1], D[2], D[3], D[4], D[5], coef[0], coef[1], coef[2], coef[3],
                     coef[4], coef[5], coef[6], Aii, Aib);
  SSLAB_HIP_CHECK(hipGetLastError());
  // X = A_ii^{-1} A_ib, then T = N_b - N_i X   (N shared by all leaves: batch stride 0)
  SSLAB_TRY(inv_batched<double>(p3, Aii, p3, s3, Ainv, p3, s3, nleaves, iw, iwork, linfo, 0.0,
                                nullptr, st));
  SSLAB_TRY(gemm_strided(0, 0, p3, nb6, p3, 1.0, Ainv, p3, s3, Aib, nb6, sb, 0.0, Xs, nb6, sb,
                         nleaves, st));
  GemmLaunch G{};
  G.nd = 1;
  G.d[0] = gemm_desc(nb6, nb6, nleaves, T, nb6, (long long)nb6 * nb6, Nb, nb6, 0, 1.0);
  gemm_add_term(G.d[0], gemm_term(Ni, p3, 0, 0, Xs, nb6, sb, 0, p3, -1.0));
  SSLAB_TRY(gemm_launch(G, st));
  return kOk;
}

int hps_merge_sizes(int nmerge, int n_new, int c, long long* work_doubles, long long* work_ints) {
  SSLAB_REQUIRE(nmerge >= 1 && n_new >= 1 && c >= 1, "hps_merge_sizes: bad sizes");
  long long iw = 0, ii = 0;
  inv_sizes(c, nmerge, &iw, &ii);
  // T_cc, its inverse, R, X, E, then the inverse's workspace
  if (work_doubles)
    *work_doubles = (long long)nmerge * (2LL * c * c + 3LL * c * n_new) + iw;
  if (work_ints) *work_ints = ii + nmerge;
  return kOk;
}

int hps_merge(int nmerge, int n_a, int n_b, int n_new, int c, const double* Ta, long long sTa,
              const double* Tb, long long sTb, const int* ca, const int* cb, const int* na,
              const int* nbix, double* Tnew, long long sTnew, double* work, int* iwork, int* info,
              hipStream_t st, double* Xsol, long long sXsol) {
  SSLAB_REQUIRE(nmerge >= 1 && n_a >= 1 && n_b >= 1 && n_new >= 1 && c >= 1,
                "hps_merge: bad sizes");
  SSLAB_REQUIRE(Ta && Tb && ca && cb && na && nbix && Tnew && work && iwork,
                "hps_merge: null pointer");
  const long long c2 = (long long)c * c, cn = (long long)c * n_new;
  double* Tcc = work;
  double* Tinv = Tcc + nmerge * c2;
  double* R = Tinv + nmerge * c2;
  // X = (T^a_cc + T^b_cc)^{-1} [T_{c,new}]: the shared edge's values are -X u_new
  // (kept in Xsol, stride sXsol, when the caller wants the merge's solution operator)
  double* X = Xsol ? Xsol : R + nmerge * cn;
  const long long sX = Xsol ? sXsol : cn;
  double* E = R + 2 * nmerge * cn;
  double* iw = E + nmerge * cn;
  long long iwn = 0;
  inv_sizes(c, nmerge, nullptr, &iwn);
  int* linfo = info ? info : iwork + iwn;
  // na[i] / nbix[i]: index of merged node i in alpha / beta (-1 if it is not there).
  SSLAB_TRY(gather2(nmerge, c, c, Ta, n_a, sTa, Tb, n_b, sTb, ca, ca, cb, cb, Tcc, c, c2, st));
  SSLAB_TRY(gather2(nmerge, c, n_new, Ta, n_a, sTa, Tb, n_b, sTb, ca, na, cb, nbix, R, n_new, cn,
                    st));
  SSLAB_TRY(gather2(nmerge, n_new, c, Ta, n_a, sTa, Tb, n_b, sTb, na, ca, nbix, cb, E, c, cn, st));
  SSLAB_TRY(gather2(nmerge, n_new, n_new, Ta, n_a, sTa, Tb, n_b, sTb, na, na, nbix, nbix, Tnew,
                    n_new, sTnew, st));
  SSLAB_TRY(inv_batched<double>(c, Tcc, c, c2, Tinv, c, c2, nmerge, iw, iwork, linfo, 0.0, nullptr,
                                st));
  SSLAB_TRY(gemm_strided(0, 0, c, n_new, c, 1.0, Tinv, c, c2, R, n_new, cn, 0.0, X, n_new, sX,
                         nmerge, st));
  SSLAB_TRY(gemm_strided(0, 0, n_new, n_new, c, -1.0, E, c, cn, X, n_new, sX, 1.0, Tnew, n_new,
                         sTnew, nmerge, st));
  return kOk;
}

int hps_pair_solve(int npair, int n_s, int c, int ncols, const double* Ts, long long sTs,
                   const int* ca, const int* cb, const int* cola, const int* colb, double* Y,
                   double* work, int* iwork, int* info, hipStream_t st) {
  SSLAB_REQUIRE(npair >= 1 && n_s >= 1 && c >= 1 && ncols >= 1, "hps_pair_solve: bad sizes");
  SSLAB_REQUIRE(Ts && ca && cb && cola && colb && Y && work && iwork,
                "hps_pair_solve: null pointer");
  const long long c2 = (long long)c * c, cn = (long long)c * ncols;
  // workspace as laid out by hps_merge_sizes(npair, ncols, c)
  double* Tcc = work;
  double* Tinv = Tcc + npair * c2;
  double* R = Tinv + npair * c2;
  double* iw = R + 3 * npair * cn;
  long long iwn = 0;
  inv_sizes(c, npair, nullptr, &iwn);
  int* linfo = info ? info : iwork + iwn;
  // alpha = slab i, beta = slab i+1 (same array, offset by one slab).
  SSLAB_TRY(gather2(npair, c, c, Ts, n_s, sTs, Ts + sTs, n_s, sTs, ca, ca, cb, cb, Tcc, c, c2,
                    st));
  SSLAB_TRY(gather2(npair, c, ncols, Ts, n_s, sTs, Ts + sTs, n_s, sTs, ca, cola, cb, colb, R,
                    ncols, cn, st));
  SSLAB_TRY(inv_batched<double>(c, Tcc, c, c2, Tinv, c, c2, npair, iw, iwork, linfo, 0.0, nullptr,
                                st));
  return gemm_strided(0, 0, c, ncols, c, 1.0, Tinv, c, c2, R, ncols, cn, 0.0, Y, ncols, cn, npair,
                      st);
}

}  // namespace sslab
