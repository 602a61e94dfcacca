// HPS slab assembly (FP64, gfx950).  See hps_f64.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

int gather2(int batch, int nrows, int ncols, const double* s0, int ld0, long long st0,
            const double* s1, int ld1, long long st1, const int* rid0, const int* cid0,
            const int* rid1, const int* cid1, double* out, int ldo, long long sto,
            hipStream_t st);

int hps_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints);
int hps_leaf_dtn(int nleaves, int q, const double* Dx, const double* Dxx, const double* Dy,
                 const double* Dyy, const double* Nb, const double* Ni, const double* c11,
                 const double* c22, const double* c1, const double* c2, const double* c0,
                 double* T, double* X, double* work, int* iwork, int* info, hipStream_t st);

// 3-D leaves (q^3 Chebyshev points): D = {Dx, Dxx, Dy, Dyy, Dz, Dzz} (q x q each),
// Nb (6p^2 x 6p^2) and Ni (6p^2 x p^3) the outward-normal flux rows, coef = {c11, c22,
// c33, c1, c2, c3, c0} per leaf-interior node (entries may be null).  T: nleaves x
// 6p^2 x 6p^2; X (optional): nleaves x p^3 x 6p^2 = A_ii^{-1} A_ib.
int hps3_leaf_sizes(int nleaves, int q, long long* work_doubles, long long* work_ints);
int hps3_leaf_dtn(int nleaves, int q, const double* const* D, const double* Nb, const double* Ni,
                  const double* const* coef, double* T, double* X, double* work, int* iwork,
                  int* info, hipStream_t st);

int hps_merge_sizes(int nmerge, int n_new, int c, long long* work_doubles, long long* work_ints);
int hps_merge(int nmerge, int n_a, int n_b, int n_new, int c, const double* Ta, long long sTa,
              const double* Tb, long long sTb, const int* ca, const int* cb, const int* na,
              const int* nbix, double* Tnew, long long sTnew, double* work, int* iwork, int* info,
              hipStream_t st, double* Xsol = nullptr, long long sXsol = 0);

int hps_pair_solve(int npair, int n_s, int c, int ncols, const double* Ts, long long sTs,
                   const int* ca, const int* cb, const int* cola, const int* colb, double* Y,
                   double* work, int* iwork, int* info, hipStream_t st);

}  // namespace sslab
