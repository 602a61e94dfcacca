// Block-tridiagonal S-system factor / solve (FP64, gfx950).  See blocktri_f64.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

int rb_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
             long long* work_ints);
int rb_info_count(int nb);
int rb_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
              double* factor, double* work, int* iwork, int* info, hipStream_t st);
int rb_solve(int nb, int m, int r, const double* factor, double* V, int ldv, double* X, int ldx,
             hipStream_t st);

int thomas_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
                 long long* work_ints);
int thomas_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
                  double* factor, double* work, int* iwork, int* info, hipStream_t st);
int thomas_solve(int nb, int m, int r, const double* factor, double* V, int ldv, double* X,
                 int ldx, hipStream_t st);

}  // namespace sslab
