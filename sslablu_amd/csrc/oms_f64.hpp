// S-system right-hand sides from global boundary data (FP64).  See oms_f64.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

int oms_rhs(int nb, int m, int nxp, int ncols, const double* Y, const double* F, int ldF, int r,
            int N, double* out, int ldo, hipStream_t st);

int oms_rhs_rb0(int nb, int m, int nxp, int ncols, const double* Y, const double* Ycb,
                const double* Yct, const double* Ycl, const double* Ycr, const double* F, int ldF,
                int r, int N, double* V, int ldv, double* X, int ldx, hipStream_t st);

}  // namespace sslab
