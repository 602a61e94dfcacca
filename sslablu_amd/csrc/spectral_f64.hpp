// Spectral double-slab discretization, 2-D with variable coefficients (FP64, gfx950).
// See spectral_f64.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

int spectral2d_sizes(int nslab, int qx, int qy, long long* work_doubles, long long* work_ints);
int spectral2d_slab_maps(int nslab, int qx, int qy, const double* Dx, const double* Dxx,
                         const double* Dy, const double* Dyy, const double* c11,
                         const double* c22, const double* c1, const double* c2, const double* c0,
                         double* Y, double* work, int* iwork, int* info, hipStream_t st);

}  // namespace sslab
