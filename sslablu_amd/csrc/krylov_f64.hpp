// Column-batched vector kernels of the device GMRES (FP64).  See krylov_f64.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

int col_dots(int n, int r, int k, const double* V, long long sV, int ldv, const double* W,
             int ldw, double* H, int ldh, hipStream_t st);
int col_axpy(int n, int r, int k, double alpha, const double* V, long long sV, int ldv,
             const double* H, int ldh, double* W, int ldw, hipStream_t st);
int col_scale(int n, int r, const double* W, int ldw, const double* s, double* Out, int ldo,
              hipStream_t st);

}  // namespace sslab
