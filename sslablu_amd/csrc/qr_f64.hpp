// Batched FP64 Householder QR with optional column pivoting.  See qr_f64.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

constexpr int kQrMaxRows = 1024;
constexpr int kQrMaxCols = 1024;

// In place: A_b (m x n, row-major, lda) -> R (upper trapezoid, zeros below).  Q (m x m,
// ldq, optional) receives the complete orthogonal factor; perm (n ints, optional)
// enables column pivoting and receives the column order: A_in[:, perm] = Q R.
int geqrf_batched(int m, int n, double* A, int lda, long long sA, double* Q, int ldq, long long sQ,
                  int* perm, long long sPerm, int batch, hipStream_t st);

// For every parent b < nparent and block blk < nblk: the bs x bs diagonal block
// (blk, blk) of Y_b (ldy, stride sY) += X[b*nblk + blk] (contiguous bs x bs blocks).
int add_diag_blocks(int bs, int nblk, int nparent, const double* X, double* Y, int ldy,
                    long long sY, hipStream_t st);

}  // namespace sslab
