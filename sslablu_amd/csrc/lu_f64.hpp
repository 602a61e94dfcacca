// Batched LU with partial pivoting and the solves built on it (gfx950), FP64 and FP32
// (E = double or float; FP32 for the low-precision factorizations).
//
// Conventions follow LAPACK getrf/getrs as exposed by scipy.linalg.lu_factor:
// row-major storage here, the packed factor holds unit-lower L below the diagonal
// and U on and above it, ipiv[i] (0-based) is the row swapped with row i at step i,
// and info = i+1 for the first exactly-zero pivot U[i][i] (0 when none).
#pragma once

#include "sslab_common.hpp"

namespace sslab {

// In-place blocked right-looking LU of `batch` n x n matrices A_b = A + b*sA.
// ipiv: batch x n ints (stride sP); info: batch ints (may be null).
template <typename E>
int getrf_batched(int n, E* A, int lda, long long sA, int* ipiv, long long sP,
                  int* info, int batch, hipStream_t stream);

// Solve A_b X_b = B_b with the factors from getrf_batched.  B (n x nrhs, row-major,
// leading dimension ldb, stride sB) is read; X (same shape, ldx, sX) receives the
// solution.  X may equal B.  perm: batch x n int scratch (stride n).
template <typename E>
int getrs_batched(int n, int nrhs, const E* LU, int lda, long long sA, const int* ipiv,
                  long long sP, const E* B, int ldb, long long sB, E* X, int ldx,
                  long long sX, int* perm, int batch, hipStream_t stream);

// Solve A_b^T X_b = B_b with the factors from getrf_batched (B read, X written, distinct
// from B).  W: batch x n x nrhs workspace (stride n*nrhs); perm as in getrs_batched.
template <typename E>
int getrs_trans_batched(int n, int nrhs, const E* LU, int lda, long long sA, const int* ipiv,
                        long long sP, const E* B, int ldb, long long sB, E* X, int ldx,
                        long long sX, E* W, int* perm, int batch, hipStream_t stream);

// In place X_b <- U_b^{-1} X_b for upper-triangular U_b (n x n, non-unit diagonal).
template <typename E>
int trsm_upper_batched(int n, int nrhs, const E* U, int ldu, long long sU, E* X, int ldx,
                       long long sX, int batch, hipStream_t stream);

// X_b = I (n x n, leading dimension ldx) for every batch entry.
template <typename E>
int set_identity_batched(int n, E* X, int ldx, long long sX, int batch, hipStream_t stream);

// Y_b[r, c] = X_b[r, c] for an m x n block (row-major, leading dims), batched.
template <typename E>
int copy_block_batched(int m, int n, const E* X, int ldx, long long sX, E* Y, int ldy,
                       long long sY, int batch, hipStream_t stream);

// Fill an m x n block with zeros, batched (a kernel, never hipMemset).
template <typename E>
int zero_block_batched(int m, int n, E* Y, int ldy, long long sY, int batch,
                       hipStream_t stream);

}  // namespace sslab
