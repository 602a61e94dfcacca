// Batched explicit inverses by recursive 2 x 2 block elimination (gfx950), FP64 and FP32.
// See inv.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

// Workspace of inv_batched for `batch` n x n matrices: `work` elements of E and `iwork`
// ints (the pivoted-LU fallback's factors, pivots and info included).
void inv_sizes(int n, int batch, long long* work, long long* iwork);

// X_b = A_b^{-1} for `batch` n x n matrices (row-major; A_b = A + b sA, X_b = X + b sX;
// A is read only and may not overlap X).  The inverse is formed by 2 x 2 block
// elimination down to 64 x 64 diagonal blocks, which one wave inverts by Gauss-Jordan
// with partial pivoting; every block product is an MFMA GEMM launch.  Its normwise
// backward error is then sampled (||G - A_b X_b G||_F / (||A_b||_F ||X_b G||_F) for four
// +-1 columns G, one host synchronisation); when some matrix misses `tol` (or is not
// finite) the whole batch is
// inverted again through the pivoted LU (getrf_batched + getrs_batched against I), and
// info receives getrf's per-matrix info (else zeros).  tol <= 0 selects the default
// (1e-12 in FP64, 1e-3 in FP32).  *fallback (may be null) is set to 1 when the LU path
// ran.  work / iwork: at least inv_sizes(n, batch).
//
// mode kInvDeferred: no synchronisation and no fallback -- the check writes flags[b] = 1
// for every matrix that misses tol (else 0, device ints) and info (if any) is zeroed; the
// caller reads the flags once after a chain of such calls and redoes the chain with
// kInvLU (the pivoted LU alone) when one is set.
enum InvMode { kInvChecked = 0, kInvDeferred = 1, kInvLU = 2 };
template <typename E>
int inv_batched(int n, const E* A, int lda, long long sA, E* X, int ldx, long long sX, int batch,
                E* work, int* iwork, int* info, double tol, int* fallback, hipStream_t st,
                int mode = kInvChecked, int* flags = nullptr);

}  // namespace sslab
