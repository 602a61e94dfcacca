// Block-tridiagonal S-system factor / solve (FP64, gfx950).  See blocktri_f64.hip.
#pragma once

#include "gemm.hpp"
#include "sslab_common.hpp"

namespace sslab {

int rb_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
             long long* work_ints);
int rb_info_count(int nb);
// Shape of the red-black plan for nb blocks of size m: the number of elimination levels
// (returned), the odd-node count of every level (nodd, at least that many entries; may be
// null) and the node count of the coarsest level (n_top).
int rb_shape(int nb, int m, int* nodd, int nodd_cap, int* n_top);
int rb_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
              double* factor, double* work, int* iwork, int* info, hipStream_t st);
int rb_factor_f32(int nb, int m, const float* T, const float* SiM, const float* SiP,
                  float* factor, float* work, int* iwork, int* info, hipStream_t st);
int rb_solve(int nb, int m, int r, const double* factor, double* V, int ldv, double* X, int ldx,
             hipStream_t st);
// Per-step tile policies of the red-black backward sweep (step u runs level L-1-u) for
// the calling thread's later solves; n = 0 clears them.  Each entry is validated as by
// gemm_set_policy.
int rb_set_backward_policies(const GemmPolicy* p, int n);
int rb_get_backward_policy_count();

int rb_solve_ex(int nb, int m, int r, const double* factor, double* V, int ldv, double* X,
                int ldx, int flags, hipStream_t st);
// Abort flag of the dependency-ordered backward sweeps on the current GPU (synchronises
// the device; clears the flag).
int rb_chain_status(int* aborted);
// Forward sweep + coarsest level of the identity-diagonal red-black solve for a V whose
// block rows are zero outside host-given column intervals [lo[i], hi[i]); X receives the
// backward sweep's entry state.  See blocktri_f64.hip.
int rb_forward_windowed(int nb, int m, int ncols, const double* factor, double* V, int ldv,
                        double* X, int ldx, const int* lo, const int* hi, hipStream_t st);
int rb_solve_multi(int nsys, const int* nb, const int* m, const int* r, const double* const* factor,
                   double* const* V, const int* ldv, double* const* X, const int* ldx,
                   hipStream_t st);
int rb_solve_f32(int nb, int m, int r, const float* factor, float* V, int ldv, float* X, int ldx,
                 hipStream_t st);
int rb_residual(int nb, int m, int r, const double* T, const double* SiM, const double* SiP,
                const double* B, int ldb, const double* X, int ldx, double* R, int ldr,
                hipStream_t st);
int rb_solve_mixed(int nb, int m, int r, const float* factor32, const double* T,
                   const double* SiM, const double* SiP, const double* B, double* X, double* R64,
                   float* V32, float* X32, int iters, hipStream_t st);
// The refinement steps of rb_solve_mixed from a given FP64 X: `iters` times r = B - S X
// (FP64), its FP32 solve and X += (FP64) correction.
int rb_refine_mixed(int nb, int m, int r, const float* factor32, const double* T,
                    const double* SiM, const double* SiP, const double* B, double* X,
                    double* R64, float* V32, float* X32, int iters, hipStream_t st);
// rb_solve_ex with an FP32 factor (flags 1, 2, 4, 8 as there; no chained sweep).
int rb_solve_f32_ex(int nb, int m, int r, const float* factor, float* V, int ldv, float* X,
                    int ldx, int flags, hipStream_t st);

// rb_solve_mixed with its workspace in a static per-GPU device array (64 MiB: 16 bytes
// per element of the nb*m x r panel, each of its three parts rounded up to 256 bytes).
int rb_solve_mixed_noalloc(int nb, int m, int r, const float* factor32, const double* T,
                           const double* SiM, const double* SiP, const double* B, double* X,
                           int iters, hipStream_t st);

int thomas_sizes(int nb, int m, long long* factor_doubles, long long* work_doubles,
                 long long* work_ints);
int thomas_factor(int nb, int m, const double* T, const double* SiM, const double* SiP,
                  double* factor, double* work, int* iwork, int* info, hipStream_t st);
int thomas_solve(int nb, int m, int r, const double* factor, double* V, int ldv, double* X,
                 int ldx, hipStream_t st);

}  // namespace sslab
