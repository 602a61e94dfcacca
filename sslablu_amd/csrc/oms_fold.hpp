// Boundary-folded forward sweep of the oms S-system solve (FP64; FP32 copy for mixed solves).  See oms_fold.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

struct FoldPlan;

int oms_fold_create(int nb, int m, int nF, const double* maps, int ldm, FoldPlan** plan,
                    hipStream_t st);
int oms_fold_apply(const FoldPlan* plan, const double* F, int ldF, int r, double* X, int ldx,
                   hipStream_t st);
// FP32 copy of the plan's maps (once; freed with the plan) for oms_fold_solve_mixed.
int oms_fold_enable_f32(FoldPlan* plan, hipStream_t st);
// Mixed-precision solve from boundary data F (nF x r, contiguous): FP32 folded solve, then
// `iters` FP64-residual refinement steps with the FP32 red-black factor (identity
// diagonal).  B = rhs(F) in FP64; work64: nb m r doubles; work32: (2 nb m + nF) r floats.
int oms_fold_solve_mixed(const FoldPlan* plan, const float* factor32, const double* SiM,
                         const double* SiP, const double* F, int ldF, int r, const double* B,
                         double* X, double* work64, float* work32, int iters, hipStream_t st);
long long oms_fold_bytes(const FoldPlan* plan);
int oms_fold_groups(const FoldPlan* plan);
void oms_fold_destroy(FoldPlan* plan);

}  // namespace sslab
