// Boundary-folded forward sweep of the oms S-system solve (FP64, gfx950).  See oms_fold.hip.
#pragma once

#include "sslab_common.hpp"

namespace sslab {

struct FoldPlan;

int oms_fold_create(int nb, int m, int nF, const double* maps, int ldm, FoldPlan** plan,
                    hipStream_t st);
int oms_fold_apply(const FoldPlan* plan, const double* F, int ldF, int r, double* X, int ldx,
                   hipStream_t st);
long long oms_fold_bytes(const FoldPlan* plan);
int oms_fold_groups(const FoldPlan* plan);
void oms_fold_destroy(FoldPlan* plan);

}  // namespace sslab
