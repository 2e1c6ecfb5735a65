// Launch wrappers of the fused FP64 kernels (mercury_fused.hip).
#pragma once
#include "mercury_kernels.h"

namespace merc {
void launch_rhs_cell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P);
void launch_cell_end(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P, unsigned long long *dt_next,
                     int compute_dt);
void launch_eface_u(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F);
void launch_edge_total(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P, int with_hall);
void launch_update_all(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const double *dt);
void launch_bcell_jedge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F);
void launch_dt_commit(hipStream_t st, unsigned long long *dt, unsigned long long *dt_next);
}  // namespace merc
