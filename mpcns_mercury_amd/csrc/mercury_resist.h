// Launch wrappers of the implicit resistive correction (mercury_resist.hip).
#pragma once
#include "mercury_kernels.h"

namespace merc {

// the unknown edges on the device: box index, axis (0..2), the i/j strides of
// the edge's block and eta
struct ResDofs {
    int n = 0;
    const int *q = nullptr, *ax = nullptr, *sI = nullptr, *sJ = nullptr;
    const double *eta = nullptr;
};

// slots of the CG scalar array
enum { CG_RZ = 0, CG_PAP, CG_ALPHA, CG_BETA, CG_RR, CG_BB, CG_NSCALAR = 8 };

int res_blocks(int n);  // blocks of the reducing kernels (length of one partial array)
void launch_res_scatter(hipStream_t st, const ResDofs &d, const double *x, double *const *F, int slot = SL_ERES);
void launch_res_pick(hipStream_t st, const ResDofs &d, double *const *F, int slot, double *x);
void launch_res_gather(hipStream_t st, const ResDofs &d, double *const *F, double *y, const double *p, const double *w, double *part);
void launch_res_rhs(hipStream_t st, const ResDofs &d, double *const *F, double dt, double *b, double *dg, double *w, double *wd);
void launch_cg_init(hipStream_t st, int n, const double *b, const double *Ax, const double *dg, const double *wd, double *r, double *p,
                    double *part, double *S);
void launch_cg_pap(hipStream_t st, int n, const double *part, double *S);
// r -= alpha Ap with the partials of r.r and r.z.w, then beta
void launch_cg_r(hipStream_t st, int n, const double *S, const double *Ap, const double *wd, double *r, double *part, double *Sw);
// x += alpha p; p = z + beta p; p into the edge slots of the unknowns
void launch_cg_pxs(hipStream_t st, const ResDofs &d, const double *S, const double *r, const double *dg, double *p, double *x,
                   double *const *F, int slot = SL_ERES);
void launch_bres_update(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, double dt);

}  // namespace merc
