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
void launch_res_rhs(hipStream_t st, const ResDofs &d, double *const *F, double dt, double *b, double *dg, double *w);
void launch_cg_init(hipStream_t st, int n, const double *b, const double *Ax, const double *dg, const double *w, double *r, double *p,
                    double *part, double *S);
void launch_cg_pap(hipStream_t st, int n, const double *part, double *S);
void launch_cg_xr(hipStream_t st, int n, const double *S, const double *p, const double *Ap, const double *dg, const double *w, double *x,
                  double *r, double *part, double *Sw);
void launch_cg_p(hipStream_t st, int n, const double *S, const double *r, const double *dg, double *p);
void launch_bres_update(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, double dt);

}  // namespace merc
