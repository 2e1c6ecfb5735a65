// Launch wrappers of the fused kernels (mercury_fused.hip); mix selects the
// mixed-precision instantiation (FP32 fluxes / EMFs / Hall / reconstruction).
#pragma once
#include "mercury_kernels.h"

namespace merc {
// R: per block ceil(m0 / L) x m1 x m2 points; each thread marches over L cells along i
void launch_rhs_cell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                     int L, bool mix);
// R: per block ceil(m0 / C) x ceil(m1 / 16) x ceil(m2 / 16) tiles, one workgroup
// each, marching over C planes along i (the LDS-tiled flux pass)
void launch_rhs_tile(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                     int C, bool mix);
// dt_next: one word per run (nruns of them), reduced per run
void launch_cell_end(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                     unsigned long long *dt_next, int compute_dt, bool mix, int nruns);
// ambi: also the node electron pressure of the ambipolar EMF (read by k_edge_total)
// ns: ion species of the run (2 or 3)
void launch_eface_u(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, bool mix, int ambi,
                    int ns);
void launch_edge_total(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                       int with_hall, bool mix);
void launch_update_all(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const double *dt,
                       bool mix, int ns);
void launch_to_double(hipStream_t st, const float *src, double *dst, long long n);
// divb_part (2 doubles per workgroup of R) and divb_mon (3 doubles), when not
// null: the div B monitor of the synchronised B faces (k_divb_monitor)
void launch_bcell_jedge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, bool mix,
                        double *divb_part = nullptr, double *divb_mon = nullptr);
void launch_divb_monitor_reset(hipStream_t st, double *divb_mon);
void launch_jcell_t(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, bool mix);
void launch_to_float(hipStream_t st, const double *src, float *dst, long long n);
// per run: the abort rule on dt_next[r], dt[r] committed, status 4 words per run
void launch_dt_commit(hipStream_t st, unsigned long long *dt, const unsigned long long *dt_next, unsigned long long *status, int nruns);
}  // namespace merc
