// Host-side launch wrappers of the FP64 kernels (mercury_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mercury_common.h"

namespace merc {

// a device MercRange plus its total thread count
struct RangeH {
    MercRange dev{};
    long long total = 0;
};

void launch_pv_uplus(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P);
void launch_bcell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F);
void launch_jedge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int s_face = SL_B, int s_edge = SL_J);
void launch_jcell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F);
// dt_bits: one word per run (nruns of them), each the minimum over its run's cells
void launch_timestep(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P, unsigned long long *dt_bits,
                     int nruns);
void launch_flux_div(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P);
void launch_sources(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P);
void launch_eface(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F);
void launch_edge_emf(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P);
void launch_hall_cell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P);
void launch_hall_edge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F);
void launch_add_hall_wall(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F);
void launch_curl_update(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const double *dt, int do_update);
// ns: ion species of the run (2 or 3)
void launch_update_u(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const double *dt, int ns);
void launch_rk2(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int save, int ns);
void launch_dec_d0(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int s_node, int s_edge);
void launch_dec_d1(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int s_edge, int s_face, double mult);
void launch_dec_d2(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int s_face, int s_cell);
// d2 B diagnostic: partial holds 2 * divb_partials() doubles, out 2 words (max|d2 B|, max|B_face| bits)
int divb_partials();
void launch_divb(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, double *partial, unsigned long long *out,
                 int write_field);

void launch_alias_reduce(hipStream_t st, const int *cls_start, const int *member, const int *msign, const int *maxis, const int *owner,
                         const int *osign, const int *oaxis, long long ncls, double *const *F, int slot0);
void launch_zero_span(hipStream_t st, double *const *F, const int *slots, int ns, long long n);
void launch_copy_multi(hipStream_t st, const int *ent, long long n, double *const *F, const MultiSets &S);
void launch_fill_multi(hipStream_t st, const int *ent, long long n, double *const *F, const MultiSets &S);
void launch_setv_multi(hipStream_t st, const int *ent, const double *val, long long n, double *const *F);
void launch_wall_multi(hipStream_t st, const int *ent, long long n, int g, double *const *F, const MercPhys &P);
// n entries (dst slot, dst, src slot, src, sign): dst <- sign * src
void launch_copy_plan(hipStream_t st, const int *ent, long long n, double *const *F);
// the same entries with every read before every write (sources that are also
// destinations read their old values): tmp holds n doubles
void launch_copy_plan_staged(hipStream_t st, const int *ent, long long n, double *const *F, double *tmp);

}  // namespace merc
