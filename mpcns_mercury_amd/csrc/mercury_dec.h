// Host-side launchers of the DEC operators on caller device arrays (mercury_dec.hip).
#pragma once
#include <hip/hip_runtime.h>

// partial maxima the no-allocation diagnostics can hold: one per workgroup of
// 256 points, so at most 33 554 432 points (1 MiB of scratch)
#define MERC_DEC_SCRATCH_DOUBLES 131072

namespace merc {

void dec_check_box(const int *dims, int nghost);
void dec_d0(const int *dims, int g, const double *node, double *const *edge, hipStream_t st);
// update = 0: face = d1(edge); update = 1: face -= dt * d1(edge)
void dec_d1(const int *dims, int g, const double *const *edge, double *const *face, int update, double dt, hipStream_t st);
void dec_d2(const int *dims, int g, const double *const *face, double *cell, hipStream_t st);
// workgroups (= partial maxima) of each diagnostic
long long dec_div_curl_groups(const int *dims);
long long dec_curl_grad_groups(const int *dims);
// partial: dec_*_groups(dims) doubles; out: one device double
void dec_div_curl(const int *dims, int g, const double *const *edge, double *partial, double *out, hipStream_t st);
void dec_curl_grad(const int *dims, int g, const double *node, double *partial, double *out, hipStream_t st);
// this GPU's copy of the static scratch array (set up on the first call per device)
double *dec_noalloc_scratch();

}  // namespace merc
