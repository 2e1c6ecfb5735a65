// Device-side records of the singular-edge assembly (built on the host from
// merc::singular::Registry, applied by the kernels of mercury_singular_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mercury_common.h"

// one singular edge: its owner DOF (arena index and axis), the owner alias's
// orientation, the mean edge vector and its cell and face ranges
struct MercSingEntry {
    int oq, oax, c0, c1, f0, f1;
    double osign, cv[3];
};
// one cell around a singular edge: arena index, the stride along the axis of
// the alias it was found from, its block; weight and that alias's orientation;
// the static Hall taper and local-resistivity factor at its centre
struct MercSingCell {
    int q, d, blk, leta_on;
    double w, orient, taper, leta;
};
// one owner face around a singular edge: inverse Hodge * orientation * incidence
struct MercSingFace {
    int q, ax;
    double coef;
};

// E (non-Hall) and E_hall owners of every singular edge; mix: U+ from the FP32 table
void launch_singular_emf(hipStream_t s, int n, const MercSingEntry *E, const MercSingCell *C, const MercBlk *B, double *const *F,
                         const float *const *G, const MercPhys &P, bool mix);
// the current (or its increment) owners of every singular edge from the face
// triplet at slot sB into the edge triplet at slot sJ
void launch_singular_current(hipStream_t s, int n, const MercSingEntry *E, const MercSingFace *Fc, double *const *F, int sB, int sJ);
// Tecplot node values of singular nodes: per stencil the |Jac|-weighted mean of
// the field over its cells (arena indices cq[start[s] .. start[s+1]))
void launch_node_stencil_mean(hipStream_t s, int n, const int *start, const int *cq, const double *field, const double *jac,
                              double *out);
