// Device side of the visualisation output (mercury_output.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mercury_common.h"

namespace merc {
// Node values of one cell field component over the owned nodes of one block
// (i in [0, m0], j in [0, m1], k in [0, m2]), k fastest: out[(i*(m1+1) + j)*(m2+1) + k].
// field and jac point at the block's first element (box layout of MercBlk).
void launch_node_average(hipStream_t st, const double *field, const double *jac, const MercBlk &b, double *out);
}  // namespace merc
