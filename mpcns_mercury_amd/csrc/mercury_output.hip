// Cell -> node values of the all-node visualisation output.
//
// IOModule::EvalValue_CellToNode_ (src/6_io/IOModule_Tecplot.cpp:322-418): the
// value at a node is the |Jac|-weighted mean of the 8 cells around it, ghost
// cells included (they hold the synchronised neighbour or boundary values), in
// the reference's cell order; a cell whose weight is not finite or not positive
// is left out, and a node with no weight gets 0.  One thread per node, k
// fastest, so the 8 cell reads of a wave are contiguous runs along k.
#include <hip/hip_runtime.h>

#include "mercury_output.h"

namespace {

__global__ void k_node_average(const double *__restrict__ f, const double *__restrict__ jac, int m0, int m1, int m2, int g, int sI,
                               int sJ, double *__restrict__ out) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = m0 + 1, nj = m1 + 1, nk = m2 + 1;
    if (n >= (long long)ni * nj * nk) return;
    const int k = (int)(n % nk);
    const long long t = n / nk;
    const int j = (int)(t % nj), i = (int)(t / nj);
    double wsum = 0.0, w = 0.0;
    // (i-1,j-1,k-1), (i,j-1,k-1), (i-1,j,k-1), (i,j,k-1), then the same at k
    for (int c = 0; c < 8; ++c) {
        const int ci = i - 1 + (c & 1), cj = j - 1 + ((c >> 1) & 1), ck = k - 1 + (c >> 2);
        const long long q = (long long)(ci + g) * sI + (long long)(cj + g) * sJ + (ck + g);
        const double wt = fabs(jac[q]);
        if (!__builtin_isfinite(wt) || wt <= 0.0) continue;
        wsum += wt * f[q];
        w += wt;
    }
    out[n] = (w > 0.0) ? wsum / w : 0.0;
}

}  // namespace

namespace merc {
void launch_node_average(hipStream_t st, const double *field, const double *jac, const MercBlk &b, double *out) {
    const long long n = (long long)(b.m[0] + 1) * (b.m[1] + 1) * (b.m[2] + 1);
    if (n <= 0) return;
    const int tpb = 256;
    hipLaunchKernelGGL(k_node_average, dim3((unsigned)((n + tpb - 1) / tpb)), dim3(tpb), 0, st, field, jac, b.m[0], b.m[1], b.m[2], b.g, b.sI,
                       b.sJ, out);
}
}  // namespace merc
