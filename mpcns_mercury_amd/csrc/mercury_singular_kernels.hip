// Singular-edge assembly: one thread per singular edge (there are a few per
// corner line of a cubed sphere, so these launches are latency, not bandwidth).
// Sums run in the registry's order, which is the reference's, so the FP64
// results equal the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include "mercury_device.h"
#include "mercury_singular_dev.h"

namespace {
using namespace merc_dev;

// AssembleSingularEdgeEMF_NonHall_ (Z4_Mercury/src/physics/MercurySolver_MagExplicit_Ideal.cpp:57-130)
// and AssembleSingularEdgeEMF_HallExplicit_ (MercurySolver_MagExplicit_Hall.cpp:32-56)
// through assemble_cell_field_to_local_owners (src/7_metric/SingularEdgeRegistry.cpp:288-310):
// owner = owner sign * sum over the cells of weight * the cell's EMF along the
// mean edge vector.  The electron pressure and the number limiter take the H+
// and Na+ species, as there.
template <bool MIX>
__global__ void __launch_bounds__(64) k_singular_emf(int n, const MercSingEntry *__restrict__ E, const MercSingCell *__restrict__ C,
                                                     const MercBlk *__restrict__ B, double *const *F_, const float *const *G_,
                                                     MercPhys P) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const MercSingEntry x = E[e];
    const double cv0 = x.cv[0], cv1 = x.cv[1], cv2 = x.cv[2];
    double s = 0.0, sh = 0.0;
    for (int t = x.c0; t < x.c1; ++t) {
        const MercSingCell c = C[t];
        const long long q = c.q;
        double u0, u1, u2;
        if constexpr (MIX) u0 = G[SL_UP][q], u1 = G[SL_UP + 1][q], u2 = G[SL_UP + 2][q];
        else u0 = F[SL_UP][q], u1 = F[SL_UP + 1][q], u2 = F[SL_UP + 2][q];
        const double b0 = F[SL_BC][q], b1 = F[SL_BC + 1][q], b2 = F[SL_BC + 2][q];
        const double j0 = F[SL_JC][q], j1 = F[SL_JC + 1][q], j2 = F[SL_JC + 2][q];
        const Num num = num_limiter(P, F[SL_UH][q], F[SL_UN][q]);
        const double ex = -(u1 * b2 - u2 * b1), ey = -(u2 * b0 - u0 * b2), ez = -(u0 * b1 - u1 * b0);
        double emf = ex * cv0 + ey * cv1 + ez * cv2;
        if (P.ambipolar_enabled != 0.0) {
            const long long d = c.d;
            const double dp = 0.5 * ((F[SL_PVH + 3][q + d] + F[SL_PVN + 3][q + d]) - (F[SL_PVH + 3][q - d] + F[SL_PVN + 3][q - d]));
            emf -= P.pressure_coef * c.orient * dp / num.ne_eff;
        }
        const double jedge = j0 * cv0 + j1 * cv1 + j2 * cv2;
        const double jmag = sqrt(j0 * j0 + j1 * j1 + j2 * j2);
        if (P.arti_eta_max > 0.0) {
            double qq = (jmag - P.J_start) / dmax(P.J_on - P.J_start, 1.e-30);
            qq = dmax(0.0, dmin(1.0, qq));
            emf += P.arti_eta_max * qq * qq * (3.0 - 2.0 * qq) * jedge;
        }
        if (c.leta_on) emf += c.leta * jedge;
        s += c.w * emf;
        const double alpha = B[c.blk].hall_coef / num.ne_eff;
        const double hx = alpha * (j1 * b2 - j2 * b1), hy = alpha * (j2 * b0 - j0 * b2), hz = alpha * (j0 * b1 - j1 * b0);
        sh += c.w * (c.taper * (hx * cv0 + hy * cv1 + hz * cv2));
    }
    F[SL_E + x.oax][x.oq] = x.osign * s;
    F[SL_EH + x.oax][x.oq] = x.osign * sh;
}

// AssembleSingularEdgeCurrent_ (Z4_Mercury/src/physics/MercurySolver_Current.cpp:59-138)
// through assemble_face_triplet_to_local_owners (SingularEdgeRegistry.cpp:312-330):
// owner = owner sign * sum over the owner faces of coef * beta * B
__global__ void __launch_bounds__(64) k_singular_current(int n, const MercSingEntry *__restrict__ E, const MercSingFace *__restrict__ Fc,
                                                         double *const *F_, int sB, int sJ) {
    gdouble *const *__restrict__ F = gslots(F_);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const MercSingEntry x = E[e];
    double s = 0.0;
    for (int t = x.f0; t < x.f1; ++t) {
        const MercSingFace f = Fc[t];
        s += f.coef * F[SL_BETA + f.ax][f.q] * F[sB + f.ax][f.q];
    }
    F[sJ + x.oax][x.oq] = x.osign * s;
}

// PrepareTecplotSingularNodeAverages_ (src/6_io/IOModule_Tecplot.cpp:528-609): the
// |Jac|-weighted mean of a cell field over a singular node's cells, in stencil
// order; cells of non-finite or non-positive |Jac| are skipped, no weight gives 0
__global__ void __launch_bounds__(64) k_node_stencil_mean(int n, const int *__restrict__ start, const int *__restrict__ cq,
                                                          const double *__restrict__ v, const double *__restrict__ jac,
                                                          double *__restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double sum = 0.0, wsum = 0.0;
    for (int t = start[s]; t < start[s + 1]; ++t) {
        const int q = cq[t];
        const double w = fabs(jac[q]);
        if (!isfinite(w) || w <= 0.0) continue;
        sum += w * v[q];
        wsum += w;
    }
    out[s] = wsum > 0.0 ? sum / wsum : 0.0;
}

}  // namespace

void launch_node_stencil_mean(hipStream_t s, int n, const int *start, const int *cq, const double *field, const double *jac,
                              double *out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_node_stencil_mean, dim3((n + 63) / 64), dim3(64), 0, s, n, start, cq, field, jac, out);
}

void launch_singular_emf(hipStream_t s, int n, const MercSingEntry *E, const MercSingCell *C, const MercBlk *B, double *const *F,
                         const float *const *G, const MercPhys &P, bool mix) {
    if (n <= 0) return;
    const int nb = (n + 63) / 64;
    if (mix) hipLaunchKernelGGL(k_singular_emf<true>, dim3(nb), dim3(64), 0, s, n, E, C, B, F, G, P);
    else hipLaunchKernelGGL(k_singular_emf<false>, dim3(nb), dim3(64), 0, s, n, E, C, B, F, G, P);
}

void launch_singular_current(hipStream_t s, int n, const MercSingEntry *E, const MercSingFace *Fc, double *const *F, int sB, int sJ) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_singular_current, dim3((n + 63) / 64), dim3(64), 0, s, n, E, Fc, F, sB, sJ);
}
