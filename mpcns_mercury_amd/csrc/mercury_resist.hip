// Implicit resistive correction (Z4_Mercury/src/implicit/Mercury_ImplicitResistive.cpp):
// the per-unknown kernels of the matrix-free operator and of the conjugate
// gradient iteration.  The unknowns are a list of owned edges (box index q and
// axis); the field parts of the operator (curls, synchronisation, alias mean)
// reuse the generic kernels on the Eres / dB / dJ slots.
//
// Reductions are deterministic: the reducing kernels run a capped grid (at most
// RGRID blocks) with grid-stride loops, so each thread sums a fixed set of
// unknowns in a fixed order; every block writes its tree sum, and one block adds
// the (at most RGRID) partials in block order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mercury_device.h"
#include "mercury_resist.h"

namespace {
using namespace merc_dev;
using namespace merc;

constexpr int RB = 256;     // threads per block of the unknown-vector kernels
constexpr int RGRID = 2048; // blocks of the reducing kernels (partials per reduced quantity)

// block tree sum of v into part[blockIdx.x]
DEVI void block_sum(double v, double *__restrict__ part) {
    __shared__ double sh[RB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// edge slots s.. <- x on the unknowns (x == nullptr: 0)
__global__ void k_scatter(int n, const int *__restrict__ q, const int *__restrict__ ax, const double *__restrict__ x,
                          double *const *__restrict__ F_, int s) {
    gdouble *const *__restrict__ F = gslots(F_);
    const int t = blockIdx.x * RB + threadIdx.x;
    if (t >= n) return;
    F[s + ax[t]][q[t]] = x ? x[t] : 0.0;
}
// x <- edge slots s.. on the unknowns
__global__ void k_pick(int n, const int *__restrict__ q, const int *__restrict__ ax, double *const *__restrict__ F_, int s,
                       double *__restrict__ x) {
    gdouble *const *__restrict__ F = gslots(F_);
    const int t = blockIdx.x * RB + threadIdx.x;
    if (t >= n) return;
    x[t] = F[s + ax[t]][q[t]];
}

// y = Eres - eta * dJ on the unknowns (MatMultImplicitResistive_ :449-459);
// with w != nullptr also the block partials of sum(p * y * w)
__global__ void k_gather(int n, const int *__restrict__ q, const int *__restrict__ ax, const double *__restrict__ eta,
                         double *const *__restrict__ F_, double *__restrict__ y, const double *__restrict__ p,
                         const double *__restrict__ w, double *__restrict__ part) {
    gdouble *const *__restrict__ F = gslots(F_);
    double v = 0.0;
    for (int t = blockIdx.x * RB + threadIdx.x; t < n; t += gridDim.x * RB) {
        const double yy = F[SL_ERES + ax[t]][q[t]] - eta[t] * F[SL_DJ + ax[t]][q[t]];
        y[t] = yy;
        if (w) v += p[t] * yy * w[t];
    }
    if (w) block_sum(v, part);
}

// right-hand side b = eta * J(B*) (PackImplicitJedgeToVec_ :387-404), the
// Jacobi diagonal 1 + dt*eta*alpha*sum(beta of the 4 faces around the edge),
// the inner-product weight 1/(eta*alpha) and its ratio to the diagonal (the
// weight of r.z)
__global__ void k_rhs(int n, const int *__restrict__ q, const int *__restrict__ ax, const int *__restrict__ sI,
                      const int *__restrict__ sJ, const double *__restrict__ eta, double *const *__restrict__ F_, double dt,
                      double *__restrict__ b, double *__restrict__ dg, double *__restrict__ w, double *__restrict__ wd) {
    gdouble *const *__restrict__ F = gslots(F_);
    const int t = blockIdx.x * RB + threadIdx.x;
    if (t >= n) return;
    const long long e = q[t], di = sI[t], dj = sJ[t];
    const int a = ax[t];
    const gdouble *be = F[SL_BETA], *bh = F[SL_BETA + 1], *bz = F[SL_BETA + 2];
    double s;
    if (a == 0) s = bh[e - 1] + bh[e] + bz[e] + bz[e - dj];
    else if (a == 1) s = be[e] + be[e - 1] + bz[e - di] + bz[e];
    else s = be[e - dj] + be[e] + bh[e] + bh[e - di];
    const double ea = eta[t] * F[SL_ALPHA + a][e];
    b[t] = 0.0 + eta[t] * F[SL_J + a][e];
    dg[t] = 1.0 + dt * ea * s;
    w[t] = 1.0 / ea;
    wd[t] = w[t] / dg[t];
}

// r = b - Ax, p = z = r / dg (z is not stored: it is recomputed from r and dg);
// partials of b.b, r.r and r.z.w = r.r.(w / dg)
__global__ void k_cg_init(int n, const double *__restrict__ b, const double *__restrict__ Ax, const double *__restrict__ dg,
                          const double *__restrict__ wd, double *__restrict__ r, double *__restrict__ p, double *__restrict__ part) {
    double bb = 0.0, rr = 0.0, rz = 0.0;
    for (int t = blockIdx.x * RB + threadIdx.x; t < n; t += gridDim.x * RB) {
        const double rv = b[t] - Ax[t];
        r[t] = rv;
        p[t] = rv / dg[t];
        bb += b[t] * b[t];
        rr += rv * rv;
        rz += rv * rv * wd[t];
    }
    block_sum(bb, part);
    __syncthreads();
    block_sum(rr, part + gridDim.x);
    __syncthreads();
    block_sum(rz, part + 2 * gridDim.x);
}

// r -= alpha Ap; partials of r.r and r.z.w = r.r.(w / dg).  The x update
// waits for k_cg_pxs, which reads p anyway.
__global__ void k_cg_r(int n, const double *__restrict__ S, const double *__restrict__ Ap, const double *__restrict__ wd,
                       double *__restrict__ r, double *__restrict__ part) {
    const double al = S[CG_ALPHA];
    double rr = 0.0, rz = 0.0;
    for (int t = blockIdx.x * RB + threadIdx.x; t < n; t += gridDim.x * RB) {
        const double rv = r[t] - al * Ap[t];
        r[t] = rv;
        rr += rv * rv;
        rz += rv * rv * wd[t];
    }
    block_sum(rr, part);
    __syncthreads();
    block_sum(rz, part + gridDim.x);
}

// x += alpha p, then p = z + beta p (z = r / dg), and the new p into the edge
// slots s.. of the unknowns for the next operator application: one pass over
// p for the end of one iteration and the start of the next
__global__ void k_cg_pxs(int n, const double *__restrict__ S, const double *__restrict__ r, const double *__restrict__ dg,
                         double *__restrict__ p, double *__restrict__ x, const int *__restrict__ q, const int *__restrict__ ax,
                         double *const *__restrict__ F_, int s) {
    gdouble *const *__restrict__ F = gslots(F_);
    const int t = blockIdx.x * RB + threadIdx.x;
    if (t >= n) return;
    const double po = p[t];
    x[t] += S[CG_ALPHA] * po;
    const double pn = r[t] / dg[t] + S[CG_BETA] * po;
    p[t] = pn;
    F[s + ax[t]][q[t]] = pn;
}

// sum of nq partial arrays of length nb (in block order) and the scalar update
// mode 0 (init): bb, rr, rz; mode 1: alpha = rz / pAp; mode 2: rr, beta = rz'/rz, rz = rz'
__global__ void k_cg_scalars(int nb, int nq, const double *__restrict__ part, double *__restrict__ S, int mode) {
    __shared__ double sh[RB];
    double tot[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < nq; ++c) {
        double v = 0.0;
        for (int i = threadIdx.x; i < nb; i += RB) v += part[c * nb + i];
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = RB / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        tot[c] = sh[0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        S[CG_BB] = tot[0];
        S[CG_RR] = tot[1];
        S[CG_RZ] = tot[2];
    } else if (mode == 1) {
        S[CG_PAP] = tot[0];
        S[CG_ALPHA] = S[CG_RZ] / tot[0];
    } else {
        S[CG_RR] = tot[0];
        S[CG_BETA] = tot[1] / S[CG_RZ];
        S[CG_RZ] = tot[1];
    }
}

// B += dt * RHS_Bres on owned faces (ApplyUpdate_Euler_BfaceOnly_,
// MercurySolver_MagExplicit_Hall.cpp:435-473)
__global__ void k_bres_update(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, double dt) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q;
    if (p.j < bk.m[1] && p.k < bk.m[2]) F[SL_B + 0][q] += dt * F[SL_RHSBR + 0][q];
    if (p.i < bk.m[0] && p.k < bk.m[2]) F[SL_B + 1][q] += dt * F[SL_RHSBR + 1][q];
    if (p.i < bk.m[0] && p.j < bk.m[1]) F[SL_B + 2][q] += dt * F[SL_RHSBR + 2][q];
}

}  // namespace

namespace merc {

static inline unsigned rblk(int n) { return (unsigned)((n + RB - 1) / RB); }
static inline unsigned rgrid(int n) { return std::min(rblk(n), (unsigned)RGRID); }

int res_blocks(int n) { return (int)rgrid(n); }

void launch_res_scatter(hipStream_t st, const ResDofs &d, const double *x, double *const *F, int slot) {
    if (d.n > 0) hipLaunchKernelGGL(k_scatter, dim3(rblk(d.n)), dim3(RB), 0, st, d.n, d.q, d.ax, x, F, slot);
}
void launch_res_pick(hipStream_t st, const ResDofs &d, double *const *F, int slot, double *x) {
    if (d.n > 0) hipLaunchKernelGGL(k_pick, dim3(rblk(d.n)), dim3(RB), 0, st, d.n, d.q, d.ax, F, slot, x);
}
void launch_res_gather(hipStream_t st, const ResDofs &d, double *const *F, double *y, const double *p, const double *w, double *part) {
    if (d.n > 0) hipLaunchKernelGGL(k_gather, dim3(rgrid(d.n)), dim3(RB), 0, st, d.n, d.q, d.ax, d.eta, F, y, p, w, part);
}
void launch_res_rhs(hipStream_t st, const ResDofs &d, double *const *F, double dt, double *b, double *dg, double *w, double *wd) {
    if (d.n > 0) hipLaunchKernelGGL(k_rhs, dim3(rblk(d.n)), dim3(RB), 0, st, d.n, d.q, d.ax, d.sI, d.sJ, d.eta, F, dt, b, dg, w, wd);
}
void launch_cg_init(hipStream_t st, int n, const double *b, const double *Ax, const double *dg, const double *wd, double *r, double *p,
                    double *part, double *S) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cg_init, dim3(rgrid(n)), dim3(RB), 0, st, n, b, Ax, dg, wd, r, p, part);
    hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(RB), 0, st, (int)rgrid(n), 3, part, S, 0);
}
void launch_cg_pap(hipStream_t st, int n, const double *part, double *S) {
    if (n > 0) hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(RB), 0, st, (int)rgrid(n), 1, part, S, 1);
}
void launch_cg_r(hipStream_t st, int n, const double *S, const double *Ap, const double *wd, double *r, double *part, double *Sw) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cg_r, dim3(rgrid(n)), dim3(RB), 0, st, n, S, Ap, wd, r, part);
    hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(RB), 0, st, (int)rgrid(n), 2, part, Sw, 2);
}
void launch_cg_pxs(hipStream_t st, const ResDofs &d, const double *S, const double *r, const double *dg, double *p, double *x,
                   double *const *F, int slot) {
    if (d.n > 0) hipLaunchKernelGGL(k_cg_pxs, dim3(rblk(d.n)), dim3(RB), 0, st, d.n, S, r, dg, p, x, d.q, d.ax, F, slot);
}
void launch_bres_update(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, double dt) {
    if (R.total > 0) hipLaunchKernelGGL(k_bres_update, dim3((unsigned)((R.total + 255) / 256)), dim3(256), 0, st, R.dev, B, F, dt);
}

}  // namespace merc
