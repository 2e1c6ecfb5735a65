// Device helpers shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "mercury_common.h"

#define DEVI __device__ __forceinline__

namespace merc_dev {

DEVI double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
DEVI double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min

struct Pt {
    int b, i, j, k;
    long long q;
    int sI, sJ;
};

DEVI bool locate(const MercRange R, const MercBlk *__restrict__ B, long long t, Pt &p) {
    if (t >= R.start[R.nb]) return false;
    int lo = 0, hi = R.nb - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (R.start[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int b = lo;
    unsigned r = (unsigned)(t - R.start[b]);  // per-block boxes stay below 2^31 points
    const unsigned n1 = (unsigned)R.n[3 * b + 1], n2 = (unsigned)R.n[3 * b + 2];
    const int kk = (int)(r % n2);
    r /= n2;
    const int jj = (int)(r % n1);
    const int ii = (int)(r / n1);
    const MercBlk &bk = B[b];
    p.b = b;
    p.i = R.lo[3 * b] + ii;
    p.j = R.lo[3 * b + 1] + jj;
    p.k = R.lo[3 * b + 2] + kk;
    p.sI = bk.sI;
    p.sJ = bk.sJ;
    p.q = bk.off + (long long)(p.i + bk.g) * bk.sI + (long long)(p.j + bk.g) * bk.sJ + (p.k + bk.g);
    return true;
}

// XCD-aware workgroup order: the dispatcher deals workgroups round-robin over
// the 8 XCDs, each with its own L2; remapping gives every XCD one contiguous
// slab of the index space, so the i/j stencil neighbours of a workgroup are
// mostly resident in the same L2.
DEVI unsigned xcd_block() {
    const unsigned nwg = gridDim.x, wg = blockIdx.x;
    const unsigned q = nwg >> 3, r = nwg & 7u, x = wg & 7u, s = wg >> 3;
    return x * q + (x < r ? x : r) + s;
}
DEVI long long xcd_tid() { return (long long)xcd_block() * blockDim.x + threadIdx.x; }
#define TID xcd_tid()
#define LOCATE(R)            \
    Pt p;                    \
    if (!locate(R, B, TID, p)) return;

// Field slot tables as the kernels use them: every slot pointer addresses
// global memory (a global address is the same 64-bit value as its flat
// address), and saying so makes the compiler emit global rather than flat
// loads and stores, which wait on the vector-memory counter alone.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) float gfloat;
DEVI gdouble *const *gslots(double *const *F) { return static_cast<gdouble *const *>(static_cast<const void *>(F)); }
DEVI const gfloat *const *gslots(const float *const *G) { return static_cast<const gfloat *const *>(static_cast<const void *>(G)); }

// Minimum of the threads' time-step candidates per run, folded into dt_bits[run]
// (FP64 bits, atomicMin: positive doubles order like their bit patterns).  A
// workgroup whose located threads all belong to one run issues one atomic;
// otherwise (a workgroup across a block boundary between runs) each wave whose
// threads share a run issues one, and a wave spanning two runs falls back to one
// atomic per thread.  run < 0 marks a thread with no candidate.  Every thread of
// the workgroup must call this (it synchronises the workgroup).
DEVI void dt_min_runs(double best, int run, unsigned long long *__restrict__ dt_bits) {
    __shared__ double wmin[16];
    __shared__ int wrun[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    // the wave's run: the largest (any valid one); uniform if min == max over valid lanes
    int rmax = run, rmin = run < 0 ? 0x7fffffff : run;
    for (int off = 32; off > 0; off >>= 1) {
        rmax = max(rmax, __shfl_xor(rmax, off, 64));
        rmin = min(rmin, __shfl_xor(rmin, off, 64));
    }
    const bool wave_uniform = rmax < 0 || rmin == rmax;
    if (!wave_uniform) {
        if (run >= 0 && best < 1e100) atomicMin(dt_bits + run, (unsigned long long)__double_as_longlong(best));
        best = 1e100;
    }
    for (int off = 32; off > 0; off >>= 1) best = dmin(best, __shfl_xor(best, off, 64));
    if (lane == 0) {
        wmin[wid] = best;
        wrun[wid] = wave_uniform ? rmax : -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int r0 = -1;
        bool same = true;
        for (int w = 0; w < nw; ++w) {
            if (wrun[w] < 0) continue;
            if (r0 < 0) r0 = wrun[w];
            else if (wrun[w] != r0) same = false;
        }
        if (same) {
            if (r0 >= 0) {
                double m = wmin[0];
                for (int w = 1; w < nw; ++w) m = dmin(m, wmin[w]);
                atomicMin(dt_bits + r0, (unsigned long long)__double_as_longlong(m));
            }
        } else {
            for (int w = 0; w < nw; ++w)
                if (wrun[w] >= 0) atomicMin(dt_bits + wrun[w], (unsigned long long)__double_as_longlong(wmin[w]));
        }
    }
}

// neighbour offsets
#define OI(p) ((long long)(p).sI)
#define OJ(p) ((long long)(p).sJ)
#define AX(p, a) ((a) == 0 ? OI(p) : ((a) == 1 ? OJ(p) : 1LL))

// n: number densities, q: charge densities Z n (q = n for the singly charged
// H+ and Na+), chi = q / n_e (the weights of U+), w = q / ne_eff
struct Num {
    double n[MERC_NSMAX], q[MERC_NSMAX], ne, ne_eff, chi[MERC_NSMAX], w[MERC_NSMAX];
};

// ion species of a run: 2 (H+, Na+) or 3 (with He++)
DEVI int nspec(const MercPhys &P) { return P.n_species >= 3.0 ? 3 : 2; }

// Hall_Num_Limiter (Z4_Mercury/src/physics/MercurySolver_Tools.cpp:111-157); with
// He++ its charge density Z_He n_He joins n_e last (no reference counterpart)
DEVI Num num_limiter(const MercPhys &P, double rhoH, double rhoNa, double rhoHe = 0.0, int ns = 2) {
    Num n;
    n.n[0] = dmax(rhoH, 0.0) / P.M_H * P.M_ref;
    n.n[1] = dmax(rhoNa, 0.0) / P.M_Na * P.M_ref;
    n.q[0] = n.n[0];
    n.q[1] = n.n[1];
    n.ne = n.q[0] + n.q[1];
    n.n[2] = n.q[2] = 0.0;
    if (ns > 2) {
        n.n[2] = dmax(rhoHe, 0.0) / P.M_He * P.M_ref;
        n.q[2] = P.Z_He * n.n[2];
        n.ne = n.ne + n.q[2];
    }
    n.ne_eff = sqrt(n.ne * n.ne + P.ne_floor * P.ne_floor);
    // loops over the fixed maximum with a guard: unrolled, every index is a
    // constant and the arrays stay in registers (a loop bounded by ns would
    // index them dynamically, through scratch memory)
    n.chi[2] = n.w[2] = 0.0;
    if (n.ne > 1.0e-300) {
#pragma unroll
        for (int s = 0; s < MERC_NSMAX; ++s)
            if (s < ns) n.chi[s] = n.q[s] / n.ne;
    } else {
        n.chi[0] = 1.0;
#pragma unroll
        for (int s = 1; s < MERC_NSMAX; ++s)
            if (s < ns) n.chi[s] = 0.0;
    }
#pragma unroll
    for (int s = 0; s < MERC_NSMAX; ++s)
        if (s < ns) n.w[s] = n.q[s] / n.ne_eff;
    return n;
}

// the limiter at one point from the state slots
template <typename FT> DEVI Num num_at(const MercPhys &P, const FT &F, long long q, int ns) {
    return num_limiter(P, F[SL_UH][q], F[SL_UN][q], ns > 2 ? (double)F[SL_UHE][q] : 0.0, ns);
}

// electron pressure stand-in of the ambipolar terms: the species pressures'
// sum at a cell (p_H + p_Na in the reference, + p_He with He++)
template <typename FT> DEVI double pe_sum(const FT &F, long long q, int ns) {
    double pe = F[SL_PVH + 3][q] + F[SL_PVN + 3][q];
    if (ns > 2) pe = pe + F[SL_PVHE + 3][q];
    return pe;
}

// molar mass and state coefficient of species s
DEVI double species_mass(const MercPhys &P, int s) { return s == 0 ? P.M_H : (s == 1 ? P.M_Na : P.M_He); }
DEVI double species_coeff(const MercPhys &P, int s) { return s == 0 ? P.state_coeff_H : (s == 1 ? P.state_coeff_Na : P.state_coeff_He); }

DEVI void rusanov_face(const gdouble *__restrict__ const *F, int su, int sp, long long qL, long long qR, long long qf, int a,
                       double gm, double out[5]) {
    const double k1 = F[SL_JD + 3 * a][qf], k2 = F[SL_JD + 3 * a + 1][qf], k3 = F[SL_JD + 3 * a + 2][qf];
    double Uc[2][5], Fl[2][5], rad[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const long long qq = side ? qR : qL;
        const double rho = F[su][qq], u = F[sp][qq], v = F[sp + 1][qq], w = F[sp + 2][qq], pr = F[sp + 3][qq];
        Uc[side][0] = rho;
        Uc[side][1] = rho * u;
        Uc[side][2] = rho * v;
        Uc[side][3] = rho * w;
        Uc[side][4] = pr / (gm - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
        const double uvw = k1 * u + k2 * v + k3 * w;
        const double kk2 = (k1 * k1 + k2 * k2 + k3 * k3);
        const double c = sqrt(dmax(0.0, gm * pr / dmax(rho, 1e-30)));
        rad[side] = fabs(uvw) + c * sqrt(kk2);
        const double rhoe = pr / (gm - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
        Fl[side][0] = rho * uvw;
        Fl[side][1] = rho * uvw * u + k1 * pr;
        Fl[side][2] = rho * uvw * v + k2 * pr;
        Fl[side][3] = rho * uvw * w + k3 * pr;
        Fl[side][4] = uvw * (rhoe + pr);
    }
    const double r = dmax(rad[0], rad[1]);
#pragma unroll
    for (int m = 0; m < 5; ++m) out[m] = 0.5 * (Fl[0][m] + Fl[1][m]) - 0.5 * r * (Uc[1][m] - Uc[0][m]);
}

DEVI double u_contra(gdouble *const *__restrict__ F, const Pt &p, int axis, long long c) {
    const long long cp = c + AX(p, axis);
    const int s0 = SL_JD + 3 * axis;
    const double kx = 0.5 * (F[s0][c] + F[s0][cp]);
    const double ky = 0.5 * (F[s0 + 1][c] + F[s0 + 1][cp]);
    const double kz = 0.5 * (F[s0 + 2][c] + F[s0 + 2][cp]);
    const double inv_jac = 1.0 / (fabs(F[SL_JAC][c]) + 1.0e-30);
    return (kx * F[SL_UP][c] + ky * F[SL_UP + 1][c] + kz * F[SL_UP + 2][c]) * inv_jac;
}
DEVI double beta_ind(gdouble *const *__restrict__ F, const Pt &p, int axis, long long c) {
    return 0.5 * (F[SL_B + axis][c] + F[SL_B + axis][c + AX(p, axis)]);
}
DEVI double beta_tot(gdouble *const *__restrict__ F, const Pt &p, int axis, long long c) {
    const long long cp = c + AX(p, axis);
    return 0.5 * ((F[SL_B + axis][c] + F[SL_BADD + axis][c]) + (F[SL_B + axis][cp] + F[SL_BADD + axis][cp]));
}
DEVI double emf_flux(gdouble *const *__restrict__ F, const Pt &p, int ba, int fa, long long L, long long Rr, double bff) {
    const double btL = beta_tot(F, p, ba, L), btR = beta_tot(F, p, ba, Rr);
    const double biL = beta_ind(F, p, ba, L), biR = beta_ind(F, p, ba, Rr);
    const double ufL = u_contra(F, p, fa, L), ufR = u_contra(F, p, fa, Rr);
    const double ubL = u_contra(F, p, ba, L), ubR = u_contra(F, p, ba, Rr);
    const double GL = ufL * btL - ubL * bff, GR = ufR * btR - ubR * bff;
    const double radius = dmax(fabs(ufL), fabs(ufR));
    return 0.5 * (GL + GR) - 0.5 * radius * (biR - biL);
}

DEVI double maxu4(gdouble *const *__restrict__ F, const Pt &p, int ax, long long c0, long long c1, long long c2, long long c3) {
    double a = 0.0;
    a = dmax(a, fabs(u_contra(F, p, ax, c0)));
    a = dmax(a, fabs(u_contra(F, p, ax, c1)));
    a = dmax(a, fabs(u_contra(F, p, ax, c2)));
    a = dmax(a, fabs(u_contra(F, p, ax, c3)));
    return a;
}

}  // namespace merc_dev
