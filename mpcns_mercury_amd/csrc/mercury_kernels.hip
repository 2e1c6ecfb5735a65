// HIP kernels of the Mercury explicit step for gfx950 (FP64 form).
//
// Every kernel restates one reference loop nest and keeps its floating-point
// operation order; the library is compiled with -ffp-contract=off so that the
// device results equal the CPU restatement bit for bit wherever both execute
// the same IEEE operations (only correctly rounded +,-,*,/,sqrt appear here;
// the transcendental static inputs are precomputed on the host).
//
// Work distribution: a MercRange lists, per block, a box of (i,j,k); thread t
// finds its block by binary search over the prefix sums and unravels (i,j,k)
// with k fastest, so a wave64 reads 64 consecutive doubles of each SoA plane.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "mercury_common.h"
#include "mercury_kernels.h"

#include "mercury_device.h"

namespace {
using namespace merc_dev;

// ---------------------------------------------------------------------------
// derived fields
// ---------------------------------------------------------------------------

// calc_PV + calc_Uplus (MercurySolver_Tools.cpp:159-301), whole box
__global__ void k_pv_uplus(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, const MercPhys P) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q;
    const int ns = nspec(P);
    double pv[MERC_NSMAX][5];
#pragma unroll
    for (int s = 0; s < MERC_NSMAX && s < ns; ++s) {
        const int su = SPEC_U(s);
        const double coeff = species_coeff(P, s);
        const double rho = dmax(F[su][q], 1e-12);
        const double u = F[su + 1][q] / rho, v = F[su + 2][q] / rho, w = F[su + 3][q] / rho;
        const double E = F[su + 4][q];
        const double ke = 0.5 * rho * (u * u + v * v + w * w);
        double eint = E - ke;
        if (eint < 0.0) eint = 0.0;
        double pr = (P.gamma - 1.0) * eint;
        pr = dmax(pr, 1e-12);
        pv[s][0] = u; pv[s][1] = v; pv[s][2] = w; pv[s][3] = pr; pv[s][4] = pr / (rho * coeff);
        const int sp = SPEC_PV(s);
        for (int c = 0; c < 5; ++c) F[sp + c][q] = pv[s][c];
    }
    double r[MERC_NSMAX] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < MERC_NSMAX && s < ns; ++s) r[s] = dmax(F[SPEC_U(s)][q], 0.0);
    const Num n = num_limiter(P, r[0], r[1], r[2], ns);
    for (int c = 0; c < 3; ++c) {
        const double uH = (r[0] > 1e-20) ? pv[0][c] : 0.0;
        const double uN = (r[1] > 1e-20) ? pv[1][c] : 0.0;
        double up = n.chi[0] * uH + n.chi[1] * uN;
        if (ns > 2) up = up + n.chi[2] * ((r[2] > 1e-20) ? pv[2][c] : 0.0);
        F[SL_UP + c][q] = up;
    }
}

// calc_Bcell (MercurySolver_MagneticCell.cpp:3-81), owned cells
__global__ void k_bcell(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q, di = OI(p), dj = OJ(p);
    double pt[6], pi[6];
    pt[0] = -F[SL_B + 0][q] - F[SL_BADD + 0][q];
    pt[1] = F[SL_B + 0][q + di] + F[SL_BADD + 0][q + di];
    pt[2] = -F[SL_B + 1][q] - F[SL_BADD + 1][q];
    pt[3] = F[SL_B + 1][q + dj] + F[SL_BADD + 1][q + dj];
    pt[4] = -F[SL_B + 2][q] - F[SL_BADD + 2][q];
    pt[5] = F[SL_B + 2][q + 1] + F[SL_BADD + 2][q + 1];
    pi[0] = -F[SL_B + 0][q];
    pi[1] = F[SL_B + 0][q + di];
    pi[2] = -F[SL_B + 1][q];
    pi[3] = F[SL_B + 1][q + dj];
    pi[4] = -F[SL_B + 2][q];
    pi[5] = F[SL_B + 2][q + 1];
    double bt[3] = {0, 0, 0}, bi[3] = {0, 0, 0};
#pragma unroll
    for (int n = 0; n < 6; ++n)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double w = F[SL_WB + 6 * c + n][q];
            bt[c] += w * pt[n];
            bi[c] += w * pi[n];
        }
    for (int c = 0; c < 3; ++c) {
        F[SL_BC + c][q] = bt[c];
        F[SL_BI + c][q] = bi[c];
    }
}

// CTOperators::CurlAdjFaceToEdge (Z4_Mercury/src/operators/CTOperators.cpp:59-110)
// followed by J *= alpha (MercurySolver_Current.cpp:140-201); owned edges of the
// three axes, one thread per node-box point
// (face slots sf.., edge slots se..: B -> J, or dB -> dJ of the implicit correction)
__global__ void k_jedge(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, int sf, int se) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
#define BB(a, o) (F[SL_BETA + (a)][q + (o)] * F[sf + (a)][q + (o)])
    if (p.i < bk.m[0]) {
        const double v = (BB(1, -dk) - BB(1, 0)) + (BB(2, 0) - BB(2, -dj));
        F[se + 0][q] = (1.0 * v) * F[SL_ALPHA + 0][q];
    }
    if (p.j < bk.m[1]) {
        const double v = (BB(0, 0) - BB(0, -dk)) + (BB(2, -di) - BB(2, 0));
        F[se + 1][q] = (1.0 * v) * F[SL_ALPHA + 1][q];
    }
    if (p.k < bk.m[2]) {
        const double v = (BB(0, -dj) - BB(0, 0)) + (BB(1, 0) - BB(1, -di));
        F[se + 2][q] = (1.0 * v) * F[SL_ALPHA + 2][q];
    }
#undef BB
}

// calc_Jcell (MercurySolver_Current.cpp:203-259), owned cells
__global__ void k_jcell(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q, di = OI(p), dj = OJ(p);
    const double s[12] = {F[SL_J + 0][q], F[SL_J + 0][q + dj], F[SL_J + 0][q + 1], F[SL_J + 0][q + dj + 1],
                          F[SL_J + 1][q], F[SL_J + 1][q + di], F[SL_J + 1][q + 1], F[SL_J + 1][q + di + 1],
                          F[SL_J + 2][q], F[SL_J + 2][q + di], F[SL_J + 2][q + dj], F[SL_J + 2][q + di + dj]};
    double J[3] = {0, 0, 0};
#pragma unroll
    for (int n = 0; n < 12; ++n)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[c] += F[SL_WJ + 12 * c + n][q] * s[n];
    for (int c = 0; c < 3; ++c) F[SL_JC + c][q] = J[c];
}

// ---------------------------------------------------------------------------
// time step (MercurySolver_Timestep.cpp:8-141): per owned cell and species the
// CFL bound, reduced to one global minimum.  All candidates are positive, so
// the IEEE bit patterns order like the values and an unsigned 64-bit atomic
// minimum gives the exact global minimum independent of the order.
// ---------------------------------------------------------------------------
__global__ void k_timestep(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, const MercPhys P,
                           unsigned long long *__restrict__ dt_bits) {
    gdouble *const *__restrict__ F = gslots(F_);
    double best = 1e100;
    int run = -1;
    Pt p;
    if (locate(R, B, TID, p)) {
        const long long q = p.q;
        run = B[p.b].run;
        const double V = fabs(F[SL_JAC][q]);
        if (!(V <= 0.0)) {
            const double Bx = F[SL_BC][q], By = F[SL_BC + 1][q], Bz = F[SL_BC + 2][q];
            const double B2 = Bx * Bx + By * By + Bz * Bz;
#pragma unroll
            for (int s = 0, ns_ = nspec(P); s < MERC_NSMAX && s < ns_; ++s) {
                const int su = SPEC_U(s);
                const double rho = dmax(F[su][q], 1e-2);
                const double ux = F[su + 1][q] / rho, uy = F[su + 2][q] / rho, uz = F[su + 3][q] / rho;
                const double E = F[su + 4][q];
                const double ke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
                const double pr = (P.gamma - 1.0) * (E - ke);
                const double cs2 = P.gamma * pr / rho;
                double denom = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int sd = 0; sd < 2; ++sd) {
                        const long long fq = q + (sd ? AX(p, a) : 0LL);
                        const double ax = F[SL_JD + 3 * a][fq], ay = F[SL_JD + 3 * a + 1][fq], az = F[SL_JD + 3 * a + 2][fq];
                        const double A = sqrt(ax * ax + ay * ay + az * az);
                        double term = 0.0;
                        if (!(A <= 0.0)) {
                            double nx = ax / A, ny = ay / A, nz = az / A;
                            if (!sd) { nx = -nx; ny = -ny; nz = -nz; }
                            const double un = ux * nx + uy * ny + uz * nz;
                            const double Bn = Bx * nx + By * ny + Bz * nz;
                            const double vA2 = P.inver_MA2 * B2 / rho;
                            const double vAn2 = P.inver_MA2 * (Bn * Bn) / rho;
                            const double t = cs2 + vA2;
                            double disc = t * t - 4.0 * cs2 * vAn2;
                            if (disc < 0.0) disc = 0.0;
                            const double cf = sqrt(0.5 * (t + sqrt(disc)));
                            term = (fabs(un) + cf) * A;
                        }
                        denom += term;
                    }
                if (denom > 0.0) best = dmin(best, P.cfl * V / denom);
            }
        }
    }
    // wave64 then workgroup minimum, per run
    dt_min_runs(best, run, dt_bits);
}

// ---------------------------------------------------------------------------
// fluid RHS: Scheme_U_ (MercurySolver_FluidRHS.cpp:3-202).  The flux of a face
// depends only on its two cells and its area vector, so each owned cell
// evaluates its six faces itself (the same operations as the face sweep) and
// accumulates -dF/Jac in the reference's xi, eta, zeta order; no face flux
// array round trip.
// ---------------------------------------------------------------------------
__global__ void k_flux_div(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, const MercPhys P) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q;
    const double gm = P.gamma;
    const double invJ = 1.0 / F[SL_JAC][q];
#pragma unroll
    for (int s = 0, ns_ = nspec(P); s < MERC_NSMAX && s < ns_; ++s) {
        const int su = SPEC_U(s), sp = SPEC_PV(s), sr = SPEC_R(s);
        double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) {
            const long long d = AX(p, a);
            double Fm[5], Fp[5];
            rusanov_face(F, su, sp, q - d, q, q, a, gm, Fm);
            rusanov_face(F, su, sp, q, q + d, q + d, a, gm, Fp);
#pragma unroll
            for (int m = 0; m < 5; ++m) r[m] -= (Fp[m] - Fm[m]) * invJ;
        }
        for (int m = 0; m < 5; ++m) F[sr + m][q] = r[m];
    }
}

// AddSourceToRHS_Fluid (MercurySolver_FluidRHS.cpp:206-527), owned cells
__global__ void k_sources(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, const MercPhys P) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q, di = OI(p), dj = OJ(p);
    const double upx = F[SL_UP][q], upy = F[SL_UP + 1][q], upz = F[SL_UP + 2][q];
    const double Bx = F[SL_BC][q], By = F[SL_BC + 1][q], Bz = F[SL_BC + 2][q];
    const double Jx = F[SL_JC][q], Jy = F[SL_JC + 1][q], Jz = F[SL_JC + 2][q];
    const double sjbx = Jy * Bz - Jz * By, sjby = Jz * Bx - Jx * Bz, sjbz = Jx * By - Jy * Bx;
    const int ns = nspec(P);
    const Num num = num_at(P, F, q, ns);
    double dpex = 0.0, dpey = 0.0, dpez = 0.0;
    const bool ambi = P.ambipolar_enabled != 0.0;
    if (ambi) {
        const double V = fabs(F[SL_JAC][q]);
        double d[9];
        if (V <= 0.0) {
            for (int t = 0; t < 9; ++t) d[t] = 0.0;
        } else {
            for (int c = 0; c < 3; ++c) {
                d[c] = 0.5 * (F[SL_JD + c][q - di] + F[SL_JD + c][q]) / V;
                d[3 + c] = 0.5 * (F[SL_JD + 3 + c][q - dj] + F[SL_JD + 3 + c][q]) / V;
                d[6 + c] = 0.5 * (F[SL_JD + 6 + c][q - 1] + F[SL_JD + 6 + c][q]) / V;
            }
        }
#define PE(o) pe_sum(F, q + (o), ns)
        const double pip = PE(di), pim = PE(-di), pjp = PE(dj), pjm = PE(-dj), pkp = PE(1), pkm = PE(-1);
#undef PE
        dpex = 0.5 * (d[0] * (pip - pim) + d[3] * (pjp - pjm) + d[6] * (pkp - pkm));
        dpey = 0.5 * (d[1] * (pip - pim) + d[4] * (pjp - pjm) + d[7] * (pkp - pkm));
        dpez = 0.5 * (d[2] * (pip - pim) + d[5] * (pjp - pjm) + d[8] * (pkp - pkm));
    }
    const double sss = F[SL_PHOTO][q];
    const double x = F[SL_CC][q], y = F[SL_CC + 1][q], z = F[SL_CC + 2][q];
    const double rho_cyl = sqrt(y * y + z * z);
    const bool in_shadow = (x < 0.0 && rho_cyl < 1.0);
    const double vst = in_shadow ? P.nu_shadow : P.nu_lit;
#pragma unroll
    for (int s = 0; s < MERC_NSMAX && s < ns; ++s) {
        const int su = SPEC_U(s), sp = SPEC_PV(s), sr = SPEC_R(s);
        const double rho = dmax(F[su][q], 0.0);
        const double u = F[sp][q], v = F[sp + 1][q], w = F[sp + 2][q];
        const double u2 = u * u + v * v + w * w;
        const double subx = (v - upy) * Bz - (w - upz) * By;
        const double suby = (w - upz) * Bx - (u - upx) * Bz;
        const double subz = (u - upx) * By - (v - upy) * Bx;
        const double subu = subx * u + suby * v + subz * w;
        const double sjbu = sjbx * u + sjby * v + sjbz * w;
        const double dpeu = dpex * u + dpey * v + dpez * w;
        const double Ts0 = F[sp + 4][q] * P.T_ref;
        const double Ms = species_mass(P, s);
        const double Mo = P.M_Na;
        const double b1 = vst * (P.Tn0 - Ts0) * (Ms / (Ms + Mo));
        const double b2 = (Mo / (Ms + Mo)) * vst;
        // number density in the thermal exchange, charge density in the induction force
        const double nsp = num.n[s], qs = num.q[s];
        const double chi = num.w[s];
        double r0 = F[sr][q], r1 = F[sr + 1][q], r2 = F[sr + 2][q], r3 = F[sr + 3][q], r4 = F[sr + 4][q];
        r1 += P.c_ind * qs * subx + P.c_hall * chi * sjbx;
        r2 += P.c_ind * qs * suby + P.c_hall * chi * sjby;
        r3 += P.c_ind * qs * subz + P.c_hall * chi * sjbz;
        r4 += P.c_ind * qs * subu + P.c_hall * chi * sjbu;
        if (ambi) {
            r1 -= chi * dpex;
            r2 -= chi * dpey;
            r3 -= chi * dpez;
            r4 -= chi * dpeu;
        }
        if (s == 1) {
            r0 += P.a1_Na * sss;
            r4 += P.a6 * sss * P.Tn0;
        }
        r1 -= P.a4 * rho * u * vst;
        r2 -= P.a4 * rho * v * vst;
        r3 -= P.a4 * rho * w * vst;
        r4 -= P.a4 * rho * u2 * vst;
        r4 += P.a5 * nsp * b1;
        r4 += P.a4 * rho * u2 * b2;
        F[sr][q] = r0; F[sr + 1][q] = r1; F[sr + 2][q] = r2; F[sr + 3][q] = r3; F[sr + 4][q] = r4;
    }
}

// ---------------------------------------------------------------------------
// ideal EMF: AssembleOneDirectionEMF_ (MercurySolver_MagExplicit_Ideal.cpp:433-594)
// ---------------------------------------------------------------------------
// face candidates on owned faces +-1 (clamped to the allocation +-1); range is
// the box [-1, m+2)^3 and each axis keeps its own valid sub-range
__global__ void k_eface(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const int idx[3] = {p.i, p.j, p.k};
    for (int fa = 0; fa < 3; ++fa) {
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const int ohi = bk.m[a] + (a == fa ? 1 : 0);
            const int lo = (-1 > -bk.g + 1) ? -1 : -bk.g + 1;
            const int hi = (ohi + 1 < ohi + bk.g - 1) ? ohi + 1 : ohi + bk.g - 1;
            if (idx[a] < lo || idx[a] >= hi) ok = false;
        }
        if (!ok) continue;
        const long long Rq = p.q, Lq = p.q - AX(p, fa);
        const double bff = F[SL_B + fa][Rq] + F[SL_BADD + fa][Rq];
        double e0 = 0.0, e1 = 0.0, e2 = 0.0;
        if (fa == 0) {
            const double Gex = emf_flux(F, p, 1, 0, Lq, Rq, bff);
            const double Gzx = emf_flux(F, p, 2, 0, Lq, Rq, bff);
            e1 = Gzx;
            e2 = -Gex;
        } else if (fa == 1) {
            const double Gxe = emf_flux(F, p, 0, 1, Lq, Rq, bff);
            const double Gze = emf_flux(F, p, 2, 1, Lq, Rq, bff);
            e0 = -Gze;
            e2 = Gxe;
        } else {
            const double Gxz = emf_flux(F, p, 0, 2, Lq, Rq, bff);
            const double Gez = emf_flux(F, p, 1, 2, Lq, Rq, bff);
            e0 = Gez;
            e1 = -Gxz;
        }
        F[SL_EF + 3 * fa][Rq] = e0;
        F[SL_EF + 3 * fa + 1][Rq] = e1;
        F[SL_EF + 3 * fa + 2][Rq] = e2;
    }
}

// Edge EMF on owned edges: AssembleEdgeEMF_FromFaceE_Ideal_ (Ideal.cpp:138-372,
// C_uct = 0.5), then in the reference's order AddAmbipolarEdgeEMF_
// (Ambipo.cpp:3-143), AddArtificialResistivityToEdgeEMF_ (Resis.cpp:6-65) and
// AddLocalArtificialResistivityToEdgeEMF_ (Resis.cpp:66-185)
__global__ void k_edge_emf(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, const MercPhys P) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    const double Cuct = 0.5;
    const bool ambi = P.ambipolar_enabled != 0.0;
    const bool arti = P.arti_eta_max > 0.0;
    const int ns = nspec(P);
    const bool local = P.local_enabled != 0.0 && P.local_eta_max > 0.0 && P.local_r_decay > 0.0;
    for (int a = 0; a < 3; ++a) {
        if ((a == 0 && p.i >= bk.m[0]) || (a == 1 && p.j >= bk.m[1]) || (a == 2 && p.k >= bk.m[2])) continue;
        double E;
        if (a == 0) {
            const double Ec = 0.25 * (F[SL_EF + 3][q] + F[SL_EF + 3][q - dk] + F[SL_EF + 6][q] + F[SL_EF + 6][q - dj]);
            const double ae = maxu4(F, p, 1, q, q - dj, q - dk, q - dj - dk);
            const double az = maxu4(F, p, 2, q, q - dj, q - dk, q - dj - dk);
            const double dBz = F[SL_B + 2][q] - F[SL_B + 2][q - dj];
            const double dBe = F[SL_B + 1][q] - F[SL_B + 1][q - dk];
            E = Ec + Cuct * (0.5 * ae * dBz - 0.5 * az * dBe);
        } else if (a == 1) {
            const double Ec = 0.25 * (F[SL_EF + 1][q] + F[SL_EF + 1][q - dk] + F[SL_EF + 7][q] + F[SL_EF + 7][q - di]);
            const double az = maxu4(F, p, 2, q, q - di, q - dk, q - di - dk);
            const double ax = maxu4(F, p, 0, q, q - di, q - dk, q - di - dk);
            const double dBx = F[SL_B + 0][q] - F[SL_B + 0][q - dk];
            const double dBz = F[SL_B + 2][q] - F[SL_B + 2][q - di];
            E = Ec + Cuct * (0.5 * az * dBx - 0.5 * ax * dBz);
        } else {
            const double Ec = 0.25 * (F[SL_EF + 2][q] + F[SL_EF + 2][q - dj] + F[SL_EF + 5][q] + F[SL_EF + 5][q - di]);
            const double ax = maxu4(F, p, 0, q, q - di, q - dj, q - di - dj);
            const double ae = maxu4(F, p, 1, q, q - di, q - dj, q - di - dj);
            const double dBe = F[SL_B + 1][q] - F[SL_B + 1][q - di];
            const double dBx = F[SL_B + 0][q] - F[SL_B + 0][q - dj];
            E = Ec + Cuct * (0.5 * ax * dBe - 0.5 * ae * dBx);
        }
        if (ambi) {
            const long long ea = AX(p, a);
            double pe[2];
            for (int e = 0; e < 2; ++e) {
                const long long n0 = q + (e ? ea : 0);
                double sum = 0.0;
                for (int ddi = -1; ddi <= 0; ++ddi)
                    for (int ddj = -1; ddj <= 0; ++ddj)
                        for (int ddk = -1; ddk <= 0; ++ddk) {
                            const long long c = n0 + ddi * di + ddj * dj + ddk;
                            sum += pe_sum(F, c, ns);
                        }
                pe[e] = 0.125 * sum;
            }
            double rH = 0.0, rN = 0.0, rHe = 0.0;
            const long long t1 = (a == 0) ? dj : di, t2 = (a == 2) ? dj : 1;
            for (int d1 = -1; d1 <= 0; ++d1)
                for (int d2 = -1; d2 <= 0; ++d2) {
                    const long long c = q + d1 * t1 + d2 * t2;
                    rH += F[SL_UH][c];
                    rN += F[SL_UN][c];
                    if (ns > 2) rHe += F[SL_UHE][c];
                }
            const Num n = num_limiter(P, 0.25 * rH, 0.25 * rN, 0.25 * rHe, ns);
            E -= P.pressure_coef * (pe[1] - pe[0]) / n.ne_eff;
        }
        if (arti) {
            const double J1 = F[SL_J + a][q];
            const double dx = F[SL_DR + 3 * a][q], dy = F[SL_DR + 3 * a + 1][q], dz = F[SL_DR + 3 * a + 2][q];
            const double drl = sqrt(dx * dx + dy * dy + dz * dz);
            const double Jm = fabs(J1) / dmax(drl, 1.0e-30);
            double eta;
            if (P.J_on <= P.J_start)
                eta = (Jm > P.J_start) ? P.arti_eta_max : 0.0;
            else {
                double qq = (Jm - P.J_start) / (P.J_on - P.J_start);
                qq = dmax(0.0, dmin(1.0, qq));
                eta = P.arti_eta_max * (qq * qq * (3.0 - 2.0 * qq));
            }
            E += eta * J1;
        }
        if (local) E += F[SL_LETA + a][q] * F[SL_J + a][q];
        F[SL_E + a][q] = E;
    }
}

// ---------------------------------------------------------------------------
// Hall EMF: BuildHallFaceEMF_Rusanov_diff_ (MercurySolver_MagExplicit_Hall.cpp:58-433)
// ---------------------------------------------------------------------------
__global__ void k_hall_cell(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, const MercPhys P) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q;
    const Num n = num_at(P, F, q, nspec(P));
    const double alpha = B[p.b].hall_coef / n.ne_eff;
    const double Jx = F[SL_JC][q], Jy = F[SL_JC + 1][q], Jz = F[SL_JC + 2][q];
    const double Bx = F[SL_BC][q], By = F[SL_BC + 1][q], Bz = F[SL_BC + 2][q];
    F[SL_EHC][q] = alpha * (Jy * Bz - Jz * By);
    F[SL_EHC + 1][q] = alpha * (Jz * Bx - Jx * Bz);
    F[SL_EHC + 2][q] = alpha * (Jx * By - Jy * Bx);
    F[SL_HB][q] = fabs(alpha) * sqrt(Bx * Bx + By * By + Bz * Bz);
}

__global__ void k_hall_edge(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    const double eps = 1e-14, Cwh = 0.1;
    for (int a = 0; a < 3; ++a) {
        if ((a == 0 && p.i >= bk.m[0]) || (a == 1 && p.j >= bk.m[1]) || (a == 2 && p.k >= bk.m[2])) continue;
        long long c[4];
        if (a == 0) { c[0] = q - dj - dk; c[1] = q - dj; c[2] = q - dk; c[3] = q; }
        else if (a == 1) { c[0] = q - di - dk; c[1] = q - di; c[2] = q - dk; c[3] = q; }
        else { c[0] = q - di - dj; c[1] = q - di; c[2] = q - dj; c[3] = q; }
        double Ec[3];
        for (int d = 0; d < 3; ++d) Ec[d] = 0.25 * (F[SL_EHC + d][c[0]] + F[SL_EHC + d][c[1]] + F[SL_EHC + d][c[2]] + F[SL_EHC + d][c[3]]);
        const double Ecen = Ec[0] * F[SL_DR + 3 * a][q] + Ec[1] * F[SL_DR + 3 * a + 1][q] + Ec[2] * F[SL_DR + 3 * a + 2][q];
        const double be = dmax(dmax(F[SL_HB][c[0]], F[SL_HB][c[1]]), dmax(F[SL_HB][c[2]], F[SL_HB][c[3]]));
        const double L = dmax(F[SL_DL + a][q], eps);
        double val;
        if (a == 0) {
            const double Az = 0.5 * (F[SL_AREA + 2][q] + F[SL_AREA + 2][q - dj]);
            const double ihe = L / dmax(Az, eps);
            const double mue = Cwh * be * ihe * ihe;
            const double Ae = 0.5 * (F[SL_AREA + 1][q] + F[SL_AREA + 1][q - dk]);
            const double ihz = L / dmax(Ae, eps);
            const double muz = Cwh * be * ihz * ihz;
            const double d1 = F[SL_B + 2][q] - F[SL_B + 2][q - dj];
            const double d2 = F[SL_B + 1][q] - F[SL_B + 1][q - dk];
            val = Ecen + 0.5 * mue * d1 - 0.5 * muz * d2;
        } else if (a == 1) {
            const double Ax = 0.5 * (F[SL_AREA + 0][q] + F[SL_AREA + 0][q - dk]);
            const double ihz = L / dmax(Ax, eps);
            const double muz = Cwh * be * ihz * ihz;
            const double Az = 0.5 * (F[SL_AREA + 2][q] + F[SL_AREA + 2][q - di]);
            const double ihx = L / dmax(Az, eps);
            const double mux = Cwh * be * ihx * ihx;
            const double d1 = F[SL_B + 0][q] - F[SL_B + 0][q - dk];
            const double d2 = F[SL_B + 2][q] - F[SL_B + 2][q - di];
            val = Ecen + 0.5 * muz * d1 - 0.5 * mux * d2;
        } else {
            const double Ae = 0.5 * (F[SL_AREA + 1][q] + F[SL_AREA + 1][q - di]);
            const double ihx = L / dmax(Ae, eps);
            const double mux = Cwh * be * ihx * ihx;
            const double Ax = 0.5 * (F[SL_AREA + 0][q] + F[SL_AREA + 0][q - dj]);
            const double ihe = L / dmax(Ax, eps);
            const double mue = Cwh * be * ihe * ihe;
            const double d1 = F[SL_B + 1][q] - F[SL_B + 1][q - di];
            const double d2 = F[SL_B + 0][q] - F[SL_B + 0][q - dj];
            val = Ecen + 0.5 * mux * d1 - 0.5 * mue * d2;
        }
        F[SL_EH + a][q] = F[SL_TAPER + a][q] * val;
    }
}

// E += Ehall on owned edges (MercurySolver_MagneticRHS.cpp:134-147), then the
// stationary-wall condition zeroes tangential E on owned wall edges
// (ApplyStationaryWallNonResistiveEMF_, MercurySolver_MagneticRHS.cpp:5-89)
__global__ void k_add_hall_wall(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q;
    const int idx[3] = {p.i, p.j, p.k};
    for (int a = 0; a < 3; ++a) {
        if (idx[a] >= bk.m[a]) continue;
        double e = F[SL_E + a][q] + F[SL_EH + a][q];
        for (int f = 0; f < 6; ++f) {
            const int nax = f >> 1;
            if (bk.bc[f] != MBC_WALL || nax == a) continue;
            if ((f & 1) ? (idx[nax] == bk.m[nax]) : (idx[nax] == 0)) e = 0.0;
        }
        F[SL_E + a][q] = e;
    }
}

// RHS_B = -d1(E) (CTOperators::CurlEdgeToFace, CTOperators.cpp:6-57) and the
// face part of ApplyUpdate_Euler_ (MercurySolver_Advance.cpp:122-146): B += dt*RHS_B
__global__ void k_curl_update(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                              const double *__restrict__ dtp, int do_update) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    const double dt = do_update ? dtp[bk.run] : 0.0;
    if (do_update && dt == 0.0) return;  // the run is stopped (k_dt_commit)
    if (p.j < bk.m[1] && p.k < bk.m[2]) {
        const double c = (F[SL_E + 1][q] - F[SL_E + 1][q + dk]) + (F[SL_E + 2][q + dj] - F[SL_E + 2][q]);
        const double r = -1.0 * c;
        F[SL_RHSB + 0][q] = r;
        if (do_update) F[SL_B + 0][q] += dt * r;
    }
    if (p.i < bk.m[0] && p.k < bk.m[2]) {
        const double c = (F[SL_E + 0][q + dk] - F[SL_E + 0][q]) + (F[SL_E + 2][q] - F[SL_E + 2][q + di]);
        const double r = -1.0 * c;
        F[SL_RHSB + 1][q] = r;
        if (do_update) F[SL_B + 1][q] += dt * r;
    }
    if (p.i < bk.m[0] && p.j < bk.m[1]) {
        const double c = (F[SL_E + 0][q] - F[SL_E + 0][q + dj]) + (F[SL_E + 1][q + di] - F[SL_E + 1][q]);
        const double r = -1.0 * c;
        F[SL_RHSB + 2][q] = r;
        if (do_update) F[SL_B + 2][q] += dt * r;
    }
}

// cell part of ApplyUpdate_Euler_ (MercurySolver_Advance.cpp:73-121)
__global__ void k_update_u(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, const double *__restrict__ dtp,
                           int ns) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q;
    const double dt = dtp[B[p.b].run];
    if (dt == 0.0) return;  // the run is stopped (k_dt_commit)
    for (int m = 0; m < 5; ++m)
#pragma unroll
        for (int s = 0; s < MERC_NSMAX && s < ns; ++s) F[SPEC_U(s) + m][q] += dt * F[SPEC_R(s) + m][q];
}

// SSP-RK2: save the step-start state / average 0.5*(U0 + U) on owned DOFs
__global__ void k_rk2(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, int save, int ns) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q;
    const bool cell = p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2];
    const bool face[3] = {p.j < bk.m[1] && p.k < bk.m[2], p.i < bk.m[0] && p.k < bk.m[2], p.i < bk.m[0] && p.j < bk.m[1]};
    if (save) {
        if (cell)
            for (int m = 0; m < 5; ++m)
#pragma unroll
                for (int s = 0; s < MERC_NSMAX && s < ns; ++s) F[SPEC_U0(s) + m][q] = F[SPEC_U(s) + m][q];
        for (int a = 0; a < 3; ++a)
            if (face[a]) F[SL_B0 + a][q] = F[SL_B + a][q];
    } else {
        if (cell)
            for (int m = 0; m < 5; ++m)
#pragma unroll
                for (int s = 0; s < MERC_NSMAX && s < ns; ++s) F[SPEC_U(s) + m][q] = 0.5 * (F[SPEC_U0(s) + m][q] + F[SPEC_U(s) + m][q]);
        for (int a = 0; a < 3; ++a)
            if (face[a]) F[SL_B + a][q] = 0.5 * (F[SL_B0 + a][q] + F[SL_B + a][q]);
    }
}

// ---------------------------------------------------------------------------
// synchronisation ops (halo / physical boundary / owner-alias), all index lists
// ---------------------------------------------------------------------------
// Every entry names its slot set, so one launch serves the lists of several
// fields (and of independent sync groups). Within one launch no destination is
// the source of another entry.
// COPY: entry (dst, src, set), dst <- src for every slot of the set
// The entry lists of the merged programs are padded per slot set to whole
// wavefronts (entries with dst < 0 do nothing), so the set, and with it every
// slot and slot pointer, is uniform over a wave: read once into scalar
// registers instead of a per-lane chain of table loads before each copy.
__global__ void k_copy_multi(const int *__restrict__ ent, long long n, double *const *__restrict__ F_, const MultiSets S) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = TID;
    if (t >= n) return;
    const int d = ent[3 * t], s = ent[3 * t + 1];
    const int k = __builtin_amdgcn_readfirstlane(ent[3 * t + 2]);
    if (d < 0) return;
    for (int c = 0; c < S.n[k]; ++c) F[S.s[k][c]][d] = F[S.s[k][c]][s];
}
// FILL: entry (dst, set), dst <- the set's constant per slot
__global__ void k_fill_multi(const int *__restrict__ ent, long long n, double *const *__restrict__ F_, const MultiSets S) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = TID;
    if (t >= n) return;
    const int d = ent[2 * t];
    const int k = __builtin_amdgcn_readfirstlane(ent[2 * t + 1]);
    if (d < 0) return;
    for (int c = 0; c < S.n[k]; ++c) F[S.s[k][c]][d] = S.v[k][c];
}
// SETV: entry (dst, slot) with its own value
__global__ void k_setv_multi(const int *__restrict__ ent, const double *__restrict__ val, long long n, double *const *__restrict__ F_) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = TID;
    if (t >= n) return;
    const int d = ent[2 * t];
    const int slot = __builtin_amdgcn_readfirstlane(ent[2 * t + 1]);
    if (d < 0) return;
    F[slot][d] = val[t];
}
// a caller's sync plan: entry (dst slot, dst, src slot, src, sign), dst <- sign * src
__global__ void k_copy_plan(const int *__restrict__ ent, long long n, double *const *__restrict__ F_) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int *e = ent + 5 * t;
    F[e[0]][e[1]] = (double)e[4] * F[e[2]][e[3]];
}
// a plan applied with every read before every write (its sources include
// destinations of the same plan): gather into tmp, then scatter
__global__ void k_plan_gather(const int *__restrict__ ent, long long n, double *const *__restrict__ F_, double *__restrict__ tmp) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int *e = ent + 5 * t;
    tmp[t] = (double)e[4] * F[e[2]][e[3]];
}
__global__ void k_plan_scatter(const int *__restrict__ ent, long long n, double *const *__restrict__ F_, const double *__restrict__ tmp) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int *e = ent + 5 * t;
    F[e[0]][e[1]] = tmp[t];
}
// BC_Solid_Surface_ (Boundary_Condition_Fluid.cpp:36-260): per wall cell the
// entry holds (species base slot, species (0 H+, 1 Na+, 2 He++), first layer,
// one deeper, g ghosts); He++ takes the H+ rule with its own density floor
__global__ void k_wall_multi(const int *__restrict__ ent, long long n, int g, double *const *__restrict__ F_, const MercPhys P) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = TID;
    if (t >= n) return;
    const int *e = ent + (long long)t * (4 + g);
    const int slot = __builtin_amdgcn_readfirstlane(e[0]), species = __builtin_amdgcn_readfirstlane(e[1]), is_na = species == 1;
    e += 2;
    if (e[0] < 0) return;  // padding of the entry list to whole waves
    const double gm1 = P.gamma - 1.0;
    const double rho_floor = species == 1 ? P.wall_rho_floor_Na : (species == 2 ? P.wall_rho_floor_He : P.wall_rho_floor_H);
    const double p_floor = P.wall_p_floor;
    if (is_na) {
        for (int l = 0; l < g; ++l) {
            const int d = e[2 + l];
            F[slot][d] = rho_floor;
            F[slot + 1][d] = 0.0;
            F[slot + 2][d] = 0.0;
            F[slot + 3][d] = 0.0;
            F[slot + 4][d] = p_floor / gm1;
        }
        return;
    }
    for (int layer = 0; layer <= g; ++layer) {
        const int r = layer == 0 ? e[1] : e[0];
        const int d = layer == 0 ? e[0] : e[1 + layer];
        const double rho_r = dmax(F[slot][r], rho_floor);
        const double mxr = F[slot + 1][r], myr = F[slot + 2][r], mzr = F[slot + 3][r];
        const double Er = F[slot + 4][r];
        const double rho2 = dmax(rho_r, rho_floor);
        const double ke = 0.5 * (mxr * mxr + myr * myr + mzr * mzr) / rho2;
        double eint = Er - ke;
        if (eint < 0.0) eint = 0.0;
        const double p_r = dmax(gm1 * eint, p_floor);
        F[slot][d] = rho_r;
        F[slot + 1][d] = rho_r * 0.0;
        F[slot + 2][d] = rho_r * 0.0;
        F[slot + 3][d] = rho_r * 0.0;
        const double ke_g = 0.5 * rho_r * 0.0;
        F[slot + 4][d] = p_r / gm1 + ke_g;
    }
}

// owner <- signed mean of an equivalence class (ReduceEdgeAliasCandidatesToOwners_,
// MercurySolver_Current.cpp:3-57); members summed in the listed order, each
// read from the triplet field of its own axis
__global__ void k_alias_reduce(const int *__restrict__ cls_start, const int *__restrict__ member, const int *__restrict__ msign,
                               const int *__restrict__ maxis, const int *__restrict__ owner, const int *__restrict__ osign,
                               const int *__restrict__ oaxis, long long ncls, double *const *__restrict__ F_, int slot0) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = TID;
    if (t >= ncls) return;
    const int b = cls_start[t], e = cls_start[t + 1];
    double sum = 0.0;
    for (int m = b; m < e; ++m) sum += msign[m] * F[slot0 + maxis[m]][member[m]];
    F[slot0 + oaxis[t]][owner[t]] = osign[t] * sum / (double)(e - b);
}

// ---------------------------------------------------------------------------
// DEC operators (src/8_dec/DecExteriorDerivative.cpp) on generic slots
// ---------------------------------------------------------------------------
// d0: node scalar -> edge 1-form, e = phi(end) - phi(start), owned edges
__global__ void k_dec_d0(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, int s_node, int s_edge) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q;
    if (p.i < bk.m[0]) F[s_edge + 0][q] = F[s_node][q + OI(p)] - F[s_node][q];
    if (p.j < bk.m[1]) F[s_edge + 1][q] = F[s_node][q + OJ(p)] - F[s_node][q];
    if (p.k < bk.m[2]) F[s_edge + 2][q] = F[s_node][q + 1] - F[s_node][q];
}
// d1: edge 1-form -> face 2-form (circulation), multiplier as CurlEdgeToFace
__global__ void k_dec_d1(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, int s_edge, int s_face, double mult) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    if (p.j < bk.m[1] && p.k < bk.m[2])
        F[s_face + 0][q] = mult * ((F[s_edge + 1][q] - F[s_edge + 1][q + dk]) + (F[s_edge + 2][q + dj] - F[s_edge + 2][q]));
    if (p.i < bk.m[0] && p.k < bk.m[2])
        F[s_face + 1][q] = mult * ((F[s_edge + 0][q + dk] - F[s_edge + 0][q]) + (F[s_edge + 2][q] - F[s_edge + 2][q + di]));
    if (p.i < bk.m[0] && p.j < bk.m[1])
        F[s_face + 2][q] = mult * ((F[s_edge + 0][q] - F[s_edge + 0][q + dj]) + (F[s_edge + 1][q + di] - F[s_edge + 1][q]));
}
// d2: face 2-form -> cell 3-form (outward flux sum), owned cells
__global__ void k_dec_d2(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_, int s_face, int s_cell) {
    gdouble *const *__restrict__ F = gslots(F_);
    LOCATE(R);
    const long long q = p.q;
    F[s_cell][q] = F[s_face][q + OI(p)] - F[s_face][q] + F[s_face + 1][q + OJ(p)] - F[s_face + 1][q] + F[s_face + 2][q + 1] - F[s_face + 2][q];
}

// calc_divB (MercurySolver_MagneticCell.cpp:151-241) plus max |d2 B| and max |B_face|
// calc_divB (MercurySolver_MagneticCell.cpp:151-241): d2 B per owned cell
// (optionally stored as divB = d2 B / Jac) and, for the normalised diagnostic,
// max|d2 B| and max|B_face| over owned cells / faces.  A capped grid walks the
// range with a stride; each workgroup reduces its waves in LDS and writes one
// pair of partial maxima (no atomics); k_divb_final folds the partials.  Maxima
// are exact, so the result does not depend on the order.
constexpr int DIVB_MAX_GROUPS = 2048;
__global__ void __launch_bounds__(256) k_divb(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                                              double *__restrict__ partial, int write_field) {
    gdouble *const *__restrict__ F = gslots(F_);
    double dmx = 0.0, bmx = 0.0;
    const long long total = R.start[R.nb];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        Pt p;
        locate(R, B, t, p);
        const MercBlk &bk = B[p.b];
        const long long q = p.q;
        if (p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2]) {
            const double d = (F[SL_B][q + OI(p)] - F[SL_B][q] + F[SL_B + 1][q + OJ(p)] - F[SL_B + 1][q] + F[SL_B + 2][q + 1] - F[SL_B + 2][q]);
            if (write_field) F[SL_DIVB][q] = d / F[SL_JAC][q];
            dmx = dmax(dmx, fabs(d));
        }
        if (p.j < bk.m[1] && p.k < bk.m[2]) bmx = dmax(bmx, fabs(F[SL_B][q]));
        if (p.i < bk.m[0] && p.k < bk.m[2]) bmx = dmax(bmx, fabs(F[SL_B + 1][q]));
        if (p.i < bk.m[0] && p.j < bk.m[1]) bmx = dmax(bmx, fabs(F[SL_B + 2][q]));
    }
    for (int off = 32; off > 0; off >>= 1) {
        dmx = dmax(dmx, __shfl_xor(dmx, off, 64));
        bmx = dmax(bmx, __shfl_xor(bmx, off, 64));
    }
    __shared__ double wd[4], wb[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        wd[wid] = dmx;
        wb[wid] = bmx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            dmx = dmax(dmx, wd[w]);
            bmx = dmax(bmx, wb[w]);
        }
        partial[2 * blockIdx.x] = dmx;
        partial[2 * blockIdx.x + 1] = bmx;
    }
}
__global__ void __launch_bounds__(256) k_divb_final(const double *__restrict__ partial, int ngroups, unsigned long long *__restrict__ out) {
    double dmx = 0.0, bmx = 0.0;
    for (int g = threadIdx.x; g < ngroups; g += blockDim.x) {
        dmx = dmax(dmx, partial[2 * g]);
        bmx = dmax(bmx, partial[2 * g + 1]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        dmx = dmax(dmx, __shfl_xor(dmx, off, 64));
        bmx = dmax(bmx, __shfl_xor(bmx, off, 64));
    }
    __shared__ double wd[4], wb[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        wd[wid] = dmx;
        wb[wid] = bmx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            dmx = dmax(dmx, wd[w]);
            bmx = dmax(bmx, wb[w]);
        }
        out[0] = (unsigned long long)__double_as_longlong(dmx);
        out[1] = (unsigned long long)__double_as_longlong(bmx);
    }
}

struct SlotSet {
    int n;
    int s[MERC_MAX_SYNC_SLOTS];
};
// zero a set of slots over whole blocks (replaces memset inside a captured step)
__global__ void k_zero_span(double *const *__restrict__ F_, const SlotSet S, long long n) {
    gdouble *const *__restrict__ F = gslots(F_);
    const long long t = TID;
    if (t >= n) return;
    for (int c = 0; c < S.n; ++c) F[S.s[c]][t] = 0.0;
}

// dt word initialisation inside the step
__global__ void k_dt_init(unsigned long long *dt_bits, int nruns) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nruns) dt_bits[r] = (unsigned long long)__double_as_longlong(1e100);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers (host side)
// ---------------------------------------------------------------------------
namespace merc {

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
#define LAUNCH_R(kern, Rh, ...)                                                                  \
    do {                                                                                         \
        if ((Rh).total > 0) hipLaunchKernelGGL(kern, dim3(nblk((Rh).total, 256)), dim3(256), 0, st, (Rh).dev, __VA_ARGS__); \
    } while (0)

void launch_pv_uplus(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P) { LAUNCH_R(k_pv_uplus, R, B, F, P); }
void launch_bcell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F) { LAUNCH_R(k_bcell, R, B, F); }
void launch_jedge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int sf, int se) { LAUNCH_R(k_jedge, R, B, F, sf, se); }
void launch_jcell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F) { LAUNCH_R(k_jcell, R, B, F); }
void launch_timestep(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P, unsigned long long *dt_bits,
                     int nruns) {
    hipLaunchKernelGGL(k_dt_init, dim3(nblk(nruns, 64)), dim3(64), 0, st, dt_bits, nruns);
    LAUNCH_R(k_timestep, R, B, F, P, dt_bits);
}
void launch_flux_div(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P) { LAUNCH_R(k_flux_div, R, B, F, P); }
void launch_sources(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P) { LAUNCH_R(k_sources, R, B, F, P); }
void launch_eface(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F) { LAUNCH_R(k_eface, R, B, F); }
void launch_edge_emf(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P) { LAUNCH_R(k_edge_emf, R, B, F, P); }
void launch_hall_cell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P) { LAUNCH_R(k_hall_cell, R, B, F, P); }
void launch_hall_edge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F) { LAUNCH_R(k_hall_edge, R, B, F); }
void launch_add_hall_wall(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F) { LAUNCH_R(k_add_hall_wall, R, B, F); }
void launch_curl_update(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const double *dt, int do_update) {
    LAUNCH_R(k_curl_update, R, B, F, dt, do_update);
}
void launch_update_u(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const double *dt, int ns) {
    LAUNCH_R(k_update_u, R, B, F, dt, ns);
}
void launch_rk2(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int save, int ns) { LAUNCH_R(k_rk2, R, B, F, save, ns); }
void launch_dec_d0(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int sn, int se) { LAUNCH_R(k_dec_d0, R, B, F, sn, se); }
void launch_dec_d1(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int se, int sf, double mult) { LAUNCH_R(k_dec_d1, R, B, F, se, sf, mult); }
void launch_dec_d2(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, int sf, int sc) { LAUNCH_R(k_dec_d2, R, B, F, sf, sc); }
int divb_partials() { return DIVB_MAX_GROUPS; }
void launch_divb(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, double *partial, unsigned long long *out,
                 int write_field) {
    if (R.total <= 0) return;
    const int groups = (int)std::min<long long>(DIVB_MAX_GROUPS, (R.total + 255) / 256);
    hipLaunchKernelGGL(k_divb, dim3(groups), dim3(256), 0, st, R.dev, B, F, partial, write_field);
    hipLaunchKernelGGL(k_divb_final, dim3(1), dim3(256), 0, st, partial, groups, out);
}

void launch_alias_reduce(hipStream_t st, const int *cls_start, const int *member, const int *msign, const int *maxis, const int *owner,
                         const int *osign, const int *oaxis, long long ncls, double *const *F, int slot0) {
    if (ncls > 0)
        hipLaunchKernelGGL(k_alias_reduce, dim3(nblk(ncls, 256)), dim3(256), 0, st, cls_start, member, msign, maxis, owner, osign, oaxis, ncls,
                           F, slot0);
}
void launch_copy_multi(hipStream_t st, const int *ent, long long n, double *const *F, const MultiSets &S) {
    if (n > 0) hipLaunchKernelGGL(k_copy_multi, dim3(nblk(n, 256)), dim3(256), 0, st, ent, n, F, S);
}
void launch_fill_multi(hipStream_t st, const int *ent, long long n, double *const *F, const MultiSets &S) {
    if (n > 0) hipLaunchKernelGGL(k_fill_multi, dim3(nblk(n, 256)), dim3(256), 0, st, ent, n, F, S);
}
void launch_setv_multi(hipStream_t st, const int *ent, const double *val, long long n, double *const *F) {
    if (n > 0) hipLaunchKernelGGL(k_setv_multi, dim3(nblk(n, 256)), dim3(256), 0, st, ent, val, n, F);
}
void launch_copy_plan(hipStream_t st, const int *ent, long long n, double *const *F) {
    if (n > 0) hipLaunchKernelGGL(k_copy_plan, dim3(nblk(n, 256)), dim3(256), 0, st, ent, n, F);
}
void launch_copy_plan_staged(hipStream_t st, const int *ent, long long n, double *const *F, double *tmp) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_plan_gather, dim3(nblk(n, 256)), dim3(256), 0, st, ent, n, F, tmp);
    hipLaunchKernelGGL(k_plan_scatter, dim3(nblk(n, 256)), dim3(256), 0, st, ent, n, F, (const double *)tmp);
}
void launch_wall_multi(hipStream_t st, const int *ent, long long n, int g, double *const *F, const MercPhys &P) {
    if (n > 0) hipLaunchKernelGGL(k_wall_multi, dim3(nblk(n, 256)), dim3(256), 0, st, ent, n, g, F, P);
}
void launch_zero_span(hipStream_t st, double *const *F, const int *slots, int ns, long long n) {
    if (n <= 0) return;
    SlotSet S;
    S.n = ns;
    for (int c = 0; c < ns; ++c) S.s[c] = slots[c];
    hipLaunchKernelGGL(k_zero_span, dim3(nblk(n, 256)), dim3(256), 0, st, F, S, n);
}

}  // namespace merc
