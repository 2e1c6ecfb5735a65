// Fused FP64 kernels of the Mercury explicit step.
//
// Same arithmetic as the per-function kernels of mercury_kernels.hip (and so as
// the reference), regrouped so that each pass over HBM does more work:
//   k_rhs_cell     Scheme_U_ + AddSourceToRHS_Fluid               (one cell pass)
//   k_cell_end     calc_PV + calc_Uplus + the Hall cell terms + the per-cell
//                  contravariant velocity of U+ + Compute_Timestep candidates
//   k_eface_u      AssembleOneDirectionEMF_ reading the precomputed velocities
//   k_edge_total   edge EMF: ideal UCT + ambipolar + artificial + local + Hall,
//                  then the stationary-wall zero (no periodic seams present)
//   k_update_all   both halves of ApplyUpdate_Euler_ with RHS_B = -d1(E)
//   k_bcell_jedge  calc_Bcell on owned cells + CurlAdjFaceToEdge*alpha on owned edges
// Moving a computation between passes never changes its inputs (each is
// evaluated after the same synchronisation as in the reference), so the fused
// path equals the per-function path bit for bit.
#include <hip/hip_runtime.h>

#include "mercury_device.h"
#include "mercury_fused.h"

namespace {
using namespace merc_dev;

// ---------------------------------------------------------------------------
__global__ void k_rhs_cell(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F, const MercPhys P) {
    LOCATE(R);
    const long long q = p.q, di = OI(p), dj = OJ(p);
    const double gm = P.gamma;
    const double invJ = 1.0 / F[SL_JAC][q];
    // --- Scheme_U_ (MercurySolver_FluidRHS.cpp:3-202)
    double rr[2][5];
    for (int s = 0; s < 2; ++s) {
        const int su = s ? SL_UN : SL_UH, sp = s ? SL_PVN : SL_PVH;
        double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) {
            const long long d = AX(p, a);
            double Fm[5], Fp[5];
            rusanov_face((const double *const *)F, su, sp, q - d, q, q, a, gm, Fm);
            rusanov_face((const double *const *)F, su, sp, q, q + d, q + d, a, gm, Fp);
#pragma unroll
            for (int m = 0; m < 5; ++m) r[m] -= (Fp[m] - Fm[m]) * invJ;
        }
        for (int m = 0; m < 5; ++m) rr[s][m] = r[m];
    }
    // --- AddSourceToRHS_Fluid (MercurySolver_FluidRHS.cpp:206-527)
    const double upx = F[SL_UP][q], upy = F[SL_UP + 1][q], upz = F[SL_UP + 2][q];
    const double Bx = F[SL_BC][q], By = F[SL_BC + 1][q], Bz = F[SL_BC + 2][q];
    const double Jx = F[SL_JC][q], Jy = F[SL_JC + 1][q], Jz = F[SL_JC + 2][q];
    const double sjbx = Jy * Bz - Jz * By, sjby = Jz * Bx - Jx * Bz, sjbz = Jx * By - Jy * Bx;
    const Num num = num_limiter(P, F[SL_UH][q], F[SL_UN][q]);
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
#define PE(o) (F[SL_PVH + 3][q + (o)] + F[SL_PVN + 3][q + (o)])
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
    for (int s = 0; s < 2; ++s) {
        const int su = s ? SL_UN : SL_UH, sp = s ? SL_PVN : SL_PVH, sr = s ? SL_RN : SL_RH;
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
        const double Ms = s ? P.M_Na : P.M_H;
        const double Mo = P.M_Na;
        const double b1 = vst * (P.Tn0 - Ts0) * (Ms / (Ms + Mo));
        const double b2 = (Mo / (Ms + Mo)) * vst;
        const double ns = s ? num.nNa : num.nH;
        const double chi = s ? num.wNa : num.wH;
        double r0 = rr[s][0], r1 = rr[s][1], r2 = rr[s][2], r3 = rr[s][3], r4 = rr[s][4];
        r1 += P.c_ind * ns * subx + P.c_hall * chi * sjbx;
        r2 += P.c_ind * ns * suby + P.c_hall * chi * sjby;
        r3 += P.c_ind * ns * subz + P.c_hall * chi * sjbz;
        r4 += P.c_ind * ns * subu + P.c_hall * chi * sjbu;
        if (ambi) {
            r1 -= chi * dpex;
            r2 -= chi * dpey;
            r3 -= chi * dpez;
            r4 -= chi * dpeu;
        }
        if (s) {
            r0 += P.a1_Na * sss;
            r4 += P.a6 * sss * P.Tn0;
        }
        r1 -= P.a4 * rho * u * vst;
        r2 -= P.a4 * rho * v * vst;
        r3 -= P.a4 * rho * w * vst;
        r4 -= P.a4 * rho * u2 * vst;
        r4 += P.a5 * ns * b1;
        r4 += P.a4 * rho * u2 * b2;
        F[sr][q] = r0; F[sr + 1][q] = r1; F[sr + 2][q] = r2; F[sr + 3][q] = r3; F[sr + 4][q] = r4;
    }
}

// ---------------------------------------------------------------------------
// cell quantities that only depend on the synchronised state of the point:
// PV_s, U+ (Tools.cpp:159-301), Hall E_hc and beta (Hall.cpp:87-124), the
// contravariant velocity of U+ per axis (cell_u_contra, Ideal.cpp:513-526) and,
// on owned cells when requested, the CFL candidate of Compute_Timestep.
__global__ void k_cell_end(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F, const MercPhys P,
                           unsigned long long *__restrict__ dt_bits, int compute_dt) {
    double best = 1e100;
    Pt p;
    if (locate(R, B, TID, p)) {
        const MercBlk &bk = B[p.b];
        const long long q = p.q;
        double pv[2][5];
        for (int s = 0; s < 2; ++s) {
            const int su = s ? SL_UN : SL_UH;
            const double coeff = s ? P.state_coeff_Na : P.state_coeff_H;
            const double rho = dmax(F[su][q], 1e-12);
            const double u = F[su + 1][q] / rho, v = F[su + 2][q] / rho, w = F[su + 3][q] / rho;
            const double E = F[su + 4][q];
            const double ke = 0.5 * rho * (u * u + v * v + w * w);
            double eint = E - ke;
            if (eint < 0.0) eint = 0.0;
            double pr = (P.gamma - 1.0) * eint;
            pr = dmax(pr, 1e-12);
            pv[s][0] = u; pv[s][1] = v; pv[s][2] = w; pv[s][3] = pr; pv[s][4] = pr / (rho * coeff);
            const int sp = s ? SL_PVN : SL_PVH;
            for (int c = 0; c < 5; ++c) F[sp + c][q] = pv[s][c];
        }
        const double rH = dmax(F[SL_UH][q], 0.0), rN = dmax(F[SL_UN][q], 0.0);
        const Num n = num_limiter(P, rH, rN);
        double up[3];
        for (int c = 0; c < 3; ++c) {
            const double uH = (rH > 1e-20) ? pv[0][c] : 0.0;
            const double uN = (rN > 1e-20) ? pv[1][c] : 0.0;
            up[c] = n.chiH * uH + n.chiNa * uN;
            F[SL_UP + c][q] = up[c];
        }
        // Hall cell terms (the limiter of the raw densities, Hall.cpp:92-95)
        const Num nh = num_limiter(P, F[SL_UH][q], F[SL_UN][q]);
        const double alpha = P.hall_coef / nh.ne_eff;
        const double Jx = F[SL_JC][q], Jy = F[SL_JC + 1][q], Jz = F[SL_JC + 2][q];
        const double Bx = F[SL_BC][q], By = F[SL_BC + 1][q], Bz = F[SL_BC + 2][q];
        F[SL_EHC][q] = alpha * (Jy * Bz - Jz * By);
        F[SL_EHC + 1][q] = alpha * (Jz * Bx - Jx * Bz);
        F[SL_EHC + 2][q] = alpha * (Jx * By - Jy * Bx);
        F[SL_HB][q] = fabs(alpha) * sqrt(Bx * Bx + By * By + Bz * Bz);
        // contravariant velocities (need the face of the next index on each axis)
        const int idx[3] = {p.i, p.j, p.k};
        const double inv_jac = 1.0 / (fabs(F[SL_JAC][q]) + 1.0e-30);
        for (int a = 0; a < 3; ++a) {
            if (idx[a] + 1 > bk.m[a] + bk.g) continue;
            const long long cp = q + AX(p, a);
            const int s0 = SL_JD + 3 * a;
            const double kx = 0.5 * (F[s0][q] + F[s0][cp]);
            const double ky = 0.5 * (F[s0 + 1][q] + F[s0 + 1][cp]);
            const double kz = 0.5 * (F[s0 + 2][q] + F[s0 + 2][cp]);
            F[SL_UCON + a][q] = (kx * up[0] + ky * up[1] + kz * up[2]) * inv_jac;
        }
        // time-step candidate on owned cells
        if (compute_dt && p.i >= 0 && p.j >= 0 && p.k >= 0 && p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2]) {
            const double V = fabs(F[SL_JAC][q]);
            if (!(V <= 0.0)) {
                const double B2 = Bx * Bx + By * By + Bz * Bz;
                for (int s = 0; s < 2; ++s) {
                    const int su = s ? SL_UN : SL_UH;
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
    }
    if (!compute_dt) return;
    for (int off = 32; off > 0; off >>= 1) best = dmin(best, __shfl_xor(best, off, 64));
    __shared__ double wmin[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = dmin(m, wmin[w]);
        atomicMin(dt_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// ---------------------------------------------------------------------------
DEVI double emf_flux_u(double *const *__restrict__ F, const Pt &p, int ba, int fa, long long L, long long Rr, double bff) {
    const double btL = beta_tot(F, p, ba, L), btR = beta_tot(F, p, ba, Rr);
    const double biL = beta_ind(F, p, ba, L), biR = beta_ind(F, p, ba, Rr);
    const double ufL = F[SL_UCON + fa][L], ufR = F[SL_UCON + fa][Rr];
    const double ubL = F[SL_UCON + ba][L], ubR = F[SL_UCON + ba][Rr];
    const double GL = ufL * btL - ubL * bff, GR = ufR * btR - ubR * bff;
    const double radius = dmax(fabs(ufL), fabs(ufR));
    return 0.5 * (GL + GR) - 0.5 * radius * (biR - biL);
}

__global__ void k_eface_u(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F) {
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
            const double Gex = emf_flux_u(F, p, 1, 0, Lq, Rq, bff);
            const double Gzx = emf_flux_u(F, p, 2, 0, Lq, Rq, bff);
            e1 = Gzx;
            e2 = -Gex;
        } else if (fa == 1) {
            const double Gxe = emf_flux_u(F, p, 0, 1, Lq, Rq, bff);
            const double Gze = emf_flux_u(F, p, 2, 1, Lq, Rq, bff);
            e0 = -Gze;
            e2 = Gxe;
        } else {
            const double Gxz = emf_flux_u(F, p, 0, 2, Lq, Rq, bff);
            const double Gez = emf_flux_u(F, p, 1, 2, Lq, Rq, bff);
            e0 = Gez;
            e1 = -Gxz;
        }
        F[SL_EF + 3 * fa][Rq] = e0;
        F[SL_EF + 3 * fa + 1][Rq] = e1;
        F[SL_EF + 3 * fa + 2][Rq] = e2;
    }
}

DEVI double maxu4_p(double *const *__restrict__ F, int ax, long long c0, long long c1, long long c2, long long c3) {
    double a = 0.0;
    a = dmax(a, fabs(F[SL_UCON + ax][c0]));
    a = dmax(a, fabs(F[SL_UCON + ax][c1]));
    a = dmax(a, fabs(F[SL_UCON + ax][c2]));
    a = dmax(a, fabs(F[SL_UCON + ax][c3]));
    return a;
}

// edge EMF, all contributions, owned edges; with_hall adds the Hall EMF and
// applies the wall condition (used when no periodic seam needs the alias mean
// of the non-Hall part first)
__global__ void k_edge_total(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F, const MercPhys P,
                             int with_hall) {
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    const int idx[3] = {p.i, p.j, p.k};
    const double Cuct = 0.5;
    const bool ambi = P.ambipolar_enabled != 0.0;
    const bool arti = P.arti_eta_max > 0.0;
    const bool local = P.local_enabled != 0.0 && P.local_eta_max > 0.0 && P.local_r_decay > 0.0;
    for (int a = 0; a < 3; ++a) {
        if (idx[a] >= bk.m[a]) continue;
        double E;
        if (a == 0) {
            const double Ec = 0.25 * (F[SL_EF + 3][q] + F[SL_EF + 3][q - dk] + F[SL_EF + 6][q] + F[SL_EF + 6][q - dj]);
            const double ae = maxu4_p(F, 1, q, q - dj, q - dk, q - dj - dk);
            const double az = maxu4_p(F, 2, q, q - dj, q - dk, q - dj - dk);
            const double dBz = F[SL_B + 2][q] - F[SL_B + 2][q - dj];
            const double dBe = F[SL_B + 1][q] - F[SL_B + 1][q - dk];
            E = Ec + Cuct * (0.5 * ae * dBz - 0.5 * az * dBe);
        } else if (a == 1) {
            const double Ec = 0.25 * (F[SL_EF + 1][q] + F[SL_EF + 1][q - dk] + F[SL_EF + 7][q] + F[SL_EF + 7][q - di]);
            const double az = maxu4_p(F, 2, q, q - di, q - dk, q - di - dk);
            const double ax = maxu4_p(F, 0, q, q - di, q - dk, q - di - dk);
            const double dBx = F[SL_B + 0][q] - F[SL_B + 0][q - dk];
            const double dBz = F[SL_B + 2][q] - F[SL_B + 2][q - di];
            E = Ec + Cuct * (0.5 * az * dBx - 0.5 * ax * dBz);
        } else {
            const double Ec = 0.25 * (F[SL_EF + 2][q] + F[SL_EF + 2][q - dj] + F[SL_EF + 5][q] + F[SL_EF + 5][q - di]);
            const double ax = maxu4_p(F, 0, q, q - di, q - dj, q - di - dj);
            const double ae = maxu4_p(F, 1, q, q - di, q - dj, q - di - dj);
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
                            sum += F[SL_PVH + 3][c] + F[SL_PVN + 3][c];
                        }
                pe[e] = 0.125 * sum;
            }
            double rH = 0.0, rN = 0.0;
            const long long t1 = (a == 0) ? dj : di, t2 = (a == 2) ? dj : 1;
            for (int d1 = -1; d1 <= 0; ++d1)
                for (int d2 = -1; d2 <= 0; ++d2) {
                    const long long c = q + d1 * t1 + d2 * t2;
                    rH += F[SL_UH][c];
                    rN += F[SL_UN][c];
                }
            const Num n = num_limiter(P, 0.25 * rH, 0.25 * rN);
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
        if (!with_hall) {
            F[SL_E + a][q] = E;
            continue;
        }
        // Hall EMF of this edge (Hall.cpp:126-385)
        const double eps = 1e-14, Cwh = 0.1;
        long long c[4];
        if (a == 0) { c[0] = q - dj - dk; c[1] = q - dj; c[2] = q - dk; c[3] = q; }
        else if (a == 1) { c[0] = q - di - dk; c[1] = q - di; c[2] = q - dk; c[3] = q; }
        else { c[0] = q - di - dj; c[1] = q - di; c[2] = q - dj; c[3] = q; }
        double Ecc[3];
        for (int d = 0; d < 3; ++d) Ecc[d] = 0.25 * (F[SL_EHC + d][c[0]] + F[SL_EHC + d][c[1]] + F[SL_EHC + d][c[2]] + F[SL_EHC + d][c[3]]);
        const double Ecen = Ecc[0] * F[SL_DR + 3 * a][q] + Ecc[1] * F[SL_DR + 3 * a + 1][q] + Ecc[2] * F[SL_DR + 3 * a + 2][q];
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
        const double eh = F[SL_TAPER + a][q] * val;
        F[SL_EH + a][q] = eh;
        double e = E + eh;
        for (int f = 0; f < 6; ++f) {
            const int nax = f >> 1;
            if (bk.bc[f] != MBC_WALL || nax == a) continue;
            if ((f & 1) ? (idx[nax] == bk.m[nax]) : (idx[nax] == 0)) e = 0.0;
        }
        F[SL_E + a][q] = e;
    }
}

// ---------------------------------------------------------------------------
__global__ void k_update_all(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F, const double *__restrict__ dtp) {
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    const double dt = *dtp;
    if (p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2])
        for (int m = 0; m < 5; ++m) {
            F[SL_UH + m][q] += dt * F[SL_RH + m][q];
            F[SL_UN + m][q] += dt * F[SL_RN + m][q];
        }
    if (p.j < bk.m[1] && p.k < bk.m[2]) {
        const double c = (F[SL_E + 1][q] - F[SL_E + 1][q + dk]) + (F[SL_E + 2][q + dj] - F[SL_E + 2][q]);
        const double r = -1.0 * c;
        F[SL_RHSB + 0][q] = r;
        F[SL_B + 0][q] += dt * r;
    }
    if (p.i < bk.m[0] && p.k < bk.m[2]) {
        const double c = (F[SL_E + 0][q + dk] - F[SL_E + 0][q]) + (F[SL_E + 2][q] - F[SL_E + 2][q + di]);
        const double r = -1.0 * c;
        F[SL_RHSB + 1][q] = r;
        F[SL_B + 1][q] += dt * r;
    }
    if (p.i < bk.m[0] && p.j < bk.m[1]) {
        const double c = (F[SL_E + 0][q] - F[SL_E + 0][q + dj]) + (F[SL_E + 1][q + di] - F[SL_E + 1][q]);
        const double r = -1.0 * c;
        F[SL_RHSB + 2][q] = r;
        F[SL_B + 2][q] += dt * r;
    }
}

// ---------------------------------------------------------------------------
__global__ void k_bcell_jedge(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F) {
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    if (p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2]) {
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
#define BB(a, o) (F[SL_BETA + (a)][q + (o)] * F[SL_B + (a)][q + (o)])
    if (p.i < bk.m[0]) {
        const double v = (BB(1, -dk) - BB(1, 0)) + (BB(2, 0) - BB(2, -dj));
        F[SL_J + 0][q] = (1.0 * v) * F[SL_ALPHA + 0][q];
    }
    if (p.j < bk.m[1]) {
        const double v = (BB(0, 0) - BB(0, -dk)) + (BB(2, -di) - BB(2, 0));
        F[SL_J + 1][q] = (1.0 * v) * F[SL_ALPHA + 1][q];
    }
    if (p.k < bk.m[2]) {
        const double v = (BB(0, -dj) - BB(0, 0)) + (BB(1, 0) - BB(1, -di));
        F[SL_J + 2][q] = (1.0 * v) * F[SL_ALPHA + 2][q];
    }
#undef BB
}

__global__ void k_dt_commit(unsigned long long *dt, unsigned long long *dt_next) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dt = *dt_next;
}
__global__ void k_dt_reset(unsigned long long *dt_next) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dt_next = (unsigned long long)__double_as_longlong(1e100);
}

}  // namespace

namespace merc {

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
#define LAUNCH_R(kern, Rh, ...)                                                                  \
    do {                                                                                         \
        if ((Rh).total > 0) hipLaunchKernelGGL(kern, dim3(nblk((Rh).total, 256)), dim3(256), 0, st, (Rh).dev, __VA_ARGS__); \
    } while (0)

void launch_rhs_cell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P) { LAUNCH_R(k_rhs_cell, R, B, F, P); }
void launch_cell_end(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P, unsigned long long *dt_next,
                     int compute_dt) {
    if (compute_dt) hipLaunchKernelGGL(k_dt_reset, dim3(1), dim3(64), 0, st, dt_next);
    LAUNCH_R(k_cell_end, R, B, F, P, dt_next, compute_dt);
}
void launch_eface_u(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F) { LAUNCH_R(k_eface_u, R, B, F); }
void launch_edge_total(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const MercPhys &P, int with_hall) {
    LAUNCH_R(k_edge_total, R, B, F, P, with_hall);
}
void launch_update_all(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const double *dt) { LAUNCH_R(k_update_all, R, B, F, dt); }
void launch_bcell_jedge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F) { LAUNCH_R(k_bcell_jedge, R, B, F); }
void launch_dt_commit(hipStream_t st, unsigned long long *dt, unsigned long long *dt_next) {
    hipLaunchKernelGGL(k_dt_commit, dim3(1), dim3(64), 0, st, dt, dt_next);
}

}  // namespace merc
