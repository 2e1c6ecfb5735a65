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

// precision policy: T = double (bit-exact FP64 form) or float (mixed form).
// In the mixed form the static arrays marked FP32 are read from the float
// table G, divisions and square roots use the hardware reciprocal / square
// root (about 1 ulp in FP32), and multi-term sums are compensated.
template <bool MIX> struct Prec;
template <> struct Prec<false> {
    using T = double;
    static DEVI T div(T a, T b) { return a / b; }
    static DEVI T sq(T a) { return sqrt(a); }
};
template <> struct Prec<true> {
    using T = float;
    static DEVI T div(T a, T b) { return a * __builtin_amdgcn_rcpf(b); }
    static DEVI T sq(T a) { return __builtin_amdgcn_sqrtf(a); }
};
// static array or derived intermediate element: FP64 table in the FP64 form,
// FP32 table G (same slot index) in the mixed form.  The derived intermediates
// kept in FP32 by the mixed form are PV u, v, w, T (pressure stays FP64), U+,
// the Hall cell terms, the contravariant U+, the face EMFs and the fluid RHS.
template <bool MIX>
DEVI typename Prec<MIX>::T st(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, int s, long long q) {
    if constexpr (MIX) return G[s][q];
    else return F[s][q];
}
template <bool MIX, typename V>
DEVI void wr(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, int s, long long q, V v) {
    if constexpr (MIX) const_cast<gfloat *>(G[s])[q] = (float)v;
    else F[s][q] = (double)v;
}
template <typename T> DEVI T tmax(T a, T b) { return (a < b) ? b : a; }
template <typename T> DEVI T tmin(T a, T b) { return (b < a) ? b : a; }

// compensated accumulation (Kahan) used by the mixed form for the flux
// divergence; the FP64 form keeps the reference's plain sum
template <bool MIX> struct Acc {
    using T = typename Prec<MIX>::T;
    T s = T(0), c = T(0);
    DEVI void sub(T x) {
        if constexpr (MIX) {
            const T y = -x - c;
            const T t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s -= x;
        }
    }
    DEVI double value() const { return MIX ? (double)s - (double)c : (double)s; }
};

// Rusanov flux of one face from the two sides' primitive values
// v = (rho, u, v, w, p) and the face metric (k1, k2, k3)
template <bool MIX>
DEVI void rusanov_vals(typename Prec<MIX>::T k1, typename Prec<MIX>::T k2, typename Prec<MIX>::T k3, const typename Prec<MIX>::T vL[5],
                       const typename Prec<MIX>::T vR[5], typename Prec<MIX>::T gm, typename Prec<MIX>::T out[5]) {
    using T = typename Prec<MIX>::T;
    using PR = Prec<MIX>;
    T Uc[2][5], Fl[2][5], rad[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const T *v5 = side ? vR : vL;
        const T rho = v5[0], u = v5[1], v = v5[2], w = v5[3], pr = v5[4];
        Uc[side][0] = rho;
        Uc[side][1] = rho * u;
        Uc[side][2] = rho * v;
        Uc[side][3] = rho * w;
        Uc[side][4] = PR::div(pr, gm - T(1)) + T(0.5) * rho * (u * u + v * v + w * w);
        const T uvw = k1 * u + k2 * v + k3 * w;
        const T kk2 = (k1 * k1 + k2 * k2 + k3 * k3);
        const T c = PR::sq(tmax(T(0), PR::div(gm * pr, tmax(rho, T(1e-30)))));
        rad[side] = fabs(uvw) + c * PR::sq(kk2);
        const T rhoe = PR::div(pr, gm - T(1)) + T(0.5) * rho * (u * u + v * v + w * w);
        Fl[side][0] = rho * uvw;
        Fl[side][1] = rho * uvw * u + k1 * pr;
        Fl[side][2] = rho * uvw * v + k2 * pr;
        Fl[side][3] = rho * uvw * w + k3 * pr;
        Fl[side][4] = uvw * (rhoe + pr);
    }
    const T r = tmax(rad[0], rad[1]);
#pragma unroll
    for (int m = 0; m < 5; ++m) out[m] = T(0.5) * (Fl[0][m] + Fl[1][m]) - T(0.5) * r * (Uc[1][m] - Uc[0][m]);
}

// the primitive values of species (su, sp) at point qq as the flux reads them
template <bool MIX>
DEVI void face_side(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, int su, int sp, long long qq,
                    typename Prec<MIX>::T v5[5]) {
    using T = typename Prec<MIX>::T;
    v5[0] = (T)F[su][qq];
    v5[1] = st<MIX>(F, G, sp, qq);
    v5[2] = st<MIX>(F, G, sp + 1, qq);
    v5[3] = st<MIX>(F, G, sp + 2, qq);
    v5[4] = (T)F[sp + 3][qq];
}

template <bool MIX>
DEVI void rusanov_face_t(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, int su, int sp, long long qL, long long qR,
                         long long qf, int a, typename Prec<MIX>::T gm, typename Prec<MIX>::T out[5]) {
    using T = typename Prec<MIX>::T;
    const T k1 = (T)F[SL_JD + 3 * a][qf], k2 = (T)F[SL_JD + 3 * a + 1][qf], k3 = (T)F[SL_JD + 3 * a + 2][qf];
    T vL[5], vR[5];
    face_side<MIX>(F, G, su, sp, qL, vL);
    face_side<MIX>(F, G, su, sp, qR, vR);
    rusanov_vals<MIX>(k1, k2, k3, vL, vR, gm, out);
}

// AddSourceToRHS_Fluid (MercurySolver_FluidRHS.cpp:206-527) at owned cell q:
// the quantities every species' sources share (U+, B, J x B, the number and
// charge densities, the ambipolar electron pressure gradient, the
// photoionisation rate and the lit/shadow collision rate) ...
struct SrcShared {
    double upx, upy, upz, Bx, By, Bz, sjbx, sjby, sjbz, dpex, dpey, dpez, sss, vst;
    Num num;
    bool ambi;
};
// pe(o) is the electron pressure stand-in at the neighbour o = 0..5 (+i, -i,
// +j, -j, +k, -k) of the ambipolar gradient
template <bool MIX, int NS, typename PE>
DEVI SrcShared src_shared(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, const MercPhys &P, long long q, long long di,
                          long long dj, PE pe) {
    SrcShared c;
    c.upx = st<MIX>(F, G, SL_UP, q), c.upy = st<MIX>(F, G, SL_UP + 1, q), c.upz = st<MIX>(F, G, SL_UP + 2, q);
    c.Bx = F[SL_BC][q], c.By = F[SL_BC + 1][q], c.Bz = F[SL_BC + 2][q];
    const double Jx = F[SL_JC][q], Jy = F[SL_JC + 1][q], Jz = F[SL_JC + 2][q];
    c.sjbx = Jy * c.Bz - Jz * c.By, c.sjby = Jz * c.Bx - Jx * c.Bz, c.sjbz = Jx * c.By - Jy * c.Bx;
    c.num = num_at(P, F, q, NS);
    c.dpex = 0.0, c.dpey = 0.0, c.dpez = 0.0;
    c.ambi = P.ambipolar_enabled != 0.0;
    if (c.ambi) {
        const double V = fabs(F[SL_JAC][q]);
        double d[9];
        if (V <= 0.0) {
            for (int t = 0; t < 9; ++t) d[t] = 0.0;
        } else {
            for (int k = 0; k < 3; ++k) {
                d[k] = 0.5 * (F[SL_JD + k][q - di] + F[SL_JD + k][q]) / V;
                d[3 + k] = 0.5 * (F[SL_JD + 3 + k][q - dj] + F[SL_JD + 3 + k][q]) / V;
                d[6 + k] = 0.5 * (F[SL_JD + 6 + k][q - 1] + F[SL_JD + 6 + k][q]) / V;
            }
        }
        const double pip = pe(0), pim = pe(1), pjp = pe(2), pjm = pe(3), pkp = pe(4), pkm = pe(5);
        c.dpex = 0.5 * (d[0] * (pip - pim) + d[3] * (pjp - pjm) + d[6] * (pkp - pkm));
        c.dpey = 0.5 * (d[1] * (pip - pim) + d[4] * (pjp - pjm) + d[7] * (pkp - pkm));
        c.dpez = 0.5 * (d[2] * (pip - pim) + d[5] * (pjp - pjm) + d[8] * (pkp - pkm));
    }
    c.sss = (double)st<MIX>(F, G, SL_PHOTO, q);
    const double x = F[SL_CC][q], y = F[SL_CC + 1][q], z = F[SL_CC + 2][q];
    const double rho_cyl = sqrt(y * y + z * z);
    const bool in_shadow = (x < 0.0 && rho_cyl < 1.0);
    c.vst = in_shadow ? P.nu_shadow : P.nu_lit;
    return c;
}

// ... and one species' sources added to its flux divergence r, written as its RHS
template <bool MIX>
DEVI void src_species(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, const MercPhys &P, long long q, int s,
                      const SrcShared &c, const double r[5]) {
    const int su = SPEC_U(s), sp = SPEC_PV(s), sr = SPEC_R(s);
    const double rho = dmax(F[su][q], 0.0);
    const double u = st<MIX>(F, G, sp, q), v = st<MIX>(F, G, sp + 1, q), w = st<MIX>(F, G, sp + 2, q);
    const double u2 = u * u + v * v + w * w;
    const double subx = (v - c.upy) * c.Bz - (w - c.upz) * c.By;
    const double suby = (w - c.upz) * c.Bx - (u - c.upx) * c.Bz;
    const double subz = (u - c.upx) * c.By - (v - c.upy) * c.Bx;
    const double subu = subx * u + suby * v + subz * w;
    const double sjbu = c.sjbx * u + c.sjby * v + c.sjbz * w;
    const double dpeu = c.dpex * u + c.dpey * v + c.dpez * w;
    const double Ts0 = (double)st<MIX>(F, G, sp + 4, q) * P.T_ref;
    const double Ms = species_mass(P, s);
    const double Mo = P.M_Na;
    const double b1 = c.vst * (P.Tn0 - Ts0) * (Ms / (Ms + Mo));
    const double b2 = (Mo / (Ms + Mo)) * c.vst;
    // number density in the thermal exchange, charge density in the induction force
    const double nsp = c.num.n[s], qs = c.num.q[s];
    const double chi = c.num.w[s];
    double r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4];
    r1 += P.c_ind * qs * subx + P.c_hall * chi * c.sjbx;
    r2 += P.c_ind * qs * suby + P.c_hall * chi * c.sjby;
    r3 += P.c_ind * qs * subz + P.c_hall * chi * c.sjbz;
    r4 += P.c_ind * qs * subu + P.c_hall * chi * sjbu;
    if (c.ambi) {
        r1 -= chi * c.dpex;
        r2 -= chi * c.dpey;
        r3 -= chi * c.dpez;
        r4 -= chi * dpeu;
    }
    if (s == 1) {
        r0 += P.a1_Na * c.sss;
        r4 += P.a6 * c.sss * P.Tn0;
    }
    r1 -= P.a4 * rho * u * c.vst;
    r2 -= P.a4 * rho * v * c.vst;
    r3 -= P.a4 * rho * w * c.vst;
    r4 -= P.a4 * rho * u2 * c.vst;
    r4 += P.a5 * nsp * b1;
    r4 += P.a4 * rho * u2 * b2;
    wr<MIX>(F, G, sr, q, r0); wr<MIX>(F, G, sr + 1, q, r1); wr<MIX>(F, G, sr + 2, q, r2);
    wr<MIX>(F, G, sr + 3, q, r3); wr<MIX>(F, G, sr + 4, q, r4);
}

// the sources of every species of the cell
template <bool MIX, int NS, typename PE>
DEVI void rhs_sources(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, const MercPhys &P, long long q, long long di,
                      long long dj, double (&rr)[NS][5], PE pe) {
    const SrcShared c = src_shared<MIX, NS>(F, G, P, q, di, dj, pe);
    for (int s = 0; s < NS; ++s) src_species<MIX>(F, G, P, q, s, c, rr[s]);
}

// ---------------------------------------------------------------------------
// Scheme_U_ + AddSourceToRHS_Fluid.  One thread marches over a column of up to
// L owned cells along i: the i+1/2 face flux of a cell is the i-1/2 flux of the
// next one (the same face evaluated once, with the same operands), and the
// i-neighbour planes are read while they are still in cache.  The launch range
// has ceil(m0 / L) points along i per block.
template <bool MIX, int NS>
__global__ void __launch_bounds__(256) k_rhs_cell(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                           const float *const *__restrict__ G_, const MercPhys P, int L) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    using T = typename Prec<MIX>::T;
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long di = OI(p), dj = OJ(p);
    const int i0 = p.i * L, i1 = (bk.m[0] < i0 + L) ? bk.m[0] : i0 + L;
    const long long q0 = p.q + (long long)(i0 - p.i) * di;
    const T gm = (T)P.gamma;
    // --- Scheme_U_ (MercurySolver_FluidRHS.cpp:3-202)
    T Fx[NS][5];  // i-1/2 flux of the current cell, per species
    for (int s = 0; s < NS; ++s) rusanov_face_t<MIX>(F, G, SPEC_U(s), SPEC_PV(s), q0 - di, q0, q0, 0, gm, Fx[s]);
    for (int i = i0; i < i1; ++i) {
        const long long q = q0 + (long long)(i - i0) * di;
        const T invJ = (T)(1.0 / F[SL_JAC][q]);
        double rr[NS][5];
        for (int s = 0; s < NS; ++s) {
            const int su = SPEC_U(s), sp = SPEC_PV(s);
            Acc<MIX> r[5];
            {
                T Fp[5];
                rusanov_face_t<MIX>(F, G, su, sp, q, q + di, q + di, 0, gm, Fp);
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    r[m].sub((Fp[m] - Fx[s][m]) * invJ);
                    Fx[s][m] = Fp[m];
                }
            }
            auto axis = [&](int a) {
                const long long d = AX(p, a);
                T Fm[5], Fp[5];
                rusanov_face_t<MIX>(F, G, su, sp, q - d, q, q, a, gm, Fm);
                rusanov_face_t<MIX>(F, G, su, sp, q, q + d, q + d, a, gm, Fp);
#pragma unroll
                for (int m = 0; m < 5; ++m) r[m].sub((Fp[m] - Fm[m]) * invJ);
            };
            if constexpr (MIX) {
                // the mixed body keeps one axis live at a time (unrolled it spills)
#pragma unroll 1
                for (int a = 1; a < 3; ++a) axis(a);
            } else {
                for (int a = 1; a < 3; ++a) axis(a);
            }
            for (int m = 0; m < 5; ++m) rr[s][m] = r[m].value();
        }
        rhs_sources<MIX, NS>(F, G, P, q, di, dj, rr, [&](int o) {
            const long long d = o == 0 ? di : o == 1 ? -di : o == 2 ? dj : o == 3 ? -dj : o == 4 ? 1LL : -1LL;
            return pe_sum(F, q + d, NS);
        });
    }
}

// ---------------------------------------------------------------------------
// Scheme_U_ + AddSourceToRHS_Fluid, LDS-tiled.  A workgroup owns a 16 x 16
// tile of (j, k) columns of one block, one column per thread, and marches over
// up to C planes along i.  Each plane's flux operands (rho, u, v, w, p per
// species, as the flux reads them) are staged in LDS with the tile's one-cell
// j/k halo ring; every j and k face of the tile is evaluated once, by one
// thread, and reaches both of its cells through LDS (the column kernel
// evaluates each such face twice, from global memory).  The i face is carried
// along the column as in k_rhs_cell, and the next plane is loaded while the
// current one is in use.  Same operands, arithmetic and summation order as
// k_rhs_cell, so the two kernels agree bit for bit.  The launch range has
// ceil(m0 / C) x ceil(m1 / 16) x ceil(m2 / 16) points per block, one
// workgroup each.
constexpr int RT_J = 16, RT_K = 16;
constexpr int RT_W = RT_K + 2;                   // staged row: the tile's k range and its two halo cells
constexpr int RT_S = (RT_J + 2) * RT_W;          // staged plane
constexpr int RT_FJ = (RT_J + 1) * RT_K;         // j faces of a tile plane
constexpr int RT_F = RT_FJ + RT_J * (RT_K + 1);  // and the k faces
static_assert(RT_J * RT_K == 256 && RT_J == 16 && RT_K == 16, "one thread per column, 64 halo cells");
static_assert(RT_F > 512 && RT_F <= 768, "three face slots per thread");

template <bool MIX, int NS>
__global__ void __launch_bounds__(256) k_rhs_tile(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                                                  const float *const *__restrict__ G_, const MercPhys P, int C) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    using T = typename Prec<MIX>::T;
    __shared__ T s_v[NS * 5][RT_S];  // flux operands of the staged plane
    __shared__ double s_pe[RT_S];     // electron pressure stand-in of the staged plane
    // face fluxes of SB species at a time: all of them where the LDS allows
    // (FP64 three-species: one, 63 KB with the staged plane)
    constexpr int SB = (MIX || NS == 2) ? NS : 1;
    __shared__ T s_f[SB][5][RT_F];
    Pt p;
    if (!locate(R, B, xcd_block(), p)) return;  // the tile: uniform over the workgroup
    const MercBlk &bk = B[p.b];
    const int tid = threadIdx.x, tj = tid / RT_K, tk = tid % RT_K;
    const int j0 = p.j * RT_J, k0 = p.k * RT_K, i0 = p.i * C;
    const int i1 = (bk.m[0] < i0 + C) ? bk.m[0] : i0 + C;
    const int m1 = bk.m[1], m2 = bk.m[2], g = bk.g;
    const long long di = bk.sI, dj = bk.sJ;
    // staged cell (r, c) of plane i is the point (i, j0 - 1 + r, k0 - 1 + c);
    // base0 addresses (i0, j0 - 1, k0 - 1)
    const long long base0 = bk.off + (long long)(i0 + g) * di + (long long)(j0 - 1 + g) * dj + (k0 - 1 + g);
    const int c_own = (tj + 1) * RT_W + (tk + 1);
    const long long own_off = (long long)(tj + 1) * dj + (tk + 1);
    const bool own = (j0 + tj < m1) && (k0 + tk < m2);
    // the staged points are the tile's cells up to index m (the upper halo of
    // a partial tile) and the halo ring; one in four threads stages a ring cell
    const bool ld = (j0 + tj <= m1) && (k0 + tk <= m2);
    int c_h = -1;
    long long h_off = 0;
    if ((tid & 3) == 0) {
        const int h = tid >> 2, e = h & 15, side = h >> 4;
        const int r = side == 0 ? 0 : (side == 1 ? RT_J + 1 : e + 1);
        const int c = side < 2 ? e + 1 : (side == 2 ? 0 : RT_K + 1);
        if (j0 - 1 + r <= m1 && k0 - 1 + c <= m2) {
            c_h = r * RT_W + c;
            h_off = (long long)r * dj + c;
        }
    }
    // the faces this thread evaluates: f = tid, tid + 256 and tid + 512 of the
    // tile plane's RT_F faces (j faces first, then k faces); face f of axis a
    // lies between the staged cells fl and fr and is the lower face of fr
    int fl[3], fr[3], fa[3];
    long long fo[3];
    bool fv[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int f = tid + 256 * u;
        int r0, c0;
        if (f < RT_FJ) {
            r0 = f / RT_K;
            c0 = f % RT_K + 1;
            fl[u] = r0 * RT_W + c0;
            fr[u] = fl[u] + RT_W;
            fo[u] = (long long)(r0 + 1) * dj + c0;
            fa[u] = 1;
            fv[u] = (j0 + r0 <= m1) && (k0 + c0 - 1 < m2);
        } else {
            const int e = f - RT_FJ;
            r0 = e / (RT_K + 1) + 1;
            c0 = e % (RT_K + 1);
            fl[u] = r0 * RT_W + c0;
            fr[u] = fl[u] + 1;
            fo[u] = (long long)r0 * dj + c0 + 1;
            fa[u] = 2;
            fv[u] = (f < RT_F) && (j0 + r0 - 1 < m1) && (k0 + c0 <= m2);
        }
    }
    const int fj_lo = tj * RT_K + tk, fk_lo = RT_FJ + tj * (RT_K + 1) + tk;
    const T gm = (T)P.gamma;
    auto load = [&](long long qq, T v[NS * 5], double &pe) {
#pragma unroll
        for (int s = 0; s < NS; ++s) face_side<MIX>(F, G, SPEC_U(s), SPEC_PV(s), qq, v + 5 * s);
        pe = pe_sum(F, qq, NS);
    };
    auto stage = [&](int c, const T v[NS * 5], double pe) {
#pragma unroll
        for (int t = 0; t < NS * 5; ++t) s_v[t][c] = v[t];
        s_pe[c] = pe;
    };
    // --- Scheme_U_ (MercurySolver_FluidRHS.cpp:3-202)
    T vn[NS * 5], vh[NS * 5];  // the next plane's own and ring cells
    double pen = 0.0, peh = 0.0, pe_m = 0.0;
    T Fx[NS][5];  // i-1/2 flux of the current cell, per species
    if (own) {
        T vm[NS * 5];
        load(base0 - di + own_off, vm, pe_m);
        load(base0 + own_off, vn, pen);
        const long long q = base0 + own_off;
        const T k1 = (T)F[SL_JD][q], k2 = (T)F[SL_JD + 1][q], k3 = (T)F[SL_JD + 2][q];
#pragma unroll
        for (int s = 0; s < NS; ++s) rusanov_vals<MIX>(k1, k2, k3, vm + 5 * s, vn + 5 * s, gm, Fx[s]);
    } else if (ld) {
        load(base0 + own_off, vn, pen);
    }
    if (c_h >= 0) load(base0 + h_off, vh, peh);
    if (ld) stage(c_own, vn, pen);
    if (c_h >= 0) stage(c_h, vh, peh);
    __syncthreads();
    for (int i = i0; i < i1; ++i) {
        const long long base = base0 + (long long)(i - i0) * di;
        const long long q = base + own_off;
        const double pe_c = pen;
        // plane i + 1 (a plane of the box: i + 1 <= m0)
        if (ld) load(q + di, vn, pen);
        if (c_h >= 0) load(base + di + h_off, vh, peh);
        T km[3][3];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            if (!fv[u]) continue;
            // the two candidate arrays are uniform: a per-lane select of the
            // pointer, not a per-lane load from the slot table
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const gdouble *pj = F[SL_JD + 3 + c], *pk = F[SL_JD + 6 + c];
                km[u][c] = (T)(fa[u] == 2 ? pk : pj)[base + fo[u]];
            }
        }
        T ki[3] = {T(0), T(0), T(0)};
        T invJ = T(0);
        double pjp = 0.0, pjm = 0.0, pkp = 0.0, pkm = 0.0;
        if (own) {
            ki[0] = (T)F[SL_JD][q + di];
            ki[1] = (T)F[SL_JD + 1][q + di];
            ki[2] = (T)F[SL_JD + 2][q + di];
            invJ = (T)(1.0 / F[SL_JAC][q]);
            pjp = s_pe[c_own + RT_W];
            pjm = s_pe[c_own - RT_W];
            pkp = s_pe[c_own + 1];
            pkm = s_pe[c_own - 1];
        }
        double rr[NS][5];
#pragma unroll
        for (int s0 = 0; s0 < NS; s0 += SB) {
            if (s0 > 0) __syncthreads();  // the previous batch's fluxes are consumed
#pragma unroll
            for (int s = s0; s < s0 + SB; ++s) {
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    if (!fv[u]) continue;
                    T vL[5], vR[5], out[5];
#pragma unroll
                    for (int m = 0; m < 5; ++m) {
                        vL[m] = s_v[5 * s + m][fl[u]];
                        vR[m] = s_v[5 * s + m][fr[u]];
                    }
                    rusanov_vals<MIX>(km[u][0], km[u][1], km[u][2], vL, vR, gm, out);
#pragma unroll
                    for (int m = 0; m < 5; ++m) s_f[s - s0][m][tid + 256 * u] = out[m];
                }
            }
            __syncthreads();
            if (own) {
#pragma unroll
                for (int s = s0; s < s0 + SB; ++s) {
                    Acc<MIX> r[5];
                    T vc[5], Fp[5];
#pragma unroll
                    for (int m = 0; m < 5; ++m) vc[m] = s_v[5 * s + m][c_own];
                    rusanov_vals<MIX>(ki[0], ki[1], ki[2], vc, vn + 5 * s, gm, Fp);
                    const T(&f)[5][RT_F] = s_f[s - s0];
#pragma unroll
                    for (int m = 0; m < 5; ++m) {
                        r[m].sub((Fp[m] - Fx[s][m]) * invJ);
                        Fx[s][m] = Fp[m];
                    }
#pragma unroll
                    for (int m = 0; m < 5; ++m) r[m].sub((f[m][fj_lo + RT_K] - f[m][fj_lo]) * invJ);
#pragma unroll
                    for (int m = 0; m < 5; ++m) r[m].sub((f[m][fk_lo + 1] - f[m][fk_lo]) * invJ);
#pragma unroll
                    for (int m = 0; m < 5; ++m) rr[s][m] = r[m].value();
                }
            }
        }
        // after the last barrier each thread reads only its own staged cell,
        // which only it rewrites below
        if (own) {
            rhs_sources<MIX, NS>(F, G, P, q, di, dj, rr, [&](int o) {
                return o == 0 ? pen : o == 1 ? pe_m : o == 2 ? pjp : o == 3 ? pjm : o == 4 ? pkp : pkm;
            });
        }
        pe_m = pe_c;
        if (ld) stage(c_own, vn, pen);
        if (c_h >= 0) stage(c_h, vh, peh);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// cell quantities that only depend on the synchronised state of the point:
// PV_s, U+ (Tools.cpp:159-301), Hall E_hc and beta (Hall.cpp:87-124), the
// contravariant velocity of U+ per axis (cell_u_contra, Ideal.cpp:513-526) and,
// on owned cells when requested, the CFL candidate of Compute_Timestep.
// Six waves per SIMD: the FP64 body fits in 80 VGPRs without spilling (102
// unconstrained, four waves), and the latency-bound pass runs 10 % faster
// (398 -> 359 us on shell6); seven waves would spill.
// ENS: an ensemble context (per-block Hall coefficient and per-run dt); a
// one-run context takes both from the kernel arguments and keeps every
// register of the six-wave budget for the body.
template <bool MIX, int NS, bool ENS>
__global__ void __attribute__((amdgpu_waves_per_eu(NS == 2 ? 6 : 5))) k_cell_end(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                           const float *const *__restrict__ G_, const MercPhys P, unsigned long long *__restrict__ dt_bits,
                           int compute_dt) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    using T = typename Prec<MIX>::T;
    using PR = Prec<MIX>;
    double best = 1e100;
    Pt p;
    const bool located = locate(R, B, TID, p);
    if (located) {
        const MercBlk &bk = B[p.b];
        const long long q = p.q;
        double pv[NS][5];
        for (int s = 0; s < NS; ++s) {
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
            for (int c = 0; c < 5; ++c) {
                if (c == 3) F[sp + 3][q] = pv[s][3];
                else wr<MIX>(F, G, sp + c, q, pv[s][c]);
            }
        }
        double r[NS];
        for (int s = 0; s < NS; ++s) r[s] = dmax(F[SPEC_U(s)][q], 0.0);
        // one limiter serves U+ (floored densities) and the Hall terms (raw
        // densities, Hall.cpp:92-95): it floors its inputs at 0 itself
        const Num n = num_limiter(P, r[0], r[1], NS > 2 ? r[NS - 1] : 0.0, NS);
        double up[3];
        for (int c = 0; c < 3; ++c) {
            const double uH = (r[0] > 1e-20) ? pv[0][c] : 0.0;
            const double uN = (r[1] > 1e-20) ? pv[1][c] : 0.0;
            up[c] = n.chi[0] * uH + n.chi[1] * uN;
            if constexpr (NS > 2) up[c] = up[c] + n.chi[2] * ((r[2] > 1e-20) ? pv[2][c] : 0.0);
            wr<MIX>(F, G, SL_UP + c, q, up[c]);
        }
        const Num &nh = n;
        const double Bx = F[SL_BC][q], By = F[SL_BC + 1][q], Bz = F[SL_BC + 2][q];
        {
            const T alpha = (T)((ENS ? bk.hall_coef : P.hall_coef) / nh.ne_eff);
            const T Jx = (T)F[SL_JC][q], Jy = (T)F[SL_JC + 1][q], Jz = (T)F[SL_JC + 2][q];
            const T bx = (T)Bx, by = (T)By, bz = (T)Bz;
            const T e0 = alpha * (Jy * bz - Jz * by), e1 = alpha * (Jz * bx - Jx * bz), e2 = alpha * (Jx * by - Jy * bx);
            const T hb = fabs(alpha) * PR::sq(bx * bx + by * by + bz * bz);
            wr<MIX>(F, G, SL_EHC, q, e0);
            wr<MIX>(F, G, SL_EHC + 1, q, e1);
            wr<MIX>(F, G, SL_EHC + 2, q, e2);
            wr<MIX>(F, G, SL_HB, q, hb);
        }
        // contravariant velocities (need the face of the next index on each axis)
        const int idx[3] = {p.i, p.j, p.k};
        const T inv_jac = MIX ? (T)(1.0 / (fabs(F[SL_JAC][q]) + 1.0e-30)) : (T)(1.0 / (fabs(F[SL_JAC][q]) + 1.0e-30));
        for (int a = 0; a < 3; ++a) {
            if (idx[a] + 1 > bk.m[a] + bk.g) continue;
            const long long cp = q + AX(p, a);
            const int s0 = SL_JD + 3 * a;
            const T kx = T(0.5) * ((T)F[s0][q] + (T)F[s0][cp]);
            const T ky = T(0.5) * ((T)F[s0 + 1][q] + (T)F[s0 + 1][cp]);
            const T kz = T(0.5) * ((T)F[s0 + 2][q] + (T)F[s0 + 2][cp]);
            wr<MIX>(F, G, SL_UCON + a, q, (kx * (T)up[0] + ky * (T)up[1] + kz * (T)up[2]) * inv_jac);
        }
        // time-step candidate on owned cells
        if (compute_dt && p.i >= 0 && p.j >= 0 && p.k >= 0 && p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2]) {
            const double V = fabs(F[SL_JAC][q]);
            if (!(V <= 0.0)) {
                // per species state first, then one pass over the six faces:
                // the face normal and |S| (the stored area, bitwise the same
                // sqrt) are shared by the species.  FP64 in the FP64 forms; the
                // mixed form evaluates the candidate in FP32 (the global minimum
                // stays an exact reduction of FP64 values)
                const T bx = (T)Bx, by = (T)By, bz = (T)Bz, gam = (T)P.gamma, mi = (T)P.inver_MA2;
                const T B2 = bx * bx + by * by + bz * bz;
                T rho[NS], ux[NS], uy[NS], uz[NS], cs2[NS], vA2[NS], denom[NS];
                for (int s = 0; s < NS; ++s) {
                    const int su = SPEC_U(s);
                    denom[s] = T(0);
                    rho[s] = tmax((T)F[su][q], T(1e-2));
                    ux[s] = PR::div((T)F[su + 1][q], rho[s]);
                    uy[s] = PR::div((T)F[su + 2][q], rho[s]);
                    uz[s] = PR::div((T)F[su + 3][q], rho[s]);
                    const T E = (T)F[su + 4][q];
                    const T ke = T(0.5) * rho[s] * (ux[s] * ux[s] + uy[s] * uy[s] + uz[s] * uz[s]);
                    const T pr = (gam - T(1)) * (E - ke);
                    cs2[s] = PR::div(gam * pr, rho[s]);
                    vA2[s] = PR::div(mi * B2, rho[s]);
                }
                for (int a = 0; a < 3; ++a)
                    for (int sd = 0; sd < 2; ++sd) {
                        const long long fq = q + (sd ? AX(p, a) : 0LL);
                        const T ax = (T)F[SL_JD + 3 * a][fq], ay = (T)F[SL_JD + 3 * a + 1][fq], az = (T)F[SL_JD + 3 * a + 2][fq];
                        const T A = st<MIX>(F, G, SL_AREA + a, fq);
                        if (A <= T(0)) {
                            for (int s = 0; s < NS; ++s) denom[s] += T(0);
                            continue;
                        }
                        T nx = PR::div(ax, A), ny = PR::div(ay, A), nz = PR::div(az, A);
                        if (!sd) { nx = -nx; ny = -ny; nz = -nz; }
                        const T Bn = bx * nx + by * ny + bz * nz;
                        for (int s = 0; s < NS; ++s) {
                            const T un = ux[s] * nx + uy[s] * ny + uz[s] * nz;
                            const T vAn2 = PR::div(mi * (Bn * Bn), rho[s]);
                            const T t = cs2[s] + vA2[s];
                            T disc = t * t - T(4) * cs2[s] * vAn2;
                            if (disc < T(0)) disc = T(0);
                            const T cf = PR::sq(T(0.5) * (t + PR::sq(disc)));
                            denom[s] += (fabs(un) + cf) * A;
                        }
                    }
                for (int s = 0; s < NS; ++s)
                    if (denom[s] > T(0)) best = dmin(best, P.cfl * V / (double)denom[s]);
            }
        }
    }
    if (!compute_dt) return;
    // the run is read here, not before the body: its register is not held across it
    dt_min_runs(best, located ? (ENS ? B[p.b].run : 0) : -1, dt_bits);
}

// ---------------------------------------------------------------------------
template <bool MIX>
DEVI typename Prec<MIX>::T emf_flux_u(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, const Pt &p, int ba, int fa,
                                      long long L, long long Rr, typename Prec<MIX>::T bff) {
    using T = typename Prec<MIX>::T;
    const long long dL = AX(p, ba);
    const T btL = T(0.5) * (((T)F[SL_B + ba][L] + (T)F[SL_BADD + ba][L]) + ((T)F[SL_B + ba][L + dL] + (T)F[SL_BADD + ba][L + dL]));
    const T btR = T(0.5) * (((T)F[SL_B + ba][Rr] + (T)F[SL_BADD + ba][Rr]) + ((T)F[SL_B + ba][Rr + dL] + (T)F[SL_BADD + ba][Rr + dL]));
    const T biL = T(0.5) * ((T)F[SL_B + ba][L] + (T)F[SL_B + ba][L + dL]);
    const T biR = T(0.5) * ((T)F[SL_B + ba][Rr] + (T)F[SL_B + ba][Rr + dL]);
    const T ufL = st<MIX>(F, G, SL_UCON + fa, L), ufR = st<MIX>(F, G, SL_UCON + fa, Rr);
    const T ubL = st<MIX>(F, G, SL_UCON + ba, L), ubR = st<MIX>(F, G, SL_UCON + ba, Rr);
    const T GL = ufL * btL - ubL * bff, GR = ufR * btR - ubR * bff;
    const T radius = tmax(fabs(ufL), fabs(ufR));
    return T(0.5) * (GL + GR) - T(0.5) * radius * (biR - biL);
}

// face EMF candidates (AssembleOneDirectionEMF_ x3) and, with the ambipolar
// term, the node electron pressure its edges need: the mean of p_H + p_Na over
// the 8 cells around the node, summed in the order of AddAmbipolarEdgeEMF_
// (MercurySolver_MagExplicit_Ambipo.cpp:3-143), once per node instead of once
// per edge end
template <bool MIX>
__global__ void k_eface_u(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                          const float *const *__restrict__ G_, int ambi, int ns) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    using T = typename Prec<MIX>::T;
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const int idx[3] = {p.i, p.j, p.k};
    if (ambi && p.i >= 0 && p.j >= 0 && p.k >= 0 && p.i <= bk.m[0] && p.j <= bk.m[1] && p.k <= bk.m[2]) {
        const long long di = OI(p), dj = OJ(p);
        T sum = T(0);
        for (int ddi = -1; ddi <= 0; ++ddi)
            for (int ddj = -1; ddj <= 0; ++ddj)
                for (int ddk = -1; ddk <= 0; ++ddk) {
                    const long long c = p.q + ddi * di + ddj * dj + ddk;
                    T pe = (T)F[SL_PVH + 3][c] + (T)F[SL_PVN + 3][c];
                    if (ns > 2) pe = pe + (T)F[SL_PVHE + 3][c];
                    sum += pe;
                }
        wr<MIX>(F, G, SL_PEN, p.q, T(0.125) * sum);
    }
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
        const T bff = (T)F[SL_B + fa][Rq] + (T)F[SL_BADD + fa][Rq];
        // only the two tangential components are written: the normal one is
        // zero by construction (the per-function pass stores that zero) and
        // its slot keeps the zero of the arena's initialisation
        if (fa == 0) {
            const T Gex = emf_flux_u<MIX>(F, G, p, 1, 0, Lq, Rq, bff);
            const T Gzx = emf_flux_u<MIX>(F, G, p, 2, 0, Lq, Rq, bff);
            wr<MIX>(F, G, SL_EF + 1, Rq, Gzx);
            wr<MIX>(F, G, SL_EF + 2, Rq, -Gex);
        } else if (fa == 1) {
            const T Gxe = emf_flux_u<MIX>(F, G, p, 0, 1, Lq, Rq, bff);
            const T Gze = emf_flux_u<MIX>(F, G, p, 2, 1, Lq, Rq, bff);
            wr<MIX>(F, G, SL_EF + 3, Rq, -Gze);
            wr<MIX>(F, G, SL_EF + 5, Rq, Gxe);
        } else {
            const T Gxz = emf_flux_u<MIX>(F, G, p, 0, 2, Lq, Rq, bff);
            const T Gez = emf_flux_u<MIX>(F, G, p, 1, 2, Lq, Rq, bff);
            wr<MIX>(F, G, SL_EF + 6, Rq, Gez);
            wr<MIX>(F, G, SL_EF + 7, Rq, -Gxz);
        }
    }
}

template <bool MIX>
DEVI typename Prec<MIX>::T maxu4_p(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, int ax, long long c0, long long c1,
                                   long long c2, long long c3) {
    using T = typename Prec<MIX>::T;
    T a = T(0);
    a = tmax(a, (T)fabs(st<MIX>(F, G, SL_UCON + ax, c0)));
    a = tmax(a, (T)fabs(st<MIX>(F, G, SL_UCON + ax, c1)));
    a = tmax(a, (T)fabs(st<MIX>(F, G, SL_UCON + ax, c2)));
    a = tmax(a, (T)fabs(st<MIX>(F, G, SL_UCON + ax, c3)));
    return a;
}

// edge EMF, all contributions, owned edges.  with_hall 0: the non-Hall EMF
// only; 1: the Hall EMF (also stored on its own) added and the wall condition
// applied (when no shared edge needs the alias mean of the non-Hall part
// first); 2: the non-Hall EMF and the Hall EMF stored apart, for the alias mean
// of the first before the two are added
template <bool MIX>
__global__ void k_edge_total(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                             const float *const *__restrict__ G_, const MercPhys P, int with_hall) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    using T = typename Prec<MIX>::T;
    using PR = Prec<MIX>;
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    const int idx[3] = {p.i, p.j, p.k};
    const T Cuct = T(0.5);
    const bool ambi = P.ambipolar_enabled != 0.0;
    const bool arti = P.arti_eta_max > 0.0;
    const bool local = P.local_enabled != 0.0 && P.local_eta_max > 0.0 && P.local_r_decay > 0.0;
#define FB(s, o) ((T)F[(s)][q + (o)])
#define FE(s, o) (st<MIX>(F, G, (s), q + (o)))
    for (int a = 0; a < 3; ++a) {
        if (idx[a] >= bk.m[a]) continue;
        T E;
        if (a == 0) {
            const T Ec = T(0.25) * (FE(SL_EF + 3, 0) + FE(SL_EF + 3, -dk) + FE(SL_EF + 6, 0) + FE(SL_EF + 6, -dj));
            const T ae = maxu4_p<MIX>(F, G, 1, q, q - dj, q - dk, q - dj - dk);
            const T az = maxu4_p<MIX>(F, G, 2, q, q - dj, q - dk, q - dj - dk);
            const T dBz = FB(SL_B + 2, 0) - FB(SL_B + 2, -dj);
            const T dBe = FB(SL_B + 1, 0) - FB(SL_B + 1, -dk);
            E = Ec + Cuct * (T(0.5) * ae * dBz - T(0.5) * az * dBe);
        } else if (a == 1) {
            const T Ec = T(0.25) * (FE(SL_EF + 1, 0) + FE(SL_EF + 1, -dk) + FE(SL_EF + 7, 0) + FE(SL_EF + 7, -di));
            const T az = maxu4_p<MIX>(F, G, 2, q, q - di, q - dk, q - di - dk);
            const T ax = maxu4_p<MIX>(F, G, 0, q, q - di, q - dk, q - di - dk);
            const T dBx = FB(SL_B + 0, 0) - FB(SL_B + 0, -dk);
            const T dBz = FB(SL_B + 2, 0) - FB(SL_B + 2, -di);
            E = Ec + Cuct * (T(0.5) * az * dBx - T(0.5) * ax * dBz);
        } else {
            const T Ec = T(0.25) * (FE(SL_EF + 2, 0) + FE(SL_EF + 2, -dj) + FE(SL_EF + 5, 0) + FE(SL_EF + 5, -di));
            const T ax = maxu4_p<MIX>(F, G, 0, q, q - di, q - dj, q - di - dj);
            const T ae = maxu4_p<MIX>(F, G, 1, q, q - di, q - dj, q - di - dj);
            const T dBe = FB(SL_B + 1, 0) - FB(SL_B + 1, -di);
            const T dBx = FB(SL_B + 0, 0) - FB(SL_B + 0, -dj);
            E = Ec + Cuct * (T(0.5) * ax * dBe - T(0.5) * ae * dBx);
        }
        if (ambi) {
            const long long ea = AX(p, a);
            const T pe[2] = {st<MIX>(F, G, SL_PEN, q), st<MIX>(F, G, SL_PEN, q + ea)};  // node electron pressure (k_eface_u)
            const int ns = nspec(P);
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
            E -= PR::div((T)P.pressure_coef * (pe[1] - pe[0]), (T)n.ne_eff);
        }
        if (arti) {
            const T J1 = FB(SL_J + a, 0);
            const T dx = st<MIX>(F, G, SL_DR + 3 * a, q), dy = st<MIX>(F, G, SL_DR + 3 * a + 1, q), dz = st<MIX>(F, G, SL_DR + 3 * a + 2, q);
            const T drl = PR::sq(dx * dx + dy * dy + dz * dz);
            const T Jm = PR::div(fabs(J1), tmax(drl, T(1.0e-30)));
            T eta;
            if (P.J_on <= P.J_start)
                eta = (Jm > (T)P.J_start) ? (T)P.arti_eta_max : T(0);
            else {
                T qq = PR::div(Jm - (T)P.J_start, (T)(P.J_on - P.J_start));
                qq = tmax(T(0), tmin(T(1), qq));
                eta = (T)P.arti_eta_max * (qq * qq * (T(3) - T(2) * qq));
            }
            E += eta * J1;
        }
        if (local) E += st<MIX>(F, G, SL_LETA + a, q) * FB(SL_J + a, 0);
        if (!with_hall) {
            F[SL_E + a][q] = (double)E;
            continue;
        }
        // Hall EMF of this edge (Hall.cpp:126-385)
        const T eps = T(1e-14), Cwh = T(0.1);
        long long c[4];
        if (a == 0) { c[0] = q - dj - dk; c[1] = q - dj; c[2] = q - dk; c[3] = q; }
        else if (a == 1) { c[0] = q - di - dk; c[1] = q - di; c[2] = q - dk; c[3] = q; }
        else { c[0] = q - di - dj; c[1] = q - di; c[2] = q - dj; c[3] = q; }
        T Ecc[3];
        for (int d = 0; d < 3; ++d)
            Ecc[d] = T(0.25) * (st<MIX>(F, G, SL_EHC + d, c[0]) + st<MIX>(F, G, SL_EHC + d, c[1]) + st<MIX>(F, G, SL_EHC + d, c[2]) +
                                st<MIX>(F, G, SL_EHC + d, c[3]));
        const T Ecen = Ecc[0] * st<MIX>(F, G, SL_DR + 3 * a, q) + Ecc[1] * st<MIX>(F, G, SL_DR + 3 * a + 1, q) +
                       Ecc[2] * st<MIX>(F, G, SL_DR + 3 * a + 2, q);
        const T be = tmax(tmax(st<MIX>(F, G, SL_HB, c[0]), st<MIX>(F, G, SL_HB, c[1])), tmax(st<MIX>(F, G, SL_HB, c[2]), st<MIX>(F, G, SL_HB, c[3])));
        const T L = tmax(st<MIX>(F, G, SL_DL + a, q), eps);
#define AR(s, o) st<MIX>(F, G, (s), q + (o))
        T val;
        if (a == 0) {
            const T Az = T(0.5) * (AR(SL_AREA + 2, 0) + AR(SL_AREA + 2, -dj));
            const T ihe = PR::div(L, tmax(Az, eps));
            const T mue = Cwh * be * ihe * ihe;
            const T Ae = T(0.5) * (AR(SL_AREA + 1, 0) + AR(SL_AREA + 1, -dk));
            const T ihz = PR::div(L, tmax(Ae, eps));
            const T muz = Cwh * be * ihz * ihz;
            const T d1 = FB(SL_B + 2, 0) - FB(SL_B + 2, -dj);
            const T d2 = FB(SL_B + 1, 0) - FB(SL_B + 1, -dk);
            val = Ecen + T(0.5) * mue * d1 - T(0.5) * muz * d2;
        } else if (a == 1) {
            const T Ax = T(0.5) * (AR(SL_AREA + 0, 0) + AR(SL_AREA + 0, -dk));
            const T ihz = PR::div(L, tmax(Ax, eps));
            const T muz = Cwh * be * ihz * ihz;
            const T Az = T(0.5) * (AR(SL_AREA + 2, 0) + AR(SL_AREA + 2, -di));
            const T ihx = PR::div(L, tmax(Az, eps));
            const T mux = Cwh * be * ihx * ihx;
            const T d1 = FB(SL_B + 0, 0) - FB(SL_B + 0, -dk);
            const T d2 = FB(SL_B + 2, 0) - FB(SL_B + 2, -di);
            val = Ecen + T(0.5) * muz * d1 - T(0.5) * mux * d2;
        } else {
            const T Ae = T(0.5) * (AR(SL_AREA + 1, 0) + AR(SL_AREA + 1, -di));
            const T ihx = PR::div(L, tmax(Ae, eps));
            const T mux = Cwh * be * ihx * ihx;
            const T Ax = T(0.5) * (AR(SL_AREA + 0, 0) + AR(SL_AREA + 0, -dj));
            const T ihe = PR::div(L, tmax(Ax, eps));
            const T mue = Cwh * be * ihe * ihe;
            const T d1 = FB(SL_B + 1, 0) - FB(SL_B + 1, -di);
            const T d2 = FB(SL_B + 0, 0) - FB(SL_B + 0, -dj);
            val = Ecen + T(0.5) * mux * d1 - T(0.5) * mue * d2;
        }
#undef AR
        const T eh = st<MIX>(F, G, SL_TAPER + a, q) * val;
        F[SL_EH + a][q] = (double)eh;
        if (with_hall == 2) {
            F[SL_E + a][q] = (double)E;
            continue;
        }
        T e = E + eh;
        for (int f = 0; f < 6; ++f) {
            const int nax = f >> 1;
            if (bk.bc[f] != MBC_WALL || nax == a) continue;
            if ((f & 1) ? (idx[nax] == bk.m[nax]) : (idx[nax] == 0)) e = T(0);
        }
        F[SL_E + a][q] = (double)e;
    }
#undef FB
#undef FE
}

// ---------------------------------------------------------------------------
// (the RHS comes from the FP32 table in the mixed form; the accumulation is FP64)
template <bool MIX>
__global__ void k_update_all(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                             const float *const *__restrict__ G_, const double *__restrict__ dtp, int ns) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    LOCATE(R);
    const MercBlk &bk = B[p.b];
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    const double dt = dtp[bk.run];
    if (dt == 0.0) return;  // the run is stopped (k_dt_commit)
    if (p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2])
        for (int m = 0; m < 5; ++m) {
            // the fluid RHS is read here only, RHS_B written here and only downloaded:
            // non-temporal, so they do not displace reused data from the caches
#pragma unroll
            for (int s = 0; s < MERC_NSMAX && s < ns; ++s) {
                if constexpr (MIX) F[SPEC_U(s) + m][q] += dt * (double)__builtin_nontemporal_load(&G[SPEC_R(s) + m][q]);
                else F[SPEC_U(s) + m][q] += dt * __builtin_nontemporal_load(&F[SPEC_R(s) + m][q]);
            }
        }
    if (p.j < bk.m[1] && p.k < bk.m[2]) {
        const double c = (F[SL_E + 1][q] - F[SL_E + 1][q + dk]) + (F[SL_E + 2][q + dj] - F[SL_E + 2][q]);
        const double r = -1.0 * c;
        __builtin_nontemporal_store(r, &F[SL_RHSB + 0][q]);
        F[SL_B + 0][q] += dt * r;
    }
    if (p.i < bk.m[0] && p.k < bk.m[2]) {
        const double c = (F[SL_E + 0][q + dk] - F[SL_E + 0][q]) + (F[SL_E + 2][q] - F[SL_E + 2][q + di]);
        const double r = -1.0 * c;
        __builtin_nontemporal_store(r, &F[SL_RHSB + 1][q]);
        F[SL_B + 1][q] += dt * r;
    }
    if (p.i < bk.m[0] && p.j < bk.m[1]) {
        const double c = (F[SL_E + 0][q] - F[SL_E + 0][q + dj]) + (F[SL_E + 1][q + di] - F[SL_E + 1][q]);
        const double r = -1.0 * c;
        __builtin_nontemporal_store(r, &F[SL_RHSB + 2][q]);
        F[SL_B + 2][q] += dt * r;
    }
}

// ---------------------------------------------------------------------------
template <bool MIX>
DEVI void bcell_jedge_point(gdouble *const *__restrict__ F, const gfloat *const *__restrict__ G, const MercBlk &bk, const Pt &p,
                            bool monitor, double &dmx, double &bmx) {
    using T = typename Prec<MIX>::T;
    const long long q = p.q, di = OI(p), dj = OJ(p), dk = 1;
    if (p.i < bk.m[0] && p.j < bk.m[1] && p.k < bk.m[2]) {
        T pt[6], pi[6];
        pt[0] = -(T)F[SL_B + 0][q] - (T)F[SL_BADD + 0][q];
        pt[1] = (T)F[SL_B + 0][q + di] + (T)F[SL_BADD + 0][q + di];
        pt[2] = -(T)F[SL_B + 1][q] - (T)F[SL_BADD + 1][q];
        pt[3] = (T)F[SL_B + 1][q + dj] + (T)F[SL_BADD + 1][q + dj];
        pt[4] = -(T)F[SL_B + 2][q] - (T)F[SL_BADD + 2][q];
        pt[5] = (T)F[SL_B + 2][q + 1] + (T)F[SL_BADD + 2][q + 1];
        pi[0] = -(T)F[SL_B + 0][q];
        pi[1] = (T)F[SL_B + 0][q + di];
        pi[2] = -(T)F[SL_B + 1][q];
        pi[3] = (T)F[SL_B + 1][q + dj];
        pi[4] = -(T)F[SL_B + 2][q];
        pi[5] = (T)F[SL_B + 2][q + 1];
        T bt[3] = {T(0), T(0), T(0)}, bi[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int n = 0; n < 6; ++n)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T w = st<MIX>(F, G, SL_WB + 6 * c + n, q);
                bt[c] += w * pt[n];
                bi[c] += w * pi[n];
            }
        for (int c = 0; c < 3; ++c) {
            F[SL_BC + c][q] = (double)bt[c];
            F[SL_BI + c][q] = (double)bi[c];
        }
        if (monitor) {
            // calc_divB's d2 B (the expression and order of k_divb)
            const double d = (F[SL_B][q + di] - F[SL_B][q] + F[SL_B + 1][q + dj] - F[SL_B + 1][q] + F[SL_B + 2][q + 1] - F[SL_B + 2][q]);
            dmx = dmax(dmx, fabs(d));
        }
    }
    if (monitor) {
        if (p.j < bk.m[1] && p.k < bk.m[2]) bmx = dmax(bmx, fabs(F[SL_B][q]));
        if (p.i < bk.m[0] && p.k < bk.m[2]) bmx = dmax(bmx, fabs(F[SL_B + 1][q]));
        if (p.i < bk.m[0] && p.j < bk.m[1]) bmx = dmax(bmx, fabs(F[SL_B + 2][q]));
    }
#define BB(a, o) (st<MIX>(F, G, SL_BETA + (a), q + (o)) * (T)F[SL_B + (a)][q + (o)])
    if (p.i < bk.m[0]) {
        const T v = (BB(1, -dk) - BB(1, 0)) + (BB(2, 0) - BB(2, -dj));
        F[SL_J + 0][q] = (double)((T(1) * v) * st<MIX>(F, G, SL_ALPHA + 0, q));
    }
    if (p.j < bk.m[1]) {
        const T v = (BB(0, 0) - BB(0, -dk)) + (BB(2, -di) - BB(2, 0));
        F[SL_J + 1][q] = (double)((T(1) * v) * st<MIX>(F, G, SL_ALPHA + 1, q));
    }
    if (p.k < bk.m[2]) {
        const T v = (BB(0, -dj) - BB(0, 0)) + (BB(1, 0) - BB(1, -di));
        F[SL_J + 2][q] = (double)((T(1) * v) * st<MIX>(F, G, SL_ALPHA + 2, q));
    }
#undef BB
}

// max of two values over the workgroup, written to part[2 * blockIdx.x .. + 1]
// (every thread of the workgroup calls it)
DEVI void group_max2(double a, double b, double *__restrict__ part) {
    __shared__ double wa[16], wb[16];
    for (int off = 32; off > 0; off >>= 1) {
        a = dmax(a, __shfl_xor(a, off, 64));
        b = dmax(b, __shfl_xor(b, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    if (lane == 0) {
        wa[wid] = a;
        wb[wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; ++w) {
            a = dmax(a, wa[w]);
            b = dmax(b, wb[w]);
        }
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

// the monitor's fold (one workgroup): mon[0] = the normalised div B of this
// pass (max |d2 B| / max |B|, as mercury_divb), mon[1] = its maximum since the
// last reset, mon[2] = max |d2 B|
__global__ void __launch_bounds__(1024) k_divb_monitor(const double *__restrict__ part, int ngroups, double *__restrict__ mon) {
    // eight independent loads in flight per thread: the fold is latency-bound
    double a = 0.0, b = 0.0;
    const int step = blockDim.x;
    int g = threadIdx.x;
    for (; g + 7 * step < ngroups; g += 8 * step) {
        double x[16];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x[2 * u] = part[2 * (g + u * step)];
            x[2 * u + 1] = part[2 * (g + u * step) + 1];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a = dmax(a, x[2 * u]);
            b = dmax(b, x[2 * u + 1]);
        }
    }
    for (; g < ngroups; g += step) {
        a = dmax(a, part[2 * g]);
        b = dmax(b, part[2 * g + 1]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a = dmax(a, __shfl_xor(a, off, 64));
        b = dmax(b, __shfl_xor(b, off, 64));
    }
    __shared__ double wa[16], wb[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        wa[wid] = a;
        wb[wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            a = dmax(a, wa[w]);
            b = dmax(b, wb[w]);
        }
        const double n = b > 0.0 ? a / b : a;
        mon[0] = n;
        mon[1] = dmax(mon[1], n);
        mon[2] = a;
    }
}
__global__ void k_divb_monitor_reset(double *mon) {
    if (threadIdx.x == 0) mon[0] = mon[1] = mon[2] = 0.0;
}

template <bool MIX>
__global__ void k_bcell_jedge(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                              const float *const *__restrict__ G_, double *__restrict__ divb_part) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    using T = typename Prec<MIX>::T;
    Pt p;
    const bool located = locate(R, B, TID, p);
    double dmx = 0.0, bmx = 0.0;  // the div B monitor: max |d2 B| over owned cells, max |B| over owned faces
    if (located) bcell_jedge_point<MIX>(F, G, B[p.b], p, divb_part != nullptr, dmx, bmx);
    if (divb_part) group_max2(dmx, bmx, divb_part);
}

// calc_Jcell (MercurySolver_Current.cpp:203-259), owned cells
template <bool MIX>
__global__ void k_jcell_t(const MercRange R, const MercBlk *__restrict__ B, double *const *__restrict__ F_,
                          const float *const *__restrict__ G_) {
    gdouble *const *__restrict__ F = gslots(F_);
    const gfloat *const *__restrict__ G = gslots(G_);
    using T = typename Prec<MIX>::T;
    LOCATE(R);
    const long long q = p.q, di = OI(p), dj = OJ(p);
    const T s[12] = {(T)F[SL_J + 0][q], (T)F[SL_J + 0][q + dj], (T)F[SL_J + 0][q + 1], (T)F[SL_J + 0][q + dj + 1],
                     (T)F[SL_J + 1][q], (T)F[SL_J + 1][q + di], (T)F[SL_J + 1][q + 1], (T)F[SL_J + 1][q + di + 1],
                     (T)F[SL_J + 2][q], (T)F[SL_J + 2][q + di], (T)F[SL_J + 2][q + dj], (T)F[SL_J + 2][q + di + dj]};
    T J[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int n = 0; n < 12; ++n)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[c] += st<MIX>(F, G, SL_WJ + 12 * c + n, q) * s[n];
    for (int c = 0; c < 3; ++c) F[SL_JC + c][q] = (double)J[c];
}

// FP32 copy of a static FP64 slot (mixed form)
__global__ void k_to_double(const float *__restrict__ src, double *__restrict__ dst, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (double)src[t];
}
__global__ void k_to_float(const double *__restrict__ src, float *__restrict__ dst, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (float)src[t];
}

// Commit the reduced dt of the step about to run, with Compute_Timestep's abort
// rule (MercurySolver_Timestep.cpp:129-140): a dt that is not finite or is below
// 1e-7 stops the run.  The stop is sticky in status[0] (1 = dt underflow, with
// the offending bits in status[2]); status[1] counts the steps committed since
// the status was cleared.  A stopped context commits dt = 0, and the update
// passes skip their writes when dt is 0 (no valid step has dt = 0), so the
// state stays that of the last completed step, as the reference exits before
// its update.
// One thread per run of the context (status: 4 words per run).
__global__ void k_dt_commit(unsigned long long *dt, const unsigned long long *dt_next, unsigned long long *status, int nruns) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    unsigned long long *st = status + 4 * r;
    const unsigned long long bits = dt_next[r];
    const double d = __longlong_as_double((long long)bits);
    if (st[0] != 0ULL) {
        dt[r] = 0ULL;
    } else if (!isfinite(d) || d < 1e-7) {
        st[0] = 1ULL;
        st[2] = bits;
        dt[r] = 0ULL;
    } else {
        dt[r] = bits;
        st[1] += 1ULL;
    }
}
__global__ void k_dt_reset(unsigned long long *dt_next, int nruns) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nruns) dt_next[r] = (unsigned long long)__double_as_longlong(1e100);
}

}  // namespace

namespace merc {

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
#define LAUNCH_R(kern, Rh, ...)                                                                  \
    do {                                                                                         \
        if ((Rh).total > 0) hipLaunchKernelGGL(kern, dim3(nblk((Rh).total, 256)), dim3(256), 0, st, (Rh).dev, __VA_ARGS__); \
    } while (0)

void launch_rhs_cell(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                     int L, bool mix) {
    const bool three = P.n_species >= 3.0;
    if (mix) {
        if (three) LAUNCH_R((k_rhs_cell<true, 3>), R, B, F, G, P, L);
        else LAUNCH_R((k_rhs_cell<true, 2>), R, B, F, G, P, L);
    } else {
        if (three) LAUNCH_R((k_rhs_cell<false, 3>), R, B, F, G, P, L);
        else LAUNCH_R((k_rhs_cell<false, 2>), R, B, F, G, P, L);
    }
}
void launch_rhs_tile(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                     int C, bool mix) {
    if (R.total <= 0) return;
    const dim3 grid((unsigned)R.total), blk(256);
    const bool three = P.n_species >= 3.0;
    if (mix) {
        if (three) hipLaunchKernelGGL((k_rhs_tile<true, 3>), grid, blk, 0, st, R.dev, B, F, G, P, C);
        else hipLaunchKernelGGL((k_rhs_tile<true, 2>), grid, blk, 0, st, R.dev, B, F, G, P, C);
    } else {
        if (three) hipLaunchKernelGGL((k_rhs_tile<false, 3>), grid, blk, 0, st, R.dev, B, F, G, P, C);
        else hipLaunchKernelGGL((k_rhs_tile<false, 2>), grid, blk, 0, st, R.dev, B, F, G, P, C);
    }
}
void launch_cell_end(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                     unsigned long long *dt_next, int compute_dt, bool mix, int nruns) {
    if (compute_dt) hipLaunchKernelGGL(k_dt_reset, dim3(nblk(nruns, 64)), dim3(64), 0, st, dt_next, nruns);
    const bool three = P.n_species >= 3.0, ens = nruns > 1;
#define CELL_END(M, N, E) LAUNCH_R((k_cell_end<M, N, E>), R, B, F, G, P, dt_next, compute_dt)
    if (mix) {
        if (three) { if (ens) CELL_END(true, 3, true); else CELL_END(true, 3, false); }
        else { if (ens) CELL_END(true, 2, true); else CELL_END(true, 2, false); }
    } else {
        if (three) { if (ens) CELL_END(false, 3, true); else CELL_END(false, 3, false); }
        else { if (ens) CELL_END(false, 2, true); else CELL_END(false, 2, false); }
    }
#undef CELL_END
}
void launch_eface_u(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, bool mix, int ambi,
                    int ns) {
    if (mix) LAUNCH_R(k_eface_u<true>, R, B, F, G, ambi, ns);
    else LAUNCH_R(k_eface_u<false>, R, B, F, G, ambi, ns);
}
void launch_edge_total(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const MercPhys &P,
                       int with_hall, bool mix) {
    if (mix) LAUNCH_R(k_edge_total<true>, R, B, F, G, P, with_hall);
    else LAUNCH_R(k_edge_total<false>, R, B, F, G, P, with_hall);
}
void launch_update_all(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, const double *dt,
                       bool mix, int ns) {
    if (mix) LAUNCH_R(k_update_all<true>, R, B, F, G, dt, ns);
    else LAUNCH_R(k_update_all<false>, R, B, F, G, dt, ns);
}
void launch_to_double(hipStream_t st, const float *src, double *dst, long long n) {
    if (n > 0) hipLaunchKernelGGL(k_to_double, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
}
void launch_bcell_jedge(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, bool mix,
                        double *divb_part, double *divb_mon) {
    if (mix) LAUNCH_R(k_bcell_jedge<true>, R, B, F, G, divb_part);
    else LAUNCH_R(k_bcell_jedge<false>, R, B, F, G, divb_part);
    if (divb_part && divb_mon && R.total > 0)
        hipLaunchKernelGGL(k_divb_monitor, dim3(1), dim3(1024), 0, st, divb_part, (int)nblk(R.total, 256), divb_mon);
}
void launch_divb_monitor_reset(hipStream_t st, double *divb_mon) {
    hipLaunchKernelGGL(k_divb_monitor_reset, dim3(1), dim3(64), 0, st, divb_mon);
}
void launch_jcell_t(hipStream_t st, const RangeH &R, const MercBlk *B, double *const *F, const float *const *G, bool mix) {
    if (mix) LAUNCH_R(k_jcell_t<true>, R, B, F, G);
    else LAUNCH_R(k_jcell_t<false>, R, B, F, G);
}
void launch_to_float(hipStream_t st, const double *src, float *dst, long long n) {
    if (n > 0) hipLaunchKernelGGL(k_to_float, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
}
void launch_dt_commit(hipStream_t st, unsigned long long *dt, const unsigned long long *dt_next, unsigned long long *status, int nruns) {
    hipLaunchKernelGGL(k_dt_commit, dim3(nblk(nruns, 64)), dim3(64), 0, st, dt, dt_next, status, nruns);
}

}  // namespace merc
