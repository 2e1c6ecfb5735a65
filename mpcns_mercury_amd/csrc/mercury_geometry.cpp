// Host-side static geometry of the Mercury step (computed once, uploaded).
//
// METRIC::compute_metric_fields (src/7_metric/Metric.cpp:78-1414),
// Block::Add_ghostmesh (src/1_grid/1_MPCNS_Grid.cpp:591-818),
// Block::calc_Dual_Grids (src/1_grid/1_MPCNS_Grid_GCL.cpp:6-24),
// SetupCellReconstructionWeights_ (Z4_Mercury/src/MercurySolver_ReconstructionSetup.cpp:29-193),
// HallRadialTaperEdge_ (Z4_Mercury/src/physics/MercurySolver_MagExplicit_Hall.cpp:6-25) and the
// local artificial eta cache (Z4_Mercury/src/physics/MercurySolver_MagExplicit_Resis.cpp:87-144).
//
// These are setup-only, FP64, and use libm transcendental functions, so they
// stay on the host; the device receives the finished per-DOF arrays.
#include "mercury_geometry.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

namespace merc {

using V3 = std::array<double, 3>;

static inline V3 sub(const V3 &a, const V3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
static inline V3 crs(const V3 &a, const V3 &b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
static inline double nrm(const V3 &a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline double dot(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double maxd(double a, double b) { return (a < b) ? b : a; }
static inline double mind(double a, double b) { return (b < a) ? b : a; }

BlockGeom::BlockGeom(const int m_[3], int g_, const double *owned_nodes) : g(g_) {
    for (int a = 0; a < 3; ++a) {
        m[a] = m_[a];
        N[a] = m[a] + 1 + 2 * g;
        Nn[a] = m[a] + 1 + 2 * (g + 1);
    }
    V = (long long)N[0] * N[1] * N[2];
    Vn = (long long)Nn[0] * Nn[1] * Nn[2];
    for (auto &c : X) c.assign(Vn, 0.0);
    const long long nn = (long long)(m[0] + 1) * (m[1] + 1) * (m[2] + 1);
    long long q = 0;
    for (int i = 0; i <= m[0]; ++i)
        for (int j = 0; j <= m[1]; ++j)
            for (int k = 0; k <= m[2]; ++k, ++q)
                for (int c = 0; c < 3; ++c) X[c][nidx(i, j, k)] = owned_nodes[c * nn + q];
    add_ghost_mesh();
}

// Block::Add_ghostmesh, three-dimensional branch: one mirror layer at a time,
// faces, then the k-, j- and i-direction edge lines.
void BlockGeom::add_ghost_mesh() {
    const int mx = m[0], my = m[1], mz = m[2];
    for (int gh = 1; gh <= g + 1; ++gh) {
        const int i0 = -gh, i1 = mx + gh, j0 = -gh, j1 = my + gh, k0 = -gh, k1 = mz + gh;
        for (auto &C : X) {
            auto P = [&](int i, int j, int k) -> double & { return C[nidx(i, j, k)]; };
            for (int j = j0 + 1; j < j1; j++)
                for (int k = k0 + 1; k < k1; k++) {
                    P(i0, j, k) = 2 * P(0, j, k) - P(-i0, j, k);
                    P(i1, j, k) = 2 * P(mx, j, k) - P(2 * mx - i1, j, k);
                }
            for (int i = i0 + 1; i < i1; i++)
                for (int k = k0 + 1; k < k1; k++) {
                    P(i, j0, k) = 2 * P(i, 0, k) - P(i, -j0, k);
                    P(i, j1, k) = 2 * P(i, my, k) - P(i, 2 * my - j1, k);
                }
            for (int i = i0 + 1; i < i1; i++)
                for (int j = j0 + 1; j < j1; j++) {
                    P(i, j, k0) = 2 * P(i, j, 0) - P(i, j, -k0);
                    P(i, j, k1) = 2 * P(i, j, mz) - P(i, j, 2 * mz - k1);
                }
            for (int k = k0 + 1; k < k1; k++) {
                P(i0, j0, k) = P(i0 + 1, j0, k) + P(i0, j0 + 1, k) - P(i0 + 1, j0 + 1, k);
                P(i1, j0, k) = P(i1 - 1, j0, k) + P(i1, j0 + 1, k) - P(i1 - 1, j0 + 1, k);
                P(i0, j1, k) = P(i0 + 1, j1, k) + P(i0, j1 - 1, k) - P(i0 + 1, j1 - 1, k);
                P(i1, j1, k) = P(i1 - 1, j1, k) + P(i1, j1 - 1, k) - P(i1 - 1, j1 - 1, k);
            }
            for (int j = j0 + 1; j < j1; j++) {
                P(i0, j, k0) = P(i0 + 1, j, k0) + P(i0, j, k0 + 1) - P(i0 + 1, j, k0 + 1);
                P(i1, j, k0) = P(i1 - 1, j, k0) + P(i1, j, k0 + 1) - P(i1 - 1, j, k0 + 1);
                P(i0, j, k1) = P(i0 + 1, j, k1) + P(i0, j, k1 - 1) - P(i0 + 1, j, k1 - 1);
                P(i1, j, k1) = P(i1 - 1, j, k1) + P(i1, j, k1 - 1) - P(i1 - 1, j, k1 - 1);
            }
            for (int i = i0; i < i1 + 1; i++) {
                P(i, j0, k0) = P(i, j0 + 1, k0) + P(i, j0, k0 + 1) - P(i, j0 + 1, k0 + 1);
                P(i, j1, k0) = P(i, j1 - 1, k0) + P(i, j1, k0 + 1) - P(i, j1 - 1, k0 + 1);
                P(i, j0, k1) = P(i, j0 + 1, k1) + P(i, j0, k1 - 1) - P(i, j0 + 1, k1 - 1);
                P(i, j1, k1) = P(i, j1 - 1, k1) + P(i, j1, k1 - 1) - P(i, j1 - 1, k1 - 1);
            }
        }
    }
}

V3 BlockGeom::node(int i, int j, int k) const {
    const long long q = nidx(i, j, k);
    return {X[0][q], X[1][q], X[2][q]};
}

void BlockGeom::compute(const MercPhys &P, SlotWriter &out) {
    const int gg = g;
    auto alloc = [&](int dx, int dy, int dz, int G, int lo[3], int hi[3]) {
        const int d[3] = {dx, dy, dz};
        for (int a = 0; a < 3; ++a) {
            lo[a] = -G;
            hi[a] = m[a] + 1 - d[a] + G;
        }
    };
    static const int DL[7][3] = {{1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    // cell centres (dual grid) for cells [-g, m+g)
    std::vector<double> cc[3];
    for (auto &c : cc) c.assign(V, 0.0);
    for (int i = -gg; i < m[0] + gg; ++i)
        for (int j = -gg; j < m[1] + gg; ++j)
            for (int k = -gg; k < m[2] + gg; ++k)
                for (int c = 0; c < 3; ++c) {
                    const auto &Q = X[c];
                    auto q = [&](int a, int b, int e) { return Q[nidx(a, b, e)]; };
                    cc[c][idx(i, j, k)] = 0.125 * (q(i + 1, j + 1, k + 1) + q(i, j + 1, k + 1) + q(i + 1, j, k + 1) + q(i, j, k + 1) +
                                                   q(i + 1, j + 1, k) + q(i, j + 1, k) + q(i + 1, j, k) + q(i, j, k));
                }
    auto cell = [&](int i, int j, int k) -> V3 {
        const long long q = idx(i, j, k);
        return {cc[0][q], cc[1][q], cc[2][q]};
    };
    for (int c = 0; c < 3; ++c) out.put(SL_CC + c, cc[c]);

    int lo[3], hi[3];
    // face area vectors JDxi/JDet/JDze, ghost g (Metric.cpp:109-211)
    auto &JD = JD_;
    for (auto &v : JD) v.assign(V, 0.0);
    for (int a = 0; a < 3; ++a) {
        alloc(DL[1 + a][0], DL[1 + a][1], DL[1 + a][2], gg, lo, hi);
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    V3 o = node(i, j, k), pa, pb, pab;
                    if (a == 0) { pa = node(i, j + 1, k); pb = node(i, j, k + 1); pab = node(i, j + 1, k + 1); }
                    else if (a == 1) { pa = node(i, j, k + 1); pb = node(i + 1, j, k); pab = node(i + 1, j, k + 1); }
                    else { pa = node(i + 1, j, k); pb = node(i, j + 1, k); pab = node(i + 1, j + 1, k); }
                    const V3 c1 = crs(sub(pa, o), sub(pb, o)), c2 = crs(sub(pb, pab), sub(pa, pab));
                    for (int c = 0; c < 3; ++c) JD[3 * a + c][idx(i, j, k)] = 0.5 * (c1[c] + c2[c]);
                }
    }
    auto jd = [&](int a, int i, int j, int k) -> V3 {
        const long long q = idx(i, j, k);
        return {JD[3 * a][q], JD[3 * a + 1][q], JD[3 * a + 2][q]};
    };
    // Jac = 1/3 sum over the six faces of A_f . (x_f - x_c), ghost g-1 (Metric.cpp:214-326)
    std::vector<double> Jac(V, 0.0);
    alloc(1, 1, 1, gg - 1, lo, hi);
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) {
                const V3 xc = cell(i, j, k);
                double Vol = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int s = 0; s < 2; ++s) {
                        const int fi = i + (a == 0 ? s : 0), fj = j + (a == 1 ? s : 0), fk = k + (a == 2 ? s : 0);
                        V3 Af = jd(a, fi, fj, fk);
                        if (!s) Af = {-Af[0], -Af[1], -Af[2]};
                        V3 n0, n1, n2, n3;
                        if (a == 0) { n0 = node(fi, j, k); n1 = node(fi, j + 1, k); n2 = node(fi, j, k + 1); n3 = node(fi, j + 1, k + 1); }
                        else if (a == 1) { n0 = node(i, fj, k); n1 = node(i + 1, fj, k); n2 = node(i, fj, k + 1); n3 = node(i + 1, fj, k + 1); }
                        else { n0 = node(i, j, fk); n1 = node(i + 1, j, fk); n2 = node(i, j + 1, fk); n3 = node(i + 1, j + 1, fk); }
                        V3 xf;
                        for (int c = 0; c < 3; ++c) xf[c] = 0.25 * (n0[c] + n1[c] + n2[c] + n3[c]);
                        Vol += dot(Af, sub(xf, xc));
                    }
                Jac[idx(i, j, k)] = Vol / 3.0;
            }
    out.put(SL_JAC, Jac);
    // |S| and face Hodge beta = |l*|/|S|, ghost g-1 (Metric.cpp:397-489)
    const double eps = 1e-25;
    for (int a = 0; a < 3; ++a) {
        std::vector<double> Area(V, 0.0), beta(V, 0.0);
        alloc(DL[1 + a][0], DL[1 + a][1], DL[1 + a][2], gg - 1, lo, hi);
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    const double Smag = nrm(jd(a, i, j, k));
                    const double Ls = nrm(sub(cell(i, j, k), cell(i - (a == 0), j - (a == 1), k - (a == 2))));
                    Area[idx(i, j, k)] = Smag;
                    beta[idx(i, j, k)] = (Smag < eps) ? 0.0 : Ls / Smag;
                }
        out.put(SL_AREA + a, Area);
        out.put(SL_BETA + a, beta);
    }
    // dl, dr and the dual-collapse-capped inverse Hodge alpha = |e|/|S*|, ghost g-1 (Metric.cpp:511-680, 1261-1326)
    const double RATIO = 1e-3, EABS = 1e-12;
    for (int a = 0; a < 3; ++a) {
        std::vector<double> dl(V, 0.0), al(V, 0.0), dr[3];
        for (auto &v : dr) v.assign(V, 0.0);
        alloc(DL[4 + a][0], DL[4 + a][1], DL[4 + a][2], gg - 1, lo, hi);
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    const V3 e = sub(node(i + (a == 0), j + (a == 1), k + (a == 2)), node(i, j, k));
                    const double L = nrm(e);
                    V3 p00, p10, p01, p11;
                    if (a == 0) { p00 = cell(i, j - 1, k - 1); p10 = cell(i, j, k - 1); p01 = cell(i, j - 1, k); p11 = cell(i, j, k); }
                    else if (a == 1) { p00 = cell(i - 1, j, k - 1); p10 = cell(i, j, k - 1); p01 = cell(i - 1, j, k); p11 = cell(i, j, k); }
                    else { p00 = cell(i - 1, j - 1, k); p10 = cell(i, j - 1, k); p01 = cell(i - 1, j, k); p11 = cell(i, j, k); }
                    const V3 a0 = crs(sub(p10, p00), sub(p01, p00)), a1 = crs(sub(p01, p11), sub(p10, p11));
                    const V3 Av = {0.5 * (a0[0] + a1[0]), 0.5 * (a0[1] + a1[1]), 0.5 * (a0[2] + a1[2])};
                    const double Amag = nrm(Av);
                    const double raw = (Amag < eps) ? 0.0 : (L / Amag);
                    const double d00_10 = nrm(sub(p10, p00)), d00_01 = nrm(sub(p01, p00)), d10_11 = nrm(sub(p11, p10)), d01_11 = nrm(sub(p11, p01));
                    const double d00_11 = nrm(sub(p11, p00)), d10_01 = nrm(sub(p01, p10));
                    const double href = maxd(maxd(maxd(d00_10, d00_01), maxd(d10_11, d01_11)), maxd(maxd(d00_11, d10_01), maxd(L, eps)));
                    const double side_min = mind(mind(d00_10, d00_01), mind(d10_11, d01_11));
                    const double diag_min = mind(d00_11, d10_01);
                    const bool touch = (L < EABS) || (side_min < RATIO * href) || (diag_min < RATIO * href) || (Amag < RATIO * href * href);
                    const long long q = idx(i, j, k);
                    dl[q] = L;
                    al[q] = touch ? 0.0 : raw;
                    for (int c = 0; c < 3; ++c) dr[c][q] = e[c];
                }
        out.put(SL_DL + a, dl);
        out.put(SL_ALPHA + a, al);
        for (int c = 0; c < 3; ++c) out.put(SL_DR + 3 * a + c, dr[c]);
    }
    for (int c = 0; c < 9; ++c) out.put(SL_JD + c, JD[c]);

    // reconstruction weights on owned cells
    {
        std::vector<double> Wb[18], Wj[36];
        for (auto &v : Wb) v.assign(V, 0.0);
        for (auto &v : Wj) v.assign(V, 0.0);
        const double epsw = 1.0e-30, eig_tol = 1.0e-10;
        for (int i = 0; i < m[0]; ++i)
            for (int j = 0; j < m[1]; ++j)
                for (int k = 0; k < m[2]; ++k) {
                    const long long q = idx(i, j, k);
                    V3 s[6];
                    {
                        const V3 a0 = jd(0, i, j, k), a1 = jd(0, i + 1, j, k), b0 = jd(1, i, j, k), b1 = jd(1, i, j + 1, k), c0 = jd(2, i, j, k),
                                 c1 = jd(2, i, j, k + 1);
                        s[0] = {-a0[0], -a0[1], -a0[2]};
                        s[1] = a1;
                        s[2] = {-b0[0], -b0[1], -b0[2]};
                        s[3] = b1;
                        s[4] = {-c0[0], -c0[1], -c0[2]};
                        s[5] = c1;
                    }
                    V3 xc;
                    {
                        const V3 n000 = node(i, j, k), n100 = node(i + 1, j, k), n010 = node(i, j + 1, k), n110 = node(i + 1, j + 1, k);
                        const V3 n001 = node(i, j, k + 1), n101 = node(i + 1, j, k + 1), n011 = node(i, j + 1, k + 1), n111 = node(i + 1, j + 1, k + 1);
                        for (int c = 0; c < 3; ++c)
                            xc[c] = 0.125 * (n000[c] + n100[c] + n010[c] + n110[c] + n001[c] + n101[c] + n011[c] + n111[c]);
                    }
                    auto fcen = [&](V3 p0, V3 p1, V3 p2, V3 p3) {
                        return V3{0.25 * (p0[0] + p1[0] + p2[0] + p3[0]), 0.25 * (p0[1] + p1[1] + p2[1] + p3[1]),
                                  0.25 * (p0[2] + p1[2] + p2[2] + p3[2])};
                    };
                    V3 fc[6];
                    fc[0] = fcen(node(i, j, k), node(i, j + 1, k), node(i, j, k + 1), node(i, j + 1, k + 1));
                    fc[1] = fcen(node(i + 1, j, k), node(i + 1, j + 1, k), node(i + 1, j, k + 1), node(i + 1, j + 1, k + 1));
                    fc[2] = fcen(node(i, j, k), node(i + 1, j, k), node(i, j, k + 1), node(i + 1, j, k + 1));
                    fc[3] = fcen(node(i, j + 1, k), node(i + 1, j + 1, k), node(i, j + 1, k + 1), node(i + 1, j + 1, k + 1));
                    fc[4] = fcen(node(i, j, k), node(i + 1, j, k), node(i, j + 1, k), node(i + 1, j + 1, k));
                    fc[5] = fcen(node(i, j, k + 1), node(i + 1, j, k + 1), node(i, j + 1, k + 1), node(i + 1, j + 1, k + 1));
                    double smag[6], mu[6], sref = 0.0;
                    for (int n = 0; n < 6; ++n) {
                        smag[n] = std::sqrt(s[n][0] * s[n][0] + s[n][1] * s[n][1] + s[n][2] * s[n][2]);
                        sref = maxd(sref, smag[n]);
                        const double dx = fc[n][0] - xc[0], dy = fc[n][1] - xc[1], dz = fc[n][2] - xc[2];
                        mu[n] = 1.0 / maxd(std::sqrt(dx * dx + dy * dy + dz * dz), 1.0e-14);
                    }
                    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    bool act[6];
                    for (int n = 0; n < 6; ++n) {
                        act[n] = smag[n] > 1.0e-10 * sref;
                        if (!act[n]) continue;
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) M[a][c] += mu[n] * s[n][a] * s[n][c];
                    }
                    double Vv[3][3], lam[3];
                    jacobi_eigen3(M, Vv, lam);
                    const double lmx = maxd(lam[0], maxd(lam[1], lam[2]));
                    if (!(lmx <= epsw)) {
                        double Mi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                        for (int t = 0; t < 3; ++t) {
                            const double inv = (lam[t] > eig_tol * lmx) ? 1.0 / lam[t] : 0.0;
                            for (int a = 0; a < 3; ++a)
                                for (int c = 0; c < 3; ++c) Mi[a][c] += Vv[a][t] * inv * Vv[c][t];
                        }
                        for (int n = 0; n < 6; ++n)
                            if (act[n])
                                for (int a = 0; a < 3; ++a) Wb[6 * a + n][q] = mu[n] * (Mi[a][0] * s[n][0] + Mi[a][1] * s[n][1] + Mi[a][2] * s[n][2]);
                    }
                    // twelve cell edges: xi (4), eta (4), zeta (4)
                    static const int ED[12][6] = {{0, 0, 0, 1, 0, 0}, {0, 1, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 1}, {0, 1, 1, 1, 1, 1},
                                                  {0, 0, 0, 0, 1, 0}, {1, 0, 0, 1, 1, 0}, {0, 0, 1, 0, 1, 1}, {1, 0, 1, 1, 1, 1},
                                                  {0, 0, 0, 0, 0, 1}, {1, 0, 0, 1, 0, 1}, {0, 1, 0, 0, 1, 1}, {1, 1, 0, 1, 1, 1}};
                    V3 l[12];
                    double lmax = 0.0;
                    for (int n = 0; n < 12; ++n) {
                        l[n] = sub(node(i + ED[n][3], j + ED[n][4], k + ED[n][5]), node(i + ED[n][0], j + ED[n][1], k + ED[n][2]));
                        lmax = maxd(lmax, std::sqrt(l[n][0] * l[n][0] + l[n][1] * l[n][1] + l[n][2] * l[n][2]));
                    }
                    if (lmax <= epsw) continue;
                    double Mj[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    bool actj[12];
                    for (int n = 0; n < 12; ++n) {
                        const double ln = std::sqrt(l[n][0] * l[n][0] + l[n][1] * l[n][1] + l[n][2] * l[n][2]);
                        actj[n] = ln > 1.0e-10 * lmax;
                        if (!actj[n]) continue;
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) Mj[a][c] += l[n][a] * l[n][c];
                    }
                    jacobi_eigen3(Mj, Vv, lam);
                    const double lme = maxd(lam[0], maxd(lam[1], lam[2]));
                    if (lme <= epsw) continue;
                    double Mi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    for (int t = 0; t < 3; ++t) {
                        const double inv = (lam[t] > eig_tol * lme) ? 1.0 / lam[t] : 0.0;
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) Mi[a][c] += Vv[a][t] * inv * Vv[c][t];
                    }
                    for (int n = 0; n < 12; ++n)
                        if (actj[n])
                            for (int a = 0; a < 3; ++a) Wj[12 * a + n][q] = Mi[a][0] * l[n][0] + Mi[a][1] * l[n][1] + Mi[a][2] * l[n][2];
                }
        for (int c = 0; c < 18; ++c) out.put(SL_WB + c, Wb[c]);
        for (int c = 0; c < 36; ++c) out.put(SL_WJ + c, Wj[c]);
    }

    // Hall taper and local artificial eta per owned edge
    for (int a = 0; a < 3; ++a) {
        std::vector<double> tp(V, 0.0), le(V, 0.0);
        const int ohi[3] = {m[0] + (a == 0 ? 0 : 1), m[1] + (a == 1 ? 0 : 1), m[2] + (a == 2 ? 0 : 1)};
        const bool local = P.local_enabled != 0.0 && P.local_eta_max > 0.0 && P.local_r_decay > 0.0;
        for (int i = 0; i < ohi[0]; ++i)
            for (int j = 0; j < ohi[1]; ++j)
                for (int k = 0; k < ohi[2]; ++k) {
                    const V3 n0 = node(i, j, k), n1 = node(i + (a == 0), j + (a == 1), k + (a == 2));
                    const double xm = 0.5 * (n0[0] + n1[0]), ym = 0.5 * (n0[1] + n1[1]), zm = 0.5 * (n0[2] + n1[2]);
                    const double r = std::sqrt(xm * xm + ym * ym + zm * zm);
                    const double width = P.taper_rmax - P.taper_rmin;
                    double t;
                    if (width <= 0.0)
                        t = (r >= P.taper_rmin) ? 1.0 : 0.0;
                    else {
                        const double mid = 0.5 * (P.taper_rmin + P.taper_rmax);
                        t = 0.5 * (1.0 + std::tanh(2.0 * (r - mid) / width));
                    }
                    const long long q = idx(i, j, k);
                    tp[q] = t;
                    double eta = 0.0;
                    if (local) {
                        const double dx = xm - P.local_center[0], dy = ym - P.local_center[1], dz = zm - P.local_center[2];
                        const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (P.local_r_cutoff > 0.0 && rr >= P.local_r_cutoff)
                            eta = 0.0;
                        else
                            eta = P.local_eta_max * std::exp(-rr / P.local_r_decay);
                    }
                    le[q] = eta;
                }
        out.put(SL_TAPER + a, tp);
        out.put(SL_LETA + a, le);
    }
}

// 3x3 symmetric Jacobi eigen-decomposition, 18 sweeps
// (Z4_Mercury/src/MercurySolver_ReconstructionSetup.cpp:34-61)
void jacobi_eigen3(const double Min[3][3], double V[3][3], double d[3]) {
    double A[3][3];
    std::memcpy(A, Min, sizeof(A));
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) V[a][b] = (a == b) ? 1.0 : 0.0;
    for (int it = 0; it < 18; ++it) {
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = A[p][q];
                if (std::abs(apq) <= 1.0e-16 * (std::abs(A[p][p]) + std::abs(A[q][q]) + 1.0)) continue;
                const double phi = 0.5 * std::atan2(2.0 * apq, A[q][q] - A[p][p]);
                // one sincos call, as g++ emits for the reference's adjacent cos/sin of
                // the same angle (glibc's sincos may differ from sin, cos in the last bit)
                double s, c;
                ::sincos(phi, &s, &c);
                const double app = A[p][p], aqq = A[q][q];
                for (int k = 0; k < 3; ++k)
                    if (k != p && k != q) {
                        const double akp = A[k][p], akq = A[k][q];
                        A[k][p] = A[p][k] = c * akp - s * akq;
                        A[k][q] = A[q][k] = s * akp + c * akq;
                    }
                A[p][p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
                A[q][q] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
                A[p][q] = A[q][p] = 0.0;
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    d[0] = A[0][0];
    d[1] = A[1][1];
    d[2] = A[2][2];
}

// DipoleField::Build_Badd_FaceFlux steps 1-2 (Z4_Mercury/include/operators/Dipole.h:128-324)
// with A_of_dipole (:25-39): the edge 1-form A.dr of the dipoles' vector
// potential (midpoint value times the edge vector) on the edge boxes of ghost
// depth 1, and its curl (d1) on the face boxes of ghost depth 1, added to the
// three face arrays in badd (one rounded addition each).  dip: 7 values per
// dipole, x0 y0 z0 mx my mz a, nondimensional.
void BlockGeom::add_dipole_flux(int nd, const double *dip, std::array<std::vector<double>, 3> &badd) const {
    if (nd <= 0) return;
    static const int DE[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::array<std::vector<double>, 3> Ae;
    for (int a = 0; a < 3; ++a) {
        Ae[a].assign(V, 0.0);
        for (int i = -1; i < m[0] + 2 - DE[a][0]; ++i)
            for (int j = -1; j < m[1] + 2 - DE[a][1]; ++j)
                for (int k = -1; k < m[2] + 2 - DE[a][2]; ++k) {
                    const V3 x1 = node(i, j, k), x2 = node(i + DE[a][0], j + DE[a][1], k + DE[a][2]);
                    const V3 xm = {(x1[0] + x2[0]) * 0.5, (x1[1] + x2[1]) * 0.5, (x1[2] + x2[2]) * 0.5};
                    const V3 dr = sub(x2, x1);
                    V3 A = {0.0, 0.0, 0.0};
                    for (int d = 0; d < nd; ++d) {
                        const double *D = dip + 7 * d;
                        const V3 r = {xm[0] - D[0], xm[1] - D[1], xm[2] - D[2]};
                        double f = (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + D[6] * D[6];
                        f = std::sqrt(f) * f;
                        const V3 c = crs({D[3], D[4], D[5]}, r);
                        for (int t = 0; t < 3; ++t) A[t] += c[t] / f;
                    }
                    Ae[a][idx(i, j, k)] = dot(A, dr);
                }
    }
    auto E = [&](int c, int i, int j, int k) { return Ae[c][idx(i, j, k)]; };
    for (int a = 0; a < 3; ++a)
        for (int i = -1; i < m[0] + 1 + (a == 0); ++i)
            for (int j = -1; j < m[1] + 1 + (a == 1); ++j)
                for (int k = -1; k < m[2] + 1 + (a == 2); ++k) {
                    double c;
                    if (a == 0) c = (E(1, i, j, k) - E(1, i, j, k + 1)) + (E(2, i, j + 1, k) - E(2, i, j, k));
                    else if (a == 1) c = (E(0, i, j, k + 1) - E(0, i, j, k)) + (E(2, i, j, k) - E(2, i + 1, j, k));
                    else c = (E(0, i, j, k) - E(0, i, j + 1, k)) + (E(1, i + 1, j, k) - E(1, i, j, k));
                    double &v = badd[a][idx(i, j, k)];
                    v = c + v;
                }
}

}  // namespace merc
