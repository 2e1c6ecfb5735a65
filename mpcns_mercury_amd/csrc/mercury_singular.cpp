// Singular-edge registry (SingularEdgeRegistry::build, src/7_metric/
// SingularEdgeRegistry.cpp:78-268; validate_or_abort :270-282).
#include "mercury_singular.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>

namespace merc {
namespace singular {
namespace {

using Key = std::tuple<int, int, int, int, int>;  // block, i, j, k, axis

bool owned(const std::array<int, 3> &m, int L, const int p[3]) {
    for (int a = 0; a < 3; ++a) {
        const int hi = m[a] + 1 - ((L == LOC_CELL) || (L >= LOC_EX && L <= LOC_EZ && L - LOC_EX == a) ||
                                          (L >= LOC_FX && L <= LOC_FZ && L - LOC_FX != a)
                                      ? 1
                                      : 0);
        if (p[a] < 0 || p[a] >= hi) return false;
    }
    return true;
}

// incidence of a face on an edge of the same block (AssembleSingularEdgeCurrent_,
// Z4_Mercury/src/physics/MercurySolver_Current.cpp:64-110): +1 or -1 as the
// face's normal circulates positively or negatively about the edge
int incidence(int ea, const int e[3], int fa, const int f[3]) {
    if (ea == 0) {
        if (fa == 1 && f[0] == e[0] && f[1] == e[1]) return f[2] == e[2] - 1 ? 1 : (f[2] == e[2] ? -1 : 0);
        if (fa == 2 && f[0] == e[0] && f[2] == e[2]) return f[1] == e[1] ? 1 : (f[1] == e[1] - 1 ? -1 : 0);
    } else if (ea == 1) {
        if (fa == 0 && f[0] == e[0] && f[1] == e[1]) return f[2] == e[2] ? 1 : (f[2] == e[2] - 1 ? -1 : 0);
        if (fa == 2 && f[1] == e[1] && f[2] == e[2]) return f[0] == e[0] - 1 ? 1 : (f[0] == e[0] ? -1 : 0);
    } else {
        if (fa == 0 && f[0] == e[0] && f[2] == e[2]) return f[1] == e[1] - 1 ? 1 : (f[1] == e[1] ? -1 : 0);
        if (fa == 1 && f[1] == e[1] && f[2] == e[2]) return f[0] == e[0] ? 1 : (f[0] == e[0] - 1 ? -1 : 0);
    }
    return 0;
}

// the four faces around an edge, in the reference's order (incident_faces,
// SingularEdgeRegistry.cpp:43-68)
void faces_around(int ax, const int e[3], int out[4][4]) {
    const int i = e[0], j = e[1], k = e[2];
    const int t[3][4][4] = {{{1, i, j, k}, {1, i, j, k - 1}, {2, i, j, k}, {2, i, j - 1, k}},
                            {{0, i, j, k}, {0, i, j, k - 1}, {2, i, j, k}, {2, i - 1, j, k}},
                            {{0, i, j, k}, {0, i, j - 1, k}, {1, i, j, k}, {1, i - 1, j, k}}};
    for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v) out[u][v] = t[ax][u][v];
}

// area of the polygon through the points ordered by angle about the tangent
// (the dual face of the edge)
double polygon_area(std::vector<std::array<double, 3>> pts, const double dr[3]) {
    std::array<double, 3> ctr{{0.0, 0.0, 0.0}};
    for (const auto &p : pts)
        for (int q = 0; q < 3; ++q) ctr[q] += p[q];
    for (int q = 0; q < 3; ++q) ctr[q] /= pts.size();
    double t[3] = {dr[0], dr[1], dr[2]};
    const double tn = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    if (tn > 0.0)
        for (double &v : t) v /= tn;
    double ref[3] = {1.0, 0.0, 0.0};
    if (!(std::abs(t[0]) < 0.9)) ref[0] = 0.0, ref[1] = 1.0;
    double u[3] = {t[1] * ref[2] - t[2] * ref[1], t[2] * ref[0] - t[0] * ref[2], t[0] * ref[1] - t[1] * ref[0]};
    const double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (un > 0.0)
        for (double &v : u) v /= un;
    const double v[3] = {t[1] * u[2] - t[2] * u[1], t[2] * u[0] - t[0] * u[2], t[0] * u[1] - t[1] * u[0]};
    auto angle = [&](const std::array<double, 3> &p) {
        double du = 0.0, dv = 0.0;
        for (int q = 0; q < 3; ++q) {
            du += (p[q] - ctr[q]) * u[q];
            dv += (p[q] - ctr[q]) * v[q];
        }
        return std::atan2(dv, du);
    };
    std::sort(pts.begin(), pts.end(), [&](const auto &a, const auto &b) { return angle(a) < angle(b); });
    double av[3] = {0.0, 0.0, 0.0};
    for (size_t n = 0; n < pts.size(); ++n) {
        const auto &a = pts[n], &b = pts[(n + 1) % pts.size()];
        av[0] += a[1] * b[2] - a[2] * b[1];
        av[1] += a[2] * b[0] - a[0] * b[2];
        av[2] += a[0] * b[1] - a[1] * b[0];
    }
    return 0.5 * std::sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
}

}  // namespace

void set_block_geo(BlockGeo &g, const std::array<int, 3> &m, int ghost, const std::vector<double> *a) {
    for (int u = 0; u < 3; ++u) g.m[u] = m[u];
    g.g = ghost;
    for (int u = 0; u < 3; ++u) g.dl[u] = a[u].data();
    for (int u = 0; u < 9; ++u) g.dr[u] = a[3 + u].data();
    g.jac = a[12].data();
    for (int u = 0; u < 3; ++u) g.cc[u] = a[13 + u].data();
    for (int u = 0; u < 3; ++u) g.area[u] = a[16 + u].data();
}

double hall_taper(const MercPhys &P, double x, double y, double z) {
    const double r = std::sqrt(x * x + y * y + z * z);
    const double width = P.taper_rmax - P.taper_rmin;
    if (width <= 0.0) return (r >= P.taper_rmin) ? 1.0 : 0.0;
    const double mid = 0.5 * (P.taper_rmin + P.taper_rmax);
    return 0.5 * (1.0 + std::tanh(2.0 * (r - mid) / width));
}

bool local_eta(const MercPhys &P, double x, double y, double z, double &eta) {
    eta = 0.0;
    if (!(P.local_enabled != 0.0 && P.local_eta_max > 0.0)) return false;
    const double dx = x - P.local_center[0], dy = y - P.local_center[1], dz = z - P.local_center[2];
    const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (!(P.local_r_cutoff <= 0.0 || rr < P.local_r_cutoff)) return false;
    eta = P.local_eta_max * std::exp(-rr / P.local_r_decay);
    return true;
}

Registry build(const patch::Topo &T, const std::vector<BlockGeo> &G) {
    Registry R;
    const patch::EdgeClasses E = patch::edge_classes(T);
    const patch::EdgeClasses FC = patch::face_classes(T);
    std::set<Key> face_alias;  // later members of face classes: not owners
    for (size_t c = 0; c + 1 < FC.start.size(); ++c)
        for (int t = FC.start[c] + 1; t < FC.start[c + 1]; ++t) {
            const patch::Dof &d = FC.mem[t];
            face_alias.insert(Key{d.b, d.i, d.j, d.k, d.f});
        }
    for (size_t c = 0; c + 1 < E.start.size(); ++c) {
        // a candidate has three or more block-local aliases; the valence is
        // classified by the owner faces around them below
        if (E.start[c + 1] - E.start[c] < 3) continue;
        Owner o{};
        o.dof = E.mem[E.start[c]];
        o.sign = E.sign[E.start[c]];
        double len = 0.0, vsum = 0.0, dr[3] = {0.0, 0.0, 0.0};
        int nlen = 0;
        std::vector<Cell> cells;
        std::vector<Face> faces;
        std::vector<int> fsign;
        std::vector<std::array<double, 3>> centers;
        std::set<Key> seen_c, seen_f;
        for (int t = E.start[c]; t < E.start[c + 1]; ++t) {
            const patch::Dof &a = E.mem[t];
            const int ax = a.f, orient = E.sign[t], b = a.b;
            const BlockGeo &g = G[b];
            const std::array<int, 3> &m = T.m[b];
            const int e[3] = {a.i, a.j, a.k};
            if (owned(m, LOC_EX + ax, e)) {
                const long long q = g.idx(e[0], e[1], e[2]);
                const double L = g.dl[ax][q];
                if (std::isfinite(L) && L > 0.0) len += L, ++nlen;
                for (int u = 0; u < 3; ++u) dr[u] += orient * g.dr[3 * ax + u][q];
            }
            const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
            for (int s0 = -1; s0 <= 0; ++s0)
                for (int s1 = -1; s1 <= 0; ++s1) {
                    int p[3] = {e[0], e[1], e[2]};
                    p[t1] += s0;
                    p[t2] += s1;
                    if (!seen_c.insert(Key{b, p[0], p[1], p[2], -1}).second || !owned(m, LOC_CELL, p)) continue;
                    const long long q = g.idx(p[0], p[1], p[2]);
                    const double v = std::abs(g.jac[q]);
                    if (!(std::isfinite(v) && v > 1.e-30)) continue;
                    cells.push_back(Cell{patch::Dof{ax, b, p[0], p[1], p[2]}, orient, v});
                    centers.push_back({g.cc[0][q], g.cc[1][q], g.cc[2][q]});
                    vsum += v;
                }
            int fl[4][4];
            faces_around(ax, e, fl);
            for (int u = 0; u < 4; ++u) {
                const int fa = fl[u][0], fp[3] = {fl[u][1], fl[u][2], fl[u][3]};
                if (!seen_f.insert(Key{b, fp[0], fp[1], fp[2], fa}).second) continue;
                if (!owned(m, LOC_FX + fa, fp) || face_alias.count(Key{b, fp[0], fp[1], fp[2], fa})) continue;
                const double av = g.area[fa][g.idx(fp[0], fp[1], fp[2])];
                if (!(std::isfinite(av) && av > 1.e-30)) continue;
                faces.push_back(Face{patch::Dof{fa, b, fp[0], fp[1], fp[2]}, 0.0});
                fsign.push_back(orient * incidence(ax, e, fa, fp));
            }
        }
        // three, or more than four, physical faces: a singular edge (four is a
        // regular interior edge, two a regular physical-boundary edge)
        const int nf = (int)faces.size();
        if (!(nf == 3 || nf > 4)) continue;
        const double plen = nlen ? len / nlen : 0.0;
        double dual = centers.size() >= 3 ? polygon_area(centers, dr) : 0.0;
        if (!(dual > 0.0)) dual = plen > 0.0 ? vsum / plen : 0.0;
        o.ih = dual > 0.0 ? plen / dual : 0.0;
        for (int u = 0; u < 3; ++u) o.cv[u] = nlen > 0 ? dr[u] / nlen : 0.0;
        if (!(plen > 0.0) || !(dual > 0.0) || !std::isfinite(o.ih) || cells.empty())
            throw std::runtime_error("singular edge at block " + std::to_string(o.dof.b) + " (" + std::to_string(o.dof.i) + "," +
                                     std::to_string(o.dof.j) + "," + std::to_string(o.dof.k) + ") axis " + std::to_string(o.dof.f) +
                                     ": no length, dual area or cell");
        o.c0 = (int)R.c.size();
        for (Cell &x : cells) {
            x.w = vsum > 0.0 ? x.w / vsum : 0.0;
            R.c.push_back(x);
        }
        o.c1 = (int)R.c.size();
        o.f0 = (int)R.f.size();
        for (int u = 0; u < nf; ++u) {
            faces[u].coef = o.ih * fsign[u];
            R.f.push_back(faces[u]);
        }
        o.f1 = (int)R.f.size();
        R.e.push_back(o);
        // the corrected scalar Hodge scale on every owned alias
        for (int t = E.start[c]; t < E.start[c + 1]; ++t) {
            const patch::Dof &a = E.mem[t];
            const int e[3] = {a.i, a.j, a.k};
            if (owned(T.m[a.b], LOC_EX + a.f, e)) {
                R.alpha_dof.push_back(a);
                R.alpha_val.push_back(o.ih);
            }
        }
    }
    return R;
}

NodeStencils node_stencils(const patch::Topo &T, const Registry &R) {
    NodeStencils S;
    if (R.e.empty()) return S;
    std::vector<long long> off;
    const std::vector<long long> rep = patch::node_reps(T, off);
    auto node_rep = [&](int b, int i, int j, int k) {
        const std::array<int, 3> &m = T.m[b];
        return rep[off[b] + ((long long)i * (m[1] + 1) + j) * (m[2] + 1) + k];
    };
    // per end-node class: the cells of the singular edges ending there
    std::map<long long, std::set<Key>> cells;
    for (const Owner &o : R.e) {
        int q[3] = {o.dof.i, o.dof.j, o.dof.k};
        const long long r0 = node_rep(o.dof.b, q[0], q[1], q[2]);
        ++q[o.dof.f];
        const long long r1 = node_rep(o.dof.b, q[0], q[1], q[2]);
        if (r0 == r1) throw std::runtime_error("singular edge whose two ends are one node");
        for (const long long r : {r0, r1})
            for (int t = o.c0; t < o.c1; ++t) {
                const patch::Dof &c = R.c[t].dof;
                cells[r].insert(Key{c.b, c.i, c.j, c.k, 0});
            }
    }
    std::map<long long, int> index;
    for (const auto &kv : cells) {
        index[kv.first] = (int)S.start.size() - 1;
        for (const Key &c : kv.second) S.cells.push_back(patch::Dof{0, std::get<0>(c), std::get<1>(c), std::get<2>(c), std::get<3>(c)});
        S.start.push_back((int)S.cells.size());
    }
    for (int b = 0; b < T.nb(); ++b) {
        const std::array<int, 3> &m = T.m[b];
        for (int i = 0; i <= m[0]; ++i)
            for (int j = 0; j <= m[1]; ++j)
                for (int k = 0; k <= m[2]; ++k) {
                    const auto it = index.find(node_rep(b, i, j, k));
                    if (it != index.end()) S.alias.push_back(patch::Dof{it->second, b, i, j, k});
                }
    }
    return S;
}

}  // namespace singular
}  // namespace merc
