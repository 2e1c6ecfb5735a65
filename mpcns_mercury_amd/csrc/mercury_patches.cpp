// Patch connections: the halo and owner -> alias plans of rotated, mirrored
// and partial-face block interfaces (host only).
//
// The halo of one sync (MercuryBoundary::Sync_, Z4_Mercury/src/boundary/
// Boundary_Sync_.cpp:11-99) is simulated on DOF identities instead of values:
// every ghost carries "the start-of-halo value of DOF x, times s".  The three
// levels are replayed as the reference runs them (regions, their order, the
// read-all-then-write component copies and the write-as-you-read triplets), so
// the final identity of each ghost is what the reference's sequence of copies
// leaves there; the device then applies one flat copy list.
#include "mercury_patches.h"

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>

#include "mercury_common.h"

namespace merc {
namespace patch {
namespace {

const int DL[8][3] = {{1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};

// node index map of a transform (HALO_TOOLS::apply_transform, src/4_halo/build/HaloBuildTools.cpp:296-310)
void tmap(const int *perm, const int *sign, const int *off, const int p[3], int out[3]) {
    for (int a = 0; a < 3; ++a) out[perm[a]] = sign[a] * p[a] + off[a];
}
// TOPO::transform_node_box (src/2_topology/TopologyOps.cpp:91-117)
void tbox(const int *perm, const int *sign, const int *off, const int lo[3], const int hi[3], int olo[3], int ohi[3]) {
    const int h[3] = {hi[0] - 1, hi[1] - 1, hi[2] - 1};
    int p0[3], p1[3];
    tmap(perm, sign, off, lo, p0);
    tmap(perm, sign, off, h, p1);
    for (int a = 0; a < 3; ++a) {
        olo[a] = std::min(p0[a], p1[a]);
        ohi[a] = std::max(p0[a], p1[a]) + 1;
    }
}
// block faces a node box lies on (boundary_dir_count, src/4_halo/build/HaloBuildEdge.cpp:9-25)
int bdir_count(const int lo[3], const int hi[3], const std::array<int, 3> &m) {
    int c = 0;
    for (int a = 0; a < 3; ++a) c += (lo[a] == 0 && hi[a] == 1) + (lo[a] == m[a] && hi[a] == m[a] + 1);
    return c;
}
// HALO_TOOLS::detect_face_direction (HaloBuildTools.cpp:12-37): 0 if on no face
int face_dir_of(const int lo[3], const int hi[3], const std::array<int, 3> &m) {
    for (int a = 0; a < 3; ++a) {
        if (lo[a] == 0 && hi[a] == 1) return -(a + 1);
        if (lo[a] == m[a] && hi[a] == m[a] + 1) return a + 1;
    }
    return 0;
}
int inv_perm(const int *perm, int d) {
    for (int a = 0; a < 3; ++a)
        if (perm[a] == d) return a;
    return d;
}
int parity(const int *perm) {
    int inv = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) inv += perm[a] > perm[b];
    return (inv & 1) ? -1 : 1;
}

// a halo region: receiving and sending block, node box on the receiver, the
// out-of-range directions of its ghost box and the transform; inv: the
// transform runs sender -> receiver (face level), else receiver -> sender
struct Region {
    int recv, send, lo[3], hi[3], nd, dirs[3], perm[3], sign[3], off[3];
    bool inv;
};
struct FacePatch {
    bool inner;
    int dir, lo[3], hi[3], nb, perm[3], sign[3], off[3];
};
struct EdgePatch {
    bool inner;
    int block, lo[3], hi[3], dir1, dir2, nb, perm[3], sign[3], off[3];
};

Region make_region(int recv, int send, const int lo[3], const int hi[3], int nd, const int *dirs, const int *perm, const int *sign,
                   const int *off, bool inv) {
    Region r{};
    r.recv = recv;
    r.send = send;
    r.nd = nd;
    r.inv = inv;
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = lo[a];
        r.hi[a] = hi[a];
        r.dirs[a] = a < nd ? dirs[a] : 0;
        r.perm[a] = perm[a];
        r.sign[a] = sign[a];
        r.off[a] = off[a];
    }
    return r;
}

// Regions of the three levels, built as the reference's topology and halo
// builders do: level 1 from every patch (Halo::build_inner_1DCorner_pattern,
// src/4_halo/build/HaloBuildFace.cpp:6-57); level 2 from the edge patches of
// every pair of one block's faces on different axes whose node boxes meet
// (TOPO::build_edge_patches, src/2_topology/TopologyEdge.cpp:45-211; the
// inner face owns the edge, else the first of the pair), kept when the edge
// lies on two faces of both blocks (build_inner_2DCorner_pattern,
// HaloBuildEdge.cpp:188-265); level 3 from the vertex patches (first pair of
// edges per node, owned by the first inner edge touching it,
// TopologyEdge.cpp:213-405; build_inner_3DCorner_pattern,
// src/4_halo/build/HaloBuildVertex.cpp:190-275).
void build_regions(const Topo &T, std::vector<Region> reg[4]) {
    for (const Patch &p : T.P) {
        int lo[3], hi[3];
        tbox(p.perm, p.sign, p.off, p.lo, p.hi, lo, hi);
        const int d = face_dir_of(lo, hi, T.m[p.nb]);
        reg[1].push_back(make_region(p.nb, p.block, lo, hi, 1, &d, p.perm, p.sign, p.off, true));
    }
    std::vector<EdgePatch> ie, pe;
    for (int b = 0; b < T.nb(); ++b) {
        std::vector<FacePatch> fl;
        for (const Patch &p : T.P) {
            if (p.block != b) continue;
            FacePatch f{};
            f.inner = true;
            f.dir = p.dir;
            f.nb = p.nb;
            for (int a = 0; a < 3; ++a) {
                f.lo[a] = p.lo[a];
                f.hi[a] = p.hi[a];
                f.perm[a] = p.perm[a];
                f.sign[a] = p.sign[a];
                f.off[a] = p.off[a];
            }
            fl.push_back(f);
        }
        for (int fc = 0; fc < 6; ++fc) {
            const int c = T.bc[6 * b + fc];
            if (c != MBC_FARFIELD && c != MBC_WALL) continue;
            FacePatch f{};
            const int a = fc / 2;
            f.inner = false;
            f.dir = (fc & 1) ? a + 1 : -(a + 1);
            f.nb = -1;
            for (int t = 0; t < 3; ++t) {
                f.lo[t] = 0;
                f.hi[t] = T.m[b][t] + 1;
            }
            f.lo[a] = (fc & 1) ? T.m[b][a] : 0;
            f.hi[a] = f.lo[a] + 1;
            fl.push_back(f);
        }
        for (size_t i = 0; i < fl.size(); ++i)
            for (size_t j = i + 1; j < fl.size(); ++j) {
                const FacePatch &f1 = fl[i], &f2 = fl[j];
                if (std::abs(f1.dir) == std::abs(f2.dir)) continue;
                EdgePatch e{};
                bool empty = false;
                for (int a = 0; a < 3; ++a) {
                    e.lo[a] = std::max(f1.lo[a], f2.lo[a]);
                    e.hi[a] = std::min(f1.hi[a], f2.hi[a]);
                    empty |= e.hi[a] <= e.lo[a];
                }
                if (empty) continue;
                const bool second = f2.inner && !f1.inner;
                const FacePatch &own = second ? f2 : f1;
                e.inner = own.inner;
                e.block = b;
                e.nb = own.nb;
                for (int a = 0; a < 3; ++a) {
                    e.perm[a] = own.perm[a];
                    e.sign[a] = own.sign[a];
                    e.off[a] = own.off[a];
                }
                e.dir1 = second ? f2.dir : f1.dir;
                e.dir2 = second ? f1.dir : f2.dir;
                (e.inner ? ie : pe).push_back(e);
            }
    }
    for (const EdgePatch &e : ie) {
        int lo[3], hi[3];
        tbox(e.perm, e.sign, e.off, e.lo, e.hi, lo, hi);
        if (bdir_count(e.lo, e.hi, T.m[e.block]) != 2 || bdir_count(lo, hi, T.m[e.nb]) != 2) continue;
        const int dirs[2] = {e.dir1, e.dir2};
        reg[2].push_back(make_region(e.block, e.nb, e.lo, e.hi, 2, dirs, e.perm, e.sign, e.off, false));
    }
    for (int b = 0; b < T.nb(); ++b) {
        std::vector<const EdgePatch *> el;
        for (const EdgePatch &e : ie)
            if (e.block == b) el.push_back(&e);
        for (const EdgePatch &e : pe)
            if (e.block == b) el.push_back(&e);
        std::vector<std::array<int, 3>> used;
        for (size_t i = 0; i < el.size(); ++i)
            for (size_t j = i + 1; j < el.size(); ++j) {
                const EdgePatch *e1 = el[i], *e2 = el[j];
                if (6 - std::abs(e1->dir1) - std::abs(e1->dir2) == 6 - std::abs(e2->dir1) - std::abs(e2->dir2)) continue;
                int lo[3], hi[3];
                bool unit = true;
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::max(e1->lo[a], e2->lo[a]);
                    hi[a] = std::min(e1->hi[a], e2->hi[a]);
                    unit &= hi[a] - lo[a] == 1;
                }
                if (!unit) continue;
                const std::array<int, 3> node{lo[0], lo[1], lo[2]};
                if (std::find(used.begin(), used.end(), node) != used.end()) continue;
                used.push_back(node);
                std::vector<const EdgePatch *> cand{e1, e2};
                for (size_t k = 0; k < el.size(); ++k) {
                    if (k == i || k == j) continue;
                    bool meet = true;
                    for (int a = 0; a < 3; ++a) meet &= el[k]->lo[a] <= lo[a] && lo[a] < el[k]->hi[a];
                    if (meet) cand.push_back(el[k]);
                }
                const EdgePatch *own = cand[0];
                for (size_t k = 1; k < cand.size(); ++k)
                    if (cand[k]->inner && !own->inner) own = cand[k];
                int dba[4] = {0, 0, 0, 0};
                for (const EdgePatch *c : cand)
                    for (int d : {c->dir1, c->dir2})
                        if (!dba[std::abs(d)]) dba[std::abs(d)] = d;
                if (!own->inner) continue;
                int nlo[3], nhi[3];
                tbox(own->perm, own->sign, own->off, lo, hi, nlo, nhi);
                if (bdir_count(lo, hi, T.m[b]) != 3 || bdir_count(nlo, nhi, T.m[own->nb]) != 3) continue;
                reg[3].push_back(make_region(b, own->nb, lo, hi, 3, dba + 1, own->perm, own->sign, own->off, false));
            }
    }
}

// DOF keys: field * total + the point's index in the block boxes [-g, m+g]
// (k fastest, blocks one after another)
struct Space {
    int g;
    std::vector<std::array<int, 3>> m;
    std::vector<long long> base;
    long long total = 0;
    explicit Space(const Topo &T) : g(T.g), m(T.m) {
        for (const auto &d : m) {
            base.push_back(total);
            total += (long long)(d[0] + 1 + 2 * g) * (d[1] + 1 + 2 * g) * (d[2] + 1 + 2 * g);
        }
    }
    long long key(int f, int b, const int p[3]) const {
        const long long N1 = m[b][1] + 1 + 2 * g, N2 = m[b][2] + 1 + 2 * g;
        return f * total + base[b] + ((long long)(p[0] + g) * N1 + (p[1] + g)) * N2 + (p[2] + g);
    }
    Dof dof(long long k) const {
        Dof d{};
        d.f = (int)(k / total);
        long long r = k % total;
        int b = 0;
        while (b + 1 < (int)base.size() && base[b + 1] <= r) ++b;
        r -= base[b];
        const long long N1 = m[b][1] + 1 + 2 * g, N2 = m[b][2] + 1 + 2 * g;
        d.b = b;
        d.i = (int)(r / (N1 * N2)) - g;
        d.j = (int)((r / N2) % N1) - g;
        d.k = (int)(r % N2) - g;
        return d;
    }
};

void owned_hi(const std::array<int, 3> &m, int L, int hi[3]) {
    for (int a = 0; a < 3; ++a) hi[a] = m[a] + 1 - DL[L][a];
}
bool in_alloc(const std::array<int, 3> &m, int g, int L, const int s[3]) {
    for (int a = 0; a < 3; ++a)
        if (s[a] < -g || s[a] >= m[a] + 1 - DL[L][a] + g) return false;
    return true;
}
// ghost box of a region at location L (LAYOUT::node_box_to_dof_box and
// apply_ghost_strip, src/0_basic/LayoutTraits.cpp:104-118, 296-303)
void region_box(const Region &r, int L, int g, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = r.lo[a];
        hi[a] = r.hi[a] - DL[L][a];
    }
    for (int t = 0; t < r.nd; ++t) {
        const int d = r.dirs[t], ax = std::abs(d) - 1;
        if (d < 0) {
            lo[ax] = -g;
            hi[ax] = 0;
        } else {
            lo[ax] = hi[ax];
            hi[ax] = lo[ax] + g;
        }
    }
}
// source index of receiver point p at location Lr: rule 0 maps DOF indices as
// node indices (HaloExchangeInner.cpp:35-52, HaloTriplet.cpp:53-84); rule 1
// maps positions (a DOF between nodes lands one index lower along a reversed axis)
void region_src(int rule, const Region &r, int Lr, const int p[3], int s[3]) {
    if (r.inv) {
        for (int a = 0; a < 3; ++a) {
            s[a] = r.sign[a] * (p[r.perm[a]] - r.off[a]);
            if (rule && r.sign[a] < 0) s[a] -= DL[Lr][r.perm[a]];
        }
    } else {
        for (int a = 0; a < 3; ++a) {
            s[r.perm[a]] = r.sign[a] * p[a] + r.off[a];
            if (rule && r.sign[a] < 0) s[r.perm[a]] -= DL[Lr][a];
        }
    }
}

// value identity of a DOF during the simulated halo: (start-of-halo DOF, sign)
using Prov = std::unordered_map<long long, std::pair<long long, int>>;
std::pair<long long, int> read(const Prov &pv, long long k) {
    auto it = pv.find(k);
    return it == pv.end() ? std::make_pair(k, 1) : it->second;
}

// Rule 1 only: the owned DOF at the position of a ghost no region wrote,
// through a chain of patch maps (across the face of the first out-of-range
// axis, through the first patch of that face holding the DOF's span along the
// in-range tangential axes and touching the face's edge along the others), at
// most 8 hops.  Triplets carry the axis and the factor of every hop.
bool chain_source(const Topo &T, int Lb, bool trip, int b, int d, const int p0[3], int &ob, int q[3], int &od, int &ofac) {
    int p[3] = {p0[0], p0[1], p0[2]}, fac = 1;
    for (int hop = 0; hop < 8; ++hop) {
        const int L = trip ? Lb + d : Lb;
        int hi[3], a = -1;
        owned_hi(T.m[b], L, hi);
        for (int t = 0; t < 3 && a < 0; ++t)
            if (p[t] < 0 || p[t] >= hi[t]) a = t;
        if (a < 0) {
            if (hop == 0) return false;
            ob = b, od = d, ofac = fac;
            for (int t = 0; t < 3; ++t) q[t] = p[t];
            return true;
        }
        const int dir = p[a] < 0 ? -(a + 1) : a + 1;
        const Patch *P = nullptr;
        for (const Patch &c : T.P) {
            if (c.block != b || c.dir != dir) continue;
            bool ok = true;
            for (int t = 0; t < 3 && ok; ++t) {
                if (t == a) continue;
                if (p[t] >= 0 && p[t] < hi[t]) ok = c.lo[t] <= p[t] && p[t] + DL[L][t] <= c.hi[t] - 1;
                else if (p[t] < 0) ok = c.lo[t] == 0;
                else ok = c.hi[t] - 1 == T.m[b][t];
            }
            if (ok) {
                P = &c;
                break;
            }
        }
        if (!P) return false;
        int np[3];
        for (int t = 0; t < 3; ++t) np[P->perm[t]] = P->sign[t] * p[t] + P->off[t] - (P->sign[t] < 0 ? DL[L][t] : 0);
        if (trip) {
            fac *= Lb == LOC_FX ? parity(P->perm) * P->sign[0] * P->sign[1] * P->sign[2] * P->sign[d] : P->sign[d];
            d = P->perm[d];
        }
        b = P->nb;
        for (int t = 0; t < 3; ++t) p[t] = np[t];
    }
    return false;
}

void complete_level(const Topo &T, const Space &S, Prov &pv, int Lb, bool trip, int stage) {
    for (int d = 0; d < (trip ? 3 : 1); ++d) {
        const int L = trip ? Lb + d : Lb;
        for (int b = 0; b < T.nb(); ++b) {
            int ohi[3], p[3];
            owned_hi(T.m[b], L, ohi);
            for (p[0] = -T.g; p[0] < ohi[0] + T.g; ++p[0])
                for (p[1] = -T.g; p[1] < ohi[1] + T.g; ++p[1])
                    for (p[2] = -T.g; p[2] < ohi[2] + T.g; ++p[2]) {
                        int nout = 0;
                        for (int t = 0; t < 3; ++t) nout += p[t] < 0 || p[t] >= ohi[t];
                        if (nout != stage) continue;
                        const long long k = S.key(d, b, p);
                        if (pv.count(k)) continue;
                        int sb, q[3], sd, fac;
                        if (!chain_source(T, Lb, trip, b, d, p, sb, q, sd, fac)) continue;
                        pv[k] = {S.key(sd, sb, q), fac};
                    }
        }
    }
}

// node classes (representative = smallest (block, i, j, k)) through the patches
struct NodeClasses {
    std::vector<long long> base;
    std::vector<long long> par, cnt;
    const Topo &T;
    explicit NodeClasses(const Topo &T_) : T(T_) {
        long long n = 0;
        for (const auto &d : T.m) {
            base.push_back(n);
            n += (long long)(d[0] + 1) * (d[1] + 1) * (d[2] + 1);
        }
        par.resize(n);
        cnt.assign(n, 0);
        for (long long x = 0; x < n; ++x) par[x] = x;
        for (const Patch &p : T.P) {
            int q[3], s[3];
            for (q[0] = p.lo[0]; q[0] < p.hi[0]; ++q[0])
                for (q[1] = p.lo[1]; q[1] < p.hi[1]; ++q[1])
                    for (q[2] = p.lo[2]; q[2] < p.hi[2]; ++q[2]) {
                        tmap(p.perm, p.sign, p.off, q, s);
                        const long long rx = find(node(p.block, q)), ry = find(node(p.nb, s));
                        if (rx != ry) par[std::max(rx, ry)] = std::min(rx, ry);
                    }
        }
        for (long long x = 0; x < n; ++x) ++cnt[find(x)];
    }
    long long node(int b, const int q[3]) const {
        return base[b] + ((long long)q[0] * (T.m[b][1] + 1) + q[1]) * (T.m[b][2] + 1) + q[2];
    }
    long long find(long long x) {
        while (par[x] != x) x = par[x] = par[par[x]];
        return x;
    }
};

struct Ent {
    long long key[4];
    int b, p[3], ax, sign;
    long long st[4][3];  // face stencil: (edge key lo, hi, coefficient), sorted
};
bool ent_less(const Ent &x, const Ent &y) {
    for (int t = 0; t < 4; ++t)
        if (x.key[t] != y.key[t]) return x.key[t] < y.key[t];
    if (x.b != y.b) return x.b < y.b;
    for (int t = 0; t < 3; ++t)
        if (x.p[t] != y.p[t]) return x.p[t] < y.p[t];
    return x.ax < y.ax;
}

// Classes of shared edges (dim 0) or faces (dim 1) (TOPO::detail::
// build_equivalence, src/2_topology/Topology.cpp:280-323): an edge's key is its
// sorted endpoint representatives, its sign +1 in that order (make_edge_key,
// TopologyQuery.cpp:104-126); a face's key is its sorted corner
// representatives, its sign +1 or -1 as its boundary-edge stencil equals or
// opposes the owner's (orient_face_members_to_owner, TopologyFaceEquiv.cpp:
// 298-351; boundary_of_face, include/2_topology/LocalIncidence.h:44-79).
// Members sorted by (block, i, j, k, axis); owner = first.
void classes(const Topo &T, int dim, std::vector<int> &start, std::vector<Dof> &mem, std::vector<int> &sign) {
    NodeClasses N(T);
    std::vector<Ent> e;
    for (int b = 0; b < T.nb(); ++b)
        for (int ax = 0; ax < 3; ++ax) {
            const int L = (dim ? LOC_FX : LOC_EX) + ax;
            int hi[3], p[3];
            owned_hi(T.m[b], L, hi);
            for (p[0] = 0; p[0] < hi[0]; ++p[0])
                for (p[1] = 0; p[1] < hi[1]; ++p[1])
                    for (p[2] = 0; p[2] < hi[2]; ++p[2]) {
                        Ent x{};
                        x.b = b, x.ax = ax, x.sign = 1;
                        for (int t = 0; t < 3; ++t) x.p[t] = p[t];
                        if (!dim) {
                            const long long r0 = N.find(N.node(b, p));
                            if (N.cnt[r0] <= 1) continue;
                            int q[3] = {p[0], p[1], p[2]};
                            ++q[ax];
                            const long long r1 = N.find(N.node(b, q));
                            if (N.cnt[r1] <= 1 || r0 == r1) continue;
                            x.key[0] = std::min(r0, r1);
                            x.key[1] = std::max(r0, r1);
                            x.sign = r0 < r1 ? 1 : -1;
                        } else {
                            const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
                            long long c[4];
                            bool ok = true;
                            for (int n = 0; n < 4 && ok; ++n) {
                                int q[3] = {p[0], p[1], p[2]};
                                q[t1] += n & 1;
                                q[t2] += n >> 1;
                                c[n] = N.find(N.node(b, q));
                                ok = N.cnt[c[n]] > 1;
                            }
                            if (!ok) continue;
                            std::sort(c, c + 4);
                            for (int n = 0; n < 4; ++n) x.key[n] = c[n];
                            // boundary edges (axis, start offset, incidence): the loop around +axis
                            const int u1 = ax == 1 ? 2 : t1, u2 = ax == 1 ? 0 : t2;
                            const int ea[4] = {u1, u2, u1, u2}, inc[4] = {1, 1, -1, -1};
                            std::vector<std::array<long long, 3>> st;
                            for (int n = 0; n < 4; ++n) {
                                int s0[3] = {p[0], p[1], p[2]};
                                if (n == 1) ++s0[u1];
                                if (n == 2) ++s0[u2];
                                int s1[3] = {s0[0], s0[1], s0[2]};
                                ++s1[ea[n]];
                                const long long a0 = N.find(N.node(b, s0)), a1 = N.find(N.node(b, s1));
                                const long long lo = std::min(a0, a1), hi2 = std::max(a0, a1), cf = inc[n] * (a0 < a1 ? 1 : -1);
                                bool hit = false;
                                for (auto &w : st)
                                    if (w[0] == lo && w[1] == hi2) w[2] += cf, hit = true;
                                if (!hit) st.push_back({lo, hi2, cf});
                            }
                            st.erase(std::remove_if(st.begin(), st.end(), [](const std::array<long long, 3> &w) { return w[2] == 0; }), st.end());
                            std::sort(st.begin(), st.end());
                            for (size_t w = 0; w < st.size(); ++w)
                                for (int z = 0; z < 3; ++z) x.st[w][z] = st[w][z];
                        }
                        e.push_back(x);
                    }
        }
    std::sort(e.begin(), e.end(), ent_less);
    start.assign(1, 0);
    mem.clear();
    sign.clear();
    for (size_t u = 0; u < e.size();) {
        size_t v = u + 1;
        while (v < e.size() && std::equal(e[v].key, e[v].key + 4, e[u].key)) ++v;
        if (v - u >= 2) {
            for (size_t w = u; w < v; ++w) {
                int sg = e[w].sign;
                if (dim) {
                    bool same = true, opp = true;
                    for (int z = 0; z < 4; ++z) {
                        same &= e[w].st[z][0] == e[u].st[z][0] && e[w].st[z][1] == e[u].st[z][1] && e[w].st[z][2] == e[u].st[z][2];
                        opp &= e[w].st[z][0] == e[u].st[z][0] && e[w].st[z][1] == e[u].st[z][1] && e[w].st[z][2] == -e[u].st[z][2];
                    }
                    if (!same && !opp) throw std::runtime_error("patch topology: a shared face's boundary does not match its owner's");
                    sg = same ? 1 : -1;
                }
                mem.push_back(Dof{e[w].ax, e[w].b, e[w].p[0], e[w].p[1], e[w].p[2]});
                sign.push_back(sg);
            }
            start.push_back((int)mem.size());
        }
        u = v;
    }
}

}  // namespace

Topo make_topo(int nb, const int *dims, int g, const int *bc, int npatch, const int *patches, int rule) {
    if (nb <= 0 || !dims || !bc || npatch < 0 || (npatch > 0 && !patches)) throw std::runtime_error("patch topology: invalid argument");
    Topo T;
    T.g = g;
    T.rule = rule;
    for (int b = 0; b < nb; ++b) T.m.push_back({dims[3 * b], dims[3 * b + 1], dims[3 * b + 2]});
    T.bc.assign(bc, bc + 6 * nb);
    for (int p = 0; p < npatch; ++p) {
        const int *r = patches + kPatchInts * (long long)p;
        Patch q{};
        q.block = r[0];
        q.dir = r[1];
        for (int a = 0; a < 3; ++a) {
            q.lo[a] = r[2 + a];
            q.hi[a] = r[5 + a];
            q.perm[a] = r[9 + a];
            q.sign[a] = r[12 + a];
            q.off[a] = r[15 + a];
        }
        q.nb = r[8];
        T.P.push_back(q);
    }
    T.validate();
    return T;
}

void Topo::validate() const {
    auto fail = [](const std::string &s) { throw std::runtime_error("patch topology: " + s); };
    if (rule != 0 && rule != 1) fail("rule must be 0 (node-index map) or 1 (positions)");
    if (g < 1) fail("the ghost depth must be at least 1");
    for (int b = 0; b < nb(); ++b) {
        for (int a = 0; a < 3; ++a)
            if (m[b][a] < 2) fail("every block needs at least 2 cells per axis");
        for (int f = 0; f < 6; ++f) {
            const int c = bc[6 * b + f];
            if (c != MBC_FARFIELD && c != MBC_WALL && c != MBC_INTERFACE)
                fail("face codes must be far field, wall or interface (periodicity is a patch from a block to itself)");
        }
    }
    for (size_t n = 0; n < P.size(); ++n) {
        const Patch &p = P[n];
        const std::string w = "patch " + std::to_string(n) + ": ";
        if (p.block < 0 || p.block >= nb() || p.nb < 0 || p.nb >= nb()) fail(w + "block out of range");
        if (p.dir == 0 || std::abs(p.dir) > 3) fail(w + "direction must be +-1, +-2 or +-3");
        int seen[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) {
            if (p.perm[a] < 0 || p.perm[a] > 2 || seen[p.perm[a]]++) fail(w + "perm is not a permutation");
            if (p.sign[a] != 1 && p.sign[a] != -1) fail(w + "signs must be +1 or -1");
            if (p.lo[a] < 0 || p.hi[a] <= p.lo[a] || p.hi[a] > m[p.block][a] + 1) fail(w + "node box outside the block");
        }
        if (face_dir_of(p.lo, p.hi, m[p.block]) != p.dir) fail(w + "node box is not on the face of its direction");
        if (bc[6 * p.block + 2 * (std::abs(p.dir) - 1) + (p.dir > 0)] != MBC_INTERFACE) fail(w + "its face is not an interface");
        int lo[3], hi[3];
        tbox(p.perm, p.sign, p.off, p.lo, p.hi, lo, hi);
        for (int a = 0; a < 3; ++a)
            if (lo[a] < 0 || hi[a] > m[p.nb][a] + 1) fail(w + "maps outside the neighbour block");
        const int dn = face_dir_of(lo, hi, m[p.nb]);
        if (dn == 0) fail(w + "maps onto no face of the neighbour");
        if (bc[6 * p.nb + 2 * (std::abs(dn) - 1) + (dn > 0)] != MBC_INTERFACE) fail(w + "the neighbour's face is not an interface");
    }
    for (int b = 0; b < nb(); ++b)
        for (int f = 0; f < 6; ++f) {
            if (bc[6 * b + f] != MBC_INTERFACE) continue;
            const int a = f / 2, d = (f & 1) ? a + 1 : -(a + 1), t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
            for (int u = 0; u <= m[b][t1]; ++u)
                for (int v = 0; v <= m[b][t2]; ++v) {
                    int q[3];
                    q[a] = (f & 1) ? m[b][a] : 0;
                    q[t1] = u;
                    q[t2] = v;
                    bool cov = false;
                    for (const Patch &p : P) {
                        if (p.block != b || p.dir != d) continue;
                        bool in = true;
                        for (int c = 0; c < 3; ++c) in &= p.lo[c] <= q[c] && q[c] < p.hi[c];
                        if (in) {
                            cov = true;
                            break;
                        }
                    }
                    if (!cov)
                        fail("block " + std::to_string(b) + " face " + std::to_string(f) + ": node (" + std::to_string(q[0]) + "," +
                             std::to_string(q[1]) + "," + std::to_string(q[2]) + ") lies in no patch");
                }
        }
}

std::vector<Entry> halo_entries(const Topo &T, int kind, int level) {
    if (kind < KIND_COMPONENT || kind > KIND_EDGE || level < 1 || level > 3) throw std::runtime_error("patch plan: invalid kind or level");
    std::vector<Region> reg[4];
    build_regions(T, reg);
    const Space S(T);
    const bool trip = kind != KIND_COMPONENT, face = kind == KIND_FACE;
    const int Lb = kind == KIND_COMPONENT ? LOC_CELL : (face ? LOC_FX : LOC_EX);
    Prov pv;
    for (int lvl = 1; lvl <= level; ++lvl) {
        if (!trip) {
            // Halo::exchange_inner / _edge / _vertex (src/4_halo/component_copy/
            // HaloExchangeInner.cpp:184-247): all regions read, then all write
            std::vector<std::pair<long long, std::pair<long long, int>>> pend;
            for (const Region &r : reg[lvl]) {
                int lo[3], hi[3], p[3], s[3];
                region_box(r, Lb, T.g, lo, hi);
                for (p[0] = lo[0]; p[0] < hi[0]; ++p[0])
                    for (p[1] = lo[1]; p[1] < hi[1]; ++p[1])
                        for (p[2] = lo[2]; p[2] < hi[2]; ++p[2]) {
                            region_src(T.rule, r, Lb, p, s);
                            if (!in_alloc(T.m[r.send], T.g, Lb, s)) continue;
                            pend.push_back({S.key(0, r.recv, p), read(pv, S.key(0, r.send, s))});
                        }
            }
            for (const auto &x : pend) pv[x.first] = x.second;
        } else {
            // Halo::exchange_inner_{face,edge,vertex}_{face_2form,edge_1form}_triplet_
            // (src/4_halo/forms/HaloTriplet.cpp:355-489, 641-829): written as read;
            // source axis and factor as the oracle's patch_exchange_triplet states
            for (int d = 0; d < 3; ++d)
                for (const Region &r : reg[lvl]) {
                    const int det = r.sign[0] * r.sign[1] * r.sign[2], par = T.rule ? parity(r.perm) : 1;
                    int sa, fac;
                    if (lvl == 1) {
                        sa = inv_perm(r.perm, d);
                        fac = face ? par * det * r.sign[sa] : r.sign[sa];
                    } else if (!face) {
                        sa = r.perm[d];
                        fac = r.sign[d];
                    } else if (T.rule) {
                        sa = r.perm[d];
                        fac = par * det * r.sign[d];
                    } else {
                        sa = inv_perm(r.perm, d);
                        fac = det * r.sign[sa];
                    }
                    int lo[3], hi[3], p[3], s[3];
                    region_box(r, Lb + d, T.g, lo, hi);
                    for (p[0] = lo[0]; p[0] < hi[0]; ++p[0])
                        for (p[1] = lo[1]; p[1] < hi[1]; ++p[1])
                            for (p[2] = lo[2]; p[2] < hi[2]; ++p[2]) {
                                region_src(T.rule, r, Lb + d, p, s);
                                if (!in_alloc(T.m[r.send], T.g, Lb + sa, s)) continue;
                                auto v = read(pv, S.key(sa, r.send, s));
                                v.second *= fac;
                                pv[S.key(d, r.recv, p)] = v;
                            }
                }
        }
        if (T.rule) complete_level(T, S, pv, Lb, trip, lvl);
    }
    std::vector<std::pair<long long, std::pair<long long, int>>> out(pv.begin(), pv.end());
    std::sort(out.begin(), out.end());
    std::vector<Entry> ent;
    for (const auto &x : out) {
        if (x.second.first == x.first && x.second.second == 1) continue;
        ent.push_back(Entry{S.dof(x.first), S.dof(x.second.first), x.second.second});
    }
    return ent;
}

std::vector<Entry> owner_entries(const Topo &T, int kind) {
    if (kind != KIND_FACE && kind != KIND_EDGE) return {};
    std::vector<int> start, sign;
    std::vector<Dof> mem;
    classes(T, kind == KIND_FACE ? 1 : 0, start, mem, sign);
    // Halo::sync_owner_alias_group (HaloSyncFramework.cpp:501-560, 695-712):
    // alias = sign(owner) * sign(alias) * owner
    std::vector<Entry> ent;
    for (size_t c = 0; c + 1 < start.size(); ++c)
        for (int t = start[c] + 1; t < start[c + 1]; ++t) ent.push_back(Entry{mem[t], mem[start[c]], sign[start[c]] * sign[t]});
    return ent;
}

EdgeClasses edge_classes(const Topo &T) {
    EdgeClasses E;
    classes(T, 0, E.start, E.mem, E.sign);
    return E;
}

std::vector<long long> node_reps(const Topo &T, std::vector<long long> &off) {
    NodeClasses N(T);
    off = N.base;
    off.push_back((long long)N.par.size());
    std::vector<long long> r(N.par.size());
    for (long long x = 0; x < (long long)r.size(); ++x) r[x] = N.find(x);
    return r;
}

EdgeClasses face_classes(const Topo &T) {
    EdgeClasses E;
    classes(T, 1, E.start, E.mem, E.sign);
    return E;
}

}  // namespace patch
}  // namespace merc
