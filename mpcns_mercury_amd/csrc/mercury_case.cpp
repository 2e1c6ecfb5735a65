// Reader of a CASE geometry directory (see mercury_case.h).
//
// File formats (paths relative to the geometry directory, rank r spelt as in
// Grid::ReadGeometry, src/1_grid/1_MPCNS_Grid.cpp:10-40: "%4d" padded with spaces):
//   grid r.grd                      Read_Grid_Binary (:125-230) / Read_Grid (:42-123)
//   boundary_condition/boundary r.txt        Read_Phy_Boundary (:232-294)
//   boundary_condition/boundary_name.fvbnd   Read_Phy_Name (:296-336)
//   boundary_condition/inner r.txt           Read_Inner_Boundary (:401-517)
//   boundary_condition/parallel r.txt        read_parallel_boundaries_for_rank
//                                            (src/2_topology/TopologyBuilder.cpp:239-318)
// Inner and parallel ranges are 1-based node indices whose sign marks the
// in-face axis roles; Pre_process (include/1_grid/Grid_Boundary.h:97-180,
// 247-330) makes them 0-based keeping the sign, and the patch transform follows
// make_index_transform_from_boundary_arrays (src/2_topology/TopologyOps.cpp:118-219).
#include "mercury_case.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace merc {
namespace casefile {
namespace {

void fail(const std::string &m) { throw std::runtime_error("case reader: " + m); }

std::string rank_suffix(int r) {
    const std::string s = std::to_string(r);
    return (s.size() < 4) ? std::string(4 - s.size(), ' ') + s : s;
}

std::string bc_path(const std::string &dir, const std::string &stem, int r) {
    return dir + "/boundary_condition/" + stem + rank_suffix(r) + ".txt";
}

bool exists(const std::string &p) {
    std::ifstream f(p, std::ios::binary);
    return (bool)f;
}

// Pre_process: absolute value minus one, sign kept
void minus1(int *p) {
    for (int i = 0; i < 3; ++i) p[i] = (p[i] > 0) ? p[i] - 1 : p[i] + 1;
}

// the normal axis of a range: direction +-(axis+1), 0 if none
int range_direction(const int *sub, const int *sup) {
    int dir = 0;
    for (int i = 0; i < 3; ++i)
        if (sub[i] == sup[i]) dir = (sub[i] == 0) ? -(i + 1) : (i + 1);
    return dir;
}

// Inner_Boundary::Pre_process roles: 1 normal, 2 negative-marked, 0 the other
void inner_roles(const int *sub, const int *sup, int *T) {
    for (int i = 0; i < 3; ++i) {
        if (sub[i] == sup[i]) T[i] = 1;
        else if (sub[i] < 0 || sup[i] < 0) T[i] = 2;
        else T[i] = 0;
    }
}

// Parallel_Boundary::Pre_process roles (three-dimensional decision tree)
void parallel_roles(const int *sub, const int *sup, int *T) {
    if (sup[2] == sub[2]) {
        T[2] = 1;
        if (sup[1] < 0 || sub[1] < 0) { T[1] = 2; T[0] = 0; }
        else { T[1] = 0; T[0] = 2; }
    } else if (sup[2] < 0 || sub[2] < 0) {
        T[2] = 2;
        if (sup[1] == sub[1]) { T[1] = 1; T[0] = 0; }
        else { T[1] = 0; T[0] = 1; }
    } else {
        T[2] = 0;
        if (sup[1] < 0 || sub[1] < 0) { T[1] = 2; T[0] = 1; }
        else { T[1] = 1; T[0] = 2; }
    }
}

void node_box(const int *sub, const int *sup, int *lo, int *hi) {
    for (int d = 0; d < 3; ++d) {
        const int s = std::abs(sub[d]), t = std::abs(sup[d]);
        lo[d] = std::min(s, t);
        hi[d] = std::max(s, t) + 1;
    }
}

void make_transform(Interface &p, const int *T, const int *nbT, const std::string &ctx) {
    int inv[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
        if (nbT[d] < 0 || nbT[d] > 2 || T[d] < 0 || T[d] > 2) fail(ctx + ": invalid axis roles");
        inv[nbT[d]] = d;
    }
    for (int d = 0; d < 3; ++d) p.perm[d] = inv[T[d]];
    for (int d = 0; d < 3; ++d) {
        const int ms = std::abs(p.sub[d]), mp = std::abs(p.sup[d]);
        const int b = p.perm[d];
        const int ts = std::abs(p.nb_sub[b]), tp = std::abs(p.nb_sup[b]);
        if (ms == mp) {
            if (ts != tp) fail(ctx + ": normal direction mismatch");
            p.sign[d] = (p.direction * p.nb_direction > 0) ? -1 : 1;
        } else {
            if (std::abs(mp - ms) != std::abs(tp - ts)) fail(ctx + ": interval length mismatch");
            p.sign[d] = ((mp - ms) * (tp - ts) > 0) ? 1 : -1;
        }
        p.offset[d] = -p.sign[d] * ms + ts;
    }
}

void read_grid(const std::string &path, bool binary, int rank, std::vector<Block> &out) {
    std::ifstream in(path, binary ? std::ios::binary : std::ios::in);
    if (!in) fail("cannot open " + path);
    int32_t nb = 0;
    std::vector<Block> blk;
    if (binary) {
        in.read(reinterpret_cast<char *>(&nb), 4);
        if (!in || nb <= 0) fail("bad block count in " + path);
        blk.resize((size_t)nb);
        for (auto &b : blk) {
            int32_t len = 0;
            in.read(reinterpret_cast<char *>(&len), 4);
            if (!in || len < 0) fail("bad block name in " + path);
            std::string s((size_t)len, '\0');
            if (len) in.read(&s[0], len);
            b.name = s.substr(0, s.find('\0'));
        }
        for (auto &b : blk) in.read(reinterpret_cast<char *>(b.nn), 12);
    } else {
        in >> nb;
        if (!in || nb <= 0) fail("bad block count in " + path);
        blk.resize((size_t)nb);
        for (auto &b : blk) in >> b.name;
        for (auto &b : blk) in >> b.nn[0] >> b.nn[1] >> b.nn[2];
    }
    if (!in) fail("truncated header in " + path);
    for (int ib = 0; ib < nb; ++ib) {
        Block &b = blk[ib];
        b.rank = rank;
        b.local = ib;
        if (b.nn[0] < 2 || b.nn[1] < 2 || b.nn[2] < 2) fail("block " + std::to_string(ib) + " of " + path + " is not three-dimensional");
        const size_t n = (size_t)b.nn[0] * b.nn[1] * b.nn[2];
        for (int c = 0; c < 3; ++c) {
            b.xyz[c].resize(n);
            if (binary) in.read(reinterpret_cast<char *>(b.xyz[c].data()), (std::streamsize)(n * sizeof(double)));
            else
                for (size_t t = 0; t < n; ++t) in >> b.xyz[c][t];
            if (!in) fail("truncated coordinates in " + path);
        }
    }
    for (auto &b : blk) out.push_back(std::move(b));
}

struct RawParallel {
    int block;  // local
    int sub[3], sup[3];
    int srid, sflag, rflag;
    int direction;
    int T[3];
};

std::vector<RawParallel> read_parallel(const std::string &path, int nblock) {
    std::vector<RawParallel> out;
    std::ifstream in(path);
    if (!in) return out;  // a rank without parallel connections may have no file
    int nb = 0, nface = 0;
    in >> nb >> nface;
    if (!in) return out;
    if (nb != nblock) fail(path + ": block count differs from the grid's");
    std::vector<int> per((size_t)nb);
    for (auto &x : per) in >> x;
    std::string line;
    std::getline(in, line);
    std::getline(in, line);
    for (int b = 0; b < nb; ++b)
        for (int f = 0; f < per[b]; ++f) {
            RawParallel r;
            r.block = b;
            in >> r.sub[0] >> r.sup[0] >> r.sub[1] >> r.sup[1] >> r.sub[2] >> r.sup[2] >> r.srid >> r.sflag >> r.rflag;
            std::string name;
            in >> name;
            if (!in) fail("truncated " + path);
            minus1(r.sub);
            minus1(r.sup);
            r.direction = range_direction(r.sub, r.sup);
            parallel_roles(r.sub, r.sup, r.T);
            out.push_back(r);
        }
    return out;
}

}  // namespace

Geometry Geometry::read(const std::string &dir, int nrank, bool binary) {
    Geometry G;
    if (nrank <= 0) {
        nrank = 0;
        while (exists(dir + "/grid" + rank_suffix(nrank) + ".grd")) ++nrank;
        if (nrank == 0) fail("no grid file in " + dir);
    }
    std::vector<int> first(nrank + 1, 0);
    for (int r = 0; r < nrank; ++r) {
        read_grid(dir + "/grid" + rank_suffix(r) + ".grd", binary, r, G.blocks);
        first[r + 1] = (int)G.blocks.size();
    }
    // boundary names: id 0 = line 2, id 1 unused, then 2, 3, ... up to BOUNDARIES
    {
        const std::string p = dir + "/boundary_condition/boundary_name.fvbnd";
        std::ifstream in(p);
        if (!in) fail("cannot open " + p);
        std::string line, tok;
        std::getline(in, line);
        std::getline(in, line);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        G.bc_names = {line, ""};
        while (in >> tok && tok != "BOUNDARIES") G.bc_names.push_back(tok);
    }
    for (int r = 0; r < nrank; ++r) {
        const int nb = first[r + 1] - first[r];
        // physical patches: closed 0-based ranges
        {
            const std::string p = bc_path(dir, "boundary", r);
            std::ifstream in(p);
            if (!in) fail("cannot open " + p);
            for (int b = 0; b < nb; ++b) {
                int n = 0;
                in >> n;
                for (int f = 0; f < n; ++f) {
                    int a[6], id;
                    in >> a[0] >> a[1] >> a[2] >> a[3] >> a[4] >> a[5] >> id;
                    if (!in) fail("truncated " + p);
                    Physical ph;
                    ph.block = first[r] + b;
                    for (int d = 0; d < 3; ++d) {
                        ph.sub[d] = std::min(a[2 * d], a[2 * d + 1]);
                        ph.sup[d] = std::max(a[2 * d], a[2 * d + 1]);
                    }
                    ph.direction = range_direction(ph.sub, ph.sup);
                    ph.bc_id = id;
                    ph.bc_name = (id >= 0 && id < (int)G.bc_names.size()) ? G.bc_names[id] : "";
                    G.physical.push_back(ph);
                }
            }
        }
        // inner patches of this rank
        {
            const std::string p = bc_path(dir, "inner", r);
            std::ifstream in(p);
            if (!in) fail("cannot open " + p);
            std::string line;
            std::getline(in, line);
            std::vector<int> per((size_t)nb);
            for (auto &x : per) in >> x;
            std::getline(in, line);
            std::getline(in, line);
            for (int b = 0; b < nb; ++b)
                for (int f = 0; f < per[b]; ++f) {
                    Interface q;
                    int src, tar;
                    std::string n1, n2;
                    in >> q.sub[0] >> q.sup[0] >> q.sub[1] >> q.sup[1] >> q.sub[2] >> q.sup[2] >> src >> n1;
                    in >> q.nb_sub[0] >> q.nb_sup[0] >> q.nb_sub[1] >> q.nb_sup[1] >> q.nb_sub[2] >> q.nb_sup[2] >> tar >> n2;
                    if (!in) fail("truncated " + p);
                    q.kind = 0;
                    q.periodic = src < 0;
                    q.block = first[r] + std::abs(src) - 1;
                    q.nb_block = first[r] + std::abs(tar) - 1;
                    if (std::abs(src) - 1 != b || std::abs(tar) < 1 || std::abs(tar) > nb)
                        fail(p + ": block numbers out of range");
                    minus1(q.sub);
                    minus1(q.sup);
                    minus1(q.nb_sub);
                    minus1(q.nb_sup);
                    q.direction = range_direction(q.sub, q.sup);
                    q.nb_direction = range_direction(q.nb_sub, q.nb_sup);
                    int T[3], nbT[3];
                    inner_roles(q.sub, q.sup, T);
                    inner_roles(q.nb_sub, q.nb_sup, nbT);
                    make_transform(q, T, nbT, p);
                    node_box(q.sub, q.sup, q.lo, q.hi);
                    node_box(q.nb_sub, q.nb_sup, q.nb_lo, q.nb_hi);
                    G.interfaces.push_back(q);
                }
        }
    }
    // parallel patches: matched by flags in the target rank's file
    std::vector<std::vector<RawParallel>> par((size_t)nrank);
    for (int r = 0; r < nrank; ++r) par[r] = read_parallel(bc_path(dir, "parallel", r), first[r + 1] - first[r]);
    for (int r = 0; r < nrank; ++r)
        for (const RawParallel &a : par[r]) {
            if (a.srid < 0 || a.srid >= nrank) fail("parallel patch of rank " + std::to_string(r) + " targets a missing rank");
            const RawParallel *m = nullptr;
            for (const RawParallel &c : par[a.srid])
                if (c.rflag == a.sflag && c.sflag == a.rflag) {
                    m = &c;
                    break;
                }
            if (!m) fail("no matching parallel patch for flags " + std::to_string(a.sflag) + "/" + std::to_string(a.rflag));
            Interface q;
            q.kind = 1;
            q.block = first[r] + a.block;
            q.nb_block = first[a.srid] + m->block;
            for (int d = 0; d < 3; ++d) {
                q.sub[d] = a.sub[d];
                q.sup[d] = a.sup[d];
                q.nb_sub[d] = m->sub[d];
                q.nb_sup[d] = m->sup[d];
            }
            q.direction = a.direction;
            q.nb_direction = m->direction;
            make_transform(q, a.T, m->T, "parallel patch of rank " + std::to_string(r));
            node_box(q.sub, q.sup, q.lo, q.hi);
            node_box(q.nb_sub, q.nb_sup, q.nb_lo, q.nb_hi);
            G.interfaces.push_back(q);
        }
    return G;
}

void Geometry::patch_topology(int *bc, std::vector<int> &patches) const {
    const int nb = (int)blocks.size();
    patches.clear();
    for (const Interface &p : interfaces) {
        const int rec[18] = {p.block, p.direction, p.lo[0], p.lo[1], p.lo[2], p.hi[0], p.hi[1], p.hi[2], p.nb_block,
                             p.perm[0], p.perm[1], p.perm[2], p.sign[0], p.sign[1], p.sign[2], p.offset[0], p.offset[1], p.offset[2]};
        patches.insert(patches.end(), rec, rec + 18);
    }
    for (int b = 0; b < nb; ++b) {
        const Block &B = blocks[b];
        for (int f = 0; f < 6; ++f) {
            const int a = f / 2, dir = (f & 1) ? a + 1 : -(a + 1);
            const std::string where = "block " + std::to_string(b) + " (rank " + std::to_string(B.rank) + ", block " +
                                      std::to_string(B.local) + ") face " + std::to_string(f);
            bool conn = false;
            for (const Interface &p : interfaces) conn |= p.block == b && p.direction == dir;
            const Physical *ph = nullptr;
            int nph = 0;
            for (const Physical &p : physical)
                if (p.block == b && p.direction == dir) ph = &p, ++nph;
            if (conn && nph) fail(where + " mixes connections and physical boundaries");
            if (conn) {
                bc[6 * b + f] = 3;
                continue;
            }
            if (nph != 1) fail(where + (nph ? " is split into several physical patches" : " has no boundary or connection"));
            for (int d = 0; d < 3; ++d)
                if (d != a && (ph->sub[d] != 0 || ph->sup[d] + 1 != B.nn[d])) fail(where + ": a physical patch must cover the whole face");
            if (ph->bc_name == "Farfield") bc[6 * b + f] = 1;
            else if (ph->bc_name == "Solid_Surface") bc[6 * b + f] = 2;
            else fail(where + ": boundary '" + ph->bc_name + "' is not supported (Farfield and Solid_Surface are)");
        }
    }
}

void Geometry::face_topology(int *bc, int *nbr, int *transform) const {
    const int nb = (int)blocks.size();
    for (int b = 0; b < nb; ++b) {
        const Block &B = blocks[b];
        for (int f = 0; f < 6; ++f) {
            const int a = f / 2, dir = (f & 1) ? a + 1 : -(a + 1);
            const std::string where = "block " + std::to_string(b) + " (rank " + std::to_string(B.rank) + ", block " +
                                      std::to_string(B.local) + ") face " + std::to_string(f);
            auto whole = [&](const int *lo, const int *hi) {
                for (int d = 0; d < 3; ++d)
                    if (d != a && (lo[d] != 0 || hi[d] != B.nn[d])) return false;
                return true;
            };
            int count = 0;
            const Physical *ph = nullptr;
            const Interface *ifc = nullptr;
            bool partial = false;
            for (const Physical &p : physical)
                if (p.block == b && p.direction == dir) {
                    ++count;
                    ph = &p;
                    int lo[3], hi[3];
                    for (int d = 0; d < 3; ++d) {
                        lo[d] = p.sub[d];
                        hi[d] = p.sup[d] + 1;
                    }
                    partial |= !whole(lo, hi);
                }
            for (const Interface &p : interfaces)
                if (p.block == b && p.direction == dir) {
                    ++count;
                    ifc = &p;
                    partial |= !whole(p.lo, p.hi);
                }
            if (count == 0) fail(where + " has no boundary or connection");
            if (count > 1 || partial)
                fail(where + " is split into " + std::to_string(count) +
                     " patches; only whole-face boundaries and connections are supported");
            int *t = transform + 9 * (6 * b + f);
            for (int d = 0; d < 3; ++d) {
                t[d] = d;
                t[3 + d] = 1;
                t[6 + d] = 0;
            }
            nbr[6 * b + f] = -1;
            if (ph) {
                if (ph->bc_name == "Farfield") bc[6 * b + f] = 1;
                else if (ph->bc_name == "Solid_Surface") bc[6 * b + f] = 2;
                else fail(where + ": boundary '" + ph->bc_name + "' is not supported (Farfield and Solid_Surface are)");
            } else {
                bc[6 * b + f] = 3;
                nbr[6 * b + f] = ifc->nb_block;
                for (int d = 0; d < 3; ++d) {
                    t[d] = ifc->perm[d];
                    t[3 + d] = ifc->sign[d];
                    t[6 + d] = ifc->offset[d];
                }
            }
        }
    }
}

}  // namespace casefile
}  // namespace merc
