// Host runtime of the MI355X Mercury explicit step and its C ABI.
//
// C++ classes keep the roles of the reference's host interfaces:
//   Field        field registry + device SoA arena      (include/3_field/Field.h:16-96)
//   Halo         periodic halo and owner/alias plans     (include/4_halo/Halo.h:15-68)
//   BoundaryCore physical ghost-fill plans               (include/5_boundary/Boundary.h:19-122)
//   DEC          d0 / d1 / d2 on device forms            (include/8_dec/DecOps.h:8-68)
//   MercurySolver the explicit step                      (Z4_Mercury/include/MercurySolver.h:31-48)
// The C ABI at the bottom (declared in include/mercury_hip.h) is the drop-in
// boundary; every entry returns an int status and never throws.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mercury_hip.h"
#include "mercury_common.h"
#include "mercury_geometry.h"
#include "mercury_kernels.h"
#include "mercury_fused.h"
#include "mercury_patches.h"
#include "mercury_singular.h"
#include "mercury_singular_dev.h"
#include "mercury_case.h"
#include "mercury_io.h"
#include "mercury_output.h"
#include "mercury_resist.h"
#include "mercury_dec.h"

static_assert(sizeof(MercuryPhysics) == sizeof(MercPhys), "MercuryPhysics / MercPhys layout mismatch");

namespace merc {

static thread_local std::string g_last_error;

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static const int DELTA[8][3] = {{1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};

// ---------------------------------------------------------------------------
// field catalogue (names of Z4_Mercury/include/0_MercuryFieldCatalog.h:70-202 and
// src/7_metric/Metric.cpp:12-76; reconstruction weights keep their names)
// ---------------------------------------------------------------------------
struct FieldSpec {
    const char *name;
    int slot, ncomp, loc;
};
static const FieldSpec kFields[] = {
    {"U_H", SL_UH, 5, LOC_CELL}, {"U_Na", SL_UN, 5, LOC_CELL}, {"PV_H", SL_PVH, 5, LOC_CELL}, {"PV_Na", SL_PVN, 5, LOC_CELL},
    {"U_plus", SL_UP, 3, LOC_CELL}, {"B_cell", SL_BC, 3, LOC_CELL}, {"Bind_cell", SL_BI, 3, LOC_CELL}, {"J_cell", SL_JC, 3, LOC_CELL},
    {"RHS_H", SL_RH, 5, LOC_CELL}, {"RHS_Na", SL_RN, 5, LOC_CELL}, {"Ehc", SL_EHC, 3, LOC_CELL}, {"hall_beta", SL_HB, 1, LOC_CELL},
    {"Jac", SL_JAC, 1, LOC_CELL}, {"Bcell_from_Bface_w", SL_WB, 18, LOC_CELL}, {"Jcell_from_Jedge_w", SL_WJ, 36, LOC_CELL},
    {"Photo_rate", SL_PHOTO, 1, LOC_CELL}, {"cell_center", SL_CC, 3, LOC_CELL}, {"divB", SL_DIVB, 1, LOC_CELL},
    {"Na", SL_NANEU, 1, LOC_CELL},
    // the third ion species (He++), no reference counterpart
    {"U_He", SL_UHE, 5, LOC_CELL}, {"PV_He", SL_PVHE, 5, LOC_CELL}, {"RHS_He", SL_RHE, 5, LOC_CELL},
    {"cell_scalar", SL_CELLS, 1, LOC_CELL},
    {"Eres_xi", SL_ERES, 1, LOC_EX}, {"Eres_eta", SL_ERES + 1, 1, LOC_EY}, {"Eres_zeta", SL_ERES + 2, 1, LOC_EZ},
    {"dJ_xi", SL_DJ, 1, LOC_EX}, {"dJ_eta", SL_DJ + 1, 1, LOC_EY}, {"dJ_zeta", SL_DJ + 2, 1, LOC_EZ},
    {"eta_res_xi", SL_ETAR, 1, LOC_EX}, {"eta_res_eta", SL_ETAR + 1, 1, LOC_EY}, {"eta_res_zeta", SL_ETAR + 2, 1, LOC_EZ},
    {"dB_xi", SL_DB, 1, LOC_FX}, {"dB_eta", SL_DB + 1, 1, LOC_FY}, {"dB_zeta", SL_DB + 2, 1, LOC_FZ},
    {"RHS_Bres_xi", SL_RHSBR, 1, LOC_FX}, {"RHS_Bres_eta", SL_RHSBR + 1, 1, LOC_FY}, {"RHS_Bres_zeta", SL_RHSBR + 2, 1, LOC_FZ},
    {"B_xi", SL_B, 1, LOC_FX}, {"B_eta", SL_B + 1, 1, LOC_FY}, {"B_zeta", SL_B + 2, 1, LOC_FZ},
    {"Badd_xi", SL_BADD, 1, LOC_FX}, {"Badd_eta", SL_BADD + 1, 1, LOC_FY}, {"Badd_zeta", SL_BADD + 2, 1, LOC_FZ},
    {"RHS_B_xi", SL_RHSB, 1, LOC_FX}, {"RHS_B_eta", SL_RHSB + 1, 1, LOC_FY}, {"RHS_B_zeta", SL_RHSB + 2, 1, LOC_FZ},
    {"JDxi", SL_JD, 3, LOC_FX}, {"JDet", SL_JD + 3, 3, LOC_FY}, {"JDze", SL_JD + 6, 3, LOC_FZ},
    {"Area_xi", SL_AREA, 1, LOC_FX}, {"Area_eta", SL_AREA + 1, 1, LOC_FY}, {"Area_zeta", SL_AREA + 2, 1, LOC_FZ},
    {"beta_xi", SL_BETA, 1, LOC_FX}, {"beta_eta", SL_BETA + 1, 1, LOC_FY}, {"beta_zeta", SL_BETA + 2, 1, LOC_FZ},
    {"Eface_xi", SL_EF, 3, LOC_FX}, {"Eface_eta", SL_EF + 3, 3, LOC_FY}, {"Eface_zeta", SL_EF + 6, 3, LOC_FZ},
    {"face_xi", SL_FACEF, 1, LOC_FX}, {"face_eta", SL_FACEF + 1, 1, LOC_FY}, {"face_zeta", SL_FACEF + 2, 1, LOC_FZ},
    {"E_xi", SL_E, 1, LOC_EX}, {"E_eta", SL_E + 1, 1, LOC_EY}, {"E_zeta", SL_E + 2, 1, LOC_EZ},
    {"Ehall_xi", SL_EH, 1, LOC_EX}, {"Ehall_eta", SL_EH + 1, 1, LOC_EY}, {"Ehall_zeta", SL_EH + 2, 1, LOC_EZ},
    {"J_xi", SL_J, 1, LOC_EX}, {"J_eta", SL_J + 1, 1, LOC_EY}, {"J_zeta", SL_J + 2, 1, LOC_EZ},
    {"dr_xi", SL_DR, 3, LOC_EX}, {"dr_eta", SL_DR + 3, 3, LOC_EY}, {"dr_zeta", SL_DR + 6, 3, LOC_EZ},
    {"dl_xi", SL_DL, 1, LOC_EX}, {"dl_eta", SL_DL + 1, 1, LOC_EY}, {"dl_zeta", SL_DL + 2, 1, LOC_EZ},
    {"alpha_xi", SL_ALPHA, 1, LOC_EX}, {"alpha_eta", SL_ALPHA + 1, 1, LOC_EY}, {"alpha_zeta", SL_ALPHA + 2, 1, LOC_EZ},
    {"taper_xi", SL_TAPER, 1, LOC_EX}, {"taper_eta", SL_TAPER + 1, 1, LOC_EY}, {"taper_zeta", SL_TAPER + 2, 1, LOC_EZ},
    {"eta_local_xi", SL_LETA, 1, LOC_EX}, {"eta_local_eta", SL_LETA + 1, 1, LOC_EY}, {"eta_local_zeta", SL_LETA + 2, 1, LOC_EZ},
    {"edge_xi", SL_EDGEF, 1, LOC_EX}, {"edge_eta", SL_EDGEF + 1, 1, LOC_EY}, {"edge_zeta", SL_EDGEF + 2, 1, LOC_EZ},
    {"node_scalar", SL_NODEF, 1, LOC_NODE},
};

static const FieldSpec *find_field(const char *name) {
    for (const auto &f : kFields)
        if (!std::strcmp(f.name, name)) return &f;
    return nullptr;
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
template <class T>
struct DBuf {
    T *p = nullptr;
    long long n = 0;
    void upload(const std::vector<T> &h) {
        release();
        n = (long long)h.size();
        if (n == 0) return;
        HIPCHK(hipMalloc(&p, sizeof(T) * n));
        HIPCHK(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

// ---------------------------------------------------------------------------
// Field: one SoA arena, slot-major, every slot spanning all blocks
// ---------------------------------------------------------------------------
class Field {
public:
    std::vector<MercBlk> blk;
    long long Vtot = 0;
    double *arena = nullptr;
    double **dptr = nullptr;  // device table of slot base pointers
    std::vector<double *> hptr;
    MercBlk *dblk = nullptr;

    // per block face: neighbour block and node-index translation of a periodic
    // self-connection or an interface (cnb = -1 on physical faces)
    std::vector<int> cnb;
    std::vector<std::array<int, 3>> coff;

    Field(int nb, const int *dims, int g, const int *bc, const int *nbr, const int *trans) {
        blk.resize(nb);
        cnb.assign(6 * nb, -1);
        coff.assign(6 * nb, {0, 0, 0});
        long long off = 0;
        for (int b = 0; b < nb; ++b) {
            MercBlk &k = blk[b];
            k.off = off;
            k.g = g;
            for (int a = 0; a < 3; ++a) {
                k.m[a] = dims[3 * b + a];
                k.N[a] = k.m[a] + 1 + 2 * g;
            }
            k.sJ = k.N[2];
            k.sI = k.N[1] * k.N[2];
            for (int f = 0; f < 6; ++f) {
                k.bc[f] = bc[6 * b + f];
                const int a = f / 2;
                if (k.bc[f] == MBC_PERIODIC) {
                    cnb[6 * b + f] = b;
                    coff[6 * b + f][a] = (f & 1) ? -k.m[a] : k.m[a];
                } else if (k.bc[f] == MBC_INTERFACE && nbr && trans) {
                    cnb[6 * b + f] = nbr[6 * b + f];
                    for (int c = 0; c < 3; ++c) coff[6 * b + f][c] = trans[9 * (6 * b + f) + 6 + c];
                }
            }
            off += (long long)k.N[0] * k.N[1] * k.N[2];
        }
        Vtot = (off + 63) / 64 * 64;
        if (Vtot >= (1LL << 31)) throw std::runtime_error("total box volume exceeds the int32 index range");
        HIPCHK(hipMalloc(&arena, sizeof(double) * Vtot * SL_NSLOT));
        HIPCHK(hipMemset(arena, 0, sizeof(double) * Vtot * SL_NSLOT));
        hptr.resize(SL_NSLOT);
        for (int s = 0; s < SL_NSLOT; ++s) hptr[s] = arena + (long long)s * Vtot;
        HIPCHK(hipMalloc(&dptr, sizeof(double *) * SL_NSLOT));
        HIPCHK(hipMemcpy(dptr, hptr.data(), sizeof(double *) * SL_NSLOT, hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&dblk, sizeof(MercBlk) * nb));
        HIPCHK(hipMemcpy(dblk, blk.data(), sizeof(MercBlk) * nb, hipMemcpyHostToDevice));
    }
    ~Field() {
        if (arena) (void)hipFree(arena);
        if (dptr) (void)hipFree(dptr);
        if (dblk) (void)hipFree(dblk);
    }
    int nb() const { return (int)blk.size(); }
    void upload_blocks() { HIPCHK(hipMemcpy(dblk, blk.data(), sizeof(MercBlk) * blk.size(), hipMemcpyHostToDevice)); }
    long long vol(int b) const { return (long long)blk[b].N[0] * blk[b].N[1] * blk[b].N[2]; }
    int gidx(int b, int i, int j, int k) const {
        const MercBlk &q = blk[b];
        return (int)(q.off + (long long)(i + q.g) * q.sI + (long long)(j + q.g) * q.sJ + (k + q.g));
    }
    void owned_hi(int b, int loc, int hi[3]) const {
        for (int a = 0; a < 3; ++a) hi[a] = blk[b].m[a] + 1 - DELTA[loc][a];
    }
    bool connected() const {
        for (int x : cnb)
            if (x >= 0) return true;
        return false;
    }
    bool has_external() const {
        for (const auto &k : blk)
            for (int f = 0; f < 6; ++f)
                if (k.bc[f] == MBC_EXTERNAL) return true;
        return false;
    }
    // (block, i, j, k) of an arena index (the inverse of gidx)
    void point_of(int x, int p[4]) const {
        int b = 0;
        while (b + 1 < nb() && blk[b + 1].off <= x) ++b;
        const MercBlk &q = blk[b];
        long long r = x - q.off;
        p[0] = b;
        p[1] = (int)(r / q.sI) - q.g;
        r %= q.sI;
        p[2] = (int)(r / q.sJ) - q.g;
        p[3] = (int)(r % q.sJ) - q.g;
    }
    // arena index of (block, i, j, k) with a check that the point lies in the block's box
    int gidx_checked(int b, int i, int j, int k, const char *what) const {
        if (b < 0 || b >= nb()) throw std::runtime_error(std::string(what) + ": block out of range");
        const MercBlk &q = blk[b];
        const int ix[3] = {i, j, k};
        for (int a = 0; a < 3; ++a)
            if (ix[a] < -q.g || ix[a] > q.m[a] + q.g) throw std::runtime_error(std::string(what) + ": point outside the block's box");
        return gidx(b, i, j, k);
    }
    bool has_interface() const {
        for (const auto &k : blk)
            for (int f = 0; f < 6; ++f)
                if (k.bc[f] == MBC_INTERFACE) return true;
        return false;
    }
    void upload_block(int slot, int b, const double *h) {
        HIPCHK(hipMemcpy(hptr[slot] + blk[b].off, h, sizeof(double) * vol(b), hipMemcpyHostToDevice));
    }
    void download_block(int slot, int b, double *h) const {
        HIPCHK(hipMemcpy(h, hptr[slot] + blk[b].off, sizeof(double) * vol(b), hipMemcpyDeviceToHost));
    }
};

// ---------------------------------------------------------------------------
// launch ranges
// ---------------------------------------------------------------------------
struct RangeOwner {
    RangeH h;
    DBuf<long long> start;
    DBuf<int> lo, n;
};

// per block box [lo, lo + n) (n = 0: no point of the block)
static void make_range_boxes(const Field &F, const std::vector<int> &lo, const std::vector<int> &n, RangeOwner &R) {
    std::vector<long long> st(F.nb() + 1, 0);
    for (int b = 0; b < F.nb(); ++b) st[b + 1] = st[b] + (long long)n[3 * b] * n[3 * b + 1] * n[3 * b + 2];
    R.start.upload(st);
    R.lo.upload(lo);
    R.n.upload(n);
    R.h.dev.nb = F.nb();
    R.h.dev.start = R.start.p;
    R.h.dev.lo = R.lo.p;
    R.h.dev.n = R.n.p;
    R.h.total = st.back();
}

static void make_range(const Field &F, int lo_off, int hi_off, RangeOwner &R) {
    // per block box [lo_off, m + hi_off) on every axis
    std::vector<long long> st(F.nb() + 1, 0);
    std::vector<int> lo(3 * F.nb()), n(3 * F.nb());
    for (int b = 0; b < F.nb(); ++b) {
        long long cnt = 1;
        for (int a = 0; a < 3; ++a) {
            lo[3 * b + a] = lo_off;
            n[3 * b + a] = F.blk[b].m[a] + hi_off - lo_off;
            cnt *= n[3 * b + a];
        }
        st[b + 1] = st[b] + cnt;
    }
    R.start.upload(st);
    R.lo.upload(lo);
    R.n.upload(n);
    R.h.dev.nb = F.nb();
    R.h.dev.start = R.start.p;
    R.h.dev.lo = R.lo.p;
    R.h.dev.n = R.n.p;
    R.h.total = st.back();
}

// ---------------------------------------------------------------------------
// synchronisation programs
// ---------------------------------------------------------------------------
enum OpType { OP_COPY, OP_FILL, OP_SETV, OP_WALLF, OP_ZERO };

// one operation of one field (or of fields sharing an index list) in one
// phase of a group's program; host lists only
struct SyncOp {
    OpType type;
    int phase = 0;
    std::vector<int> slots;
    std::vector<double> fillv;
    int is_na = 0;  // WALLF: species (0 H+, 1 Na+, 2 He++)
    std::vector<int> hidx;     // pairs for COPY, (first, deeper, ghosts) for WALLF
    std::vector<double> hval;  // SETV values
};

struct GroupProg {
    std::vector<std::unique_ptr<SyncOp>> ops;
};

// A merged launch: every op of one type in one phase of a set of independent
// groups. Entries name their slot set: COPY (dst, src, set), FILL (dst, set),
// SETV (dst, slot) + value, WALLF (slot, species, first, deeper, g ghosts).
struct MergedOp {
    OpType type;
    MultiSets S{};
    DBuf<int> ent;
    DBuf<double> val;
    long long n = 0;
};
struct MergedProg {
    std::vector<std::unique_ptr<MergedOp>> ops;
};

enum GroupId { G_UCELL = 0, G_BFACE, G_EEDGE, G_EHALL, G_JEDGE, G_BCELL, G_JCELL, G_BADD, G_ERES, G_ERES0, G_DB, G_NGROUP };
// K_ERES0: Eres with a zero far field (the linear part of the resistive operator)
enum Kind { K_UH, K_UNA, K_BFACE, K_EEDGE, K_EHALL, K_JEDGE, K_BCELL, K_BIND, K_JCELL, K_BADD, K_ERES, K_ERES0, K_DB, K_UHE };

struct GField {
    int kind, slot, ncomp, loc, level, owner_sync;
};

// sync groups (Z4_Mercury/src/boundary/Boundary_Build.cpp:564-703) with each
// field's halo level and owner policy (0_MercuryFieldCatalog.h:77-178)
static std::vector<GField> group_fields(int gid, int ns) {
    std::vector<GField> v;
    switch (gid) {
    case G_UCELL:
        v = {{K_UH, SL_UH, 5, LOC_CELL, 1, 0}, {K_UNA, SL_UN, 5, LOC_CELL, 1, 0}};
        if (ns > 2) v.push_back({K_UHE, SL_UHE, 5, LOC_CELL, 1, 0});  // He++ (no reference counterpart)
        break;
    case G_BFACE: for (int a = 0; a < 3; ++a) v.push_back({K_BFACE, SL_B + a, 1, LOC_FX + a, 3, 1}); break;
    case G_BADD: for (int a = 0; a < 3; ++a) v.push_back({K_BADD, SL_BADD + a, 1, LOC_FX + a, 3, 1}); break;
    case G_EEDGE: for (int a = 0; a < 3; ++a) v.push_back({K_EEDGE, SL_E + a, 1, LOC_EX + a, 3, 1}); break;
    case G_EHALL: for (int a = 0; a < 3; ++a) v.push_back({K_EHALL, SL_EH + a, 1, LOC_EX + a, 3, 1}); break;
    case G_JEDGE: for (int a = 0; a < 3; ++a) v.push_back({K_JEDGE, SL_J + a, 1, LOC_EX + a, 3, 1}); break;
    case G_BCELL: v = {{K_BCELL, SL_BC, 3, LOC_CELL, 3, 0}, {K_BIND, SL_BI, 3, LOC_CELL, 3, 0}}; break;
    case G_JCELL: v = {{K_JCELL, SL_JC, 3, LOC_CELL, 3, 0}}; break;
    // Eres / dB groups of the implicit correction (Boundary_Build.cpp:605-613, 674-683)
    case G_ERES: for (int a = 0; a < 3; ++a) v.push_back({K_ERES, SL_ERES + a, 1, LOC_EX + a, 3, 1}); break;
    case G_ERES0: for (int a = 0; a < 3; ++a) v.push_back({K_ERES0, SL_ERES + a, 1, LOC_EX + a, 3, 1}); break;
    case G_DB: for (int a = 0; a < 3; ++a) v.push_back({K_DB, SL_DB + a, 1, LOC_FX + a, 3, 1}); break;
    }
    return v;
}

static std::vector<int> slots_of(const GField &f) {
    std::vector<int> s;
    for (int c = 0; c < f.ncomp; ++c) s.push_back(f.slot + c);
    return s;
}

// Boxes ---------------------------------------------------------------------
struct Box {
    int lo[3], hi[3];
};

// boundary-adjacent inner slab of a full face (LayoutTraits.cpp:434-465)
static Box inner_slab(const Field &F, int b, int loc, int dir) {
    int ohi[3];
    F.owned_hi(b, loc, ohi);
    Box x;
    for (int a = 0; a < 3; ++a) {
        x.lo[a] = 0;
        x.hi[a] = ohi[a];
    }
    const int ax = std::abs(dir) - 1;
    if (dir < 0) {
        x.lo[ax] = 0;
        x.hi[ax] = 1;
    } else {
        x.lo[ax] = ohi[ax] - 1;
        x.hi[ax] = ohi[ax];
    }
    return x;
}
// ghost slab outside an inner slab (LayoutTraits.cpp:467-485)
static Box ghost_slab(const Box &in, int dir, int G) {
    Box x = in;
    const int ax = std::abs(dir) - 1;
    if (dir < 0) {
        x.hi[ax] = in.lo[ax];
        x.lo[ax] = in.lo[ax] - G;
    } else {
        x.lo[ax] = in.hi[ax];
        x.hi[ax] = in.hi[ax] + G;
    }
    return x;
}

template <class Fn>
static void for_box(const Box &x, Fn fn) {
    for (int i = x.lo[0]; i < x.hi[0]; ++i)
        for (int j = x.lo[1]; j < x.hi[1]; ++j)
            for (int k = x.lo[2]; k < x.hi[2]; ++k) fn(i, j, k);
}

static inline int face_dir(int f) { return (f / 2 + 1) * ((f & 1) ? 1 : -1); }

// ---------------------------------------------------------------------------
// BoundaryCore: physical handlers as index lists
// ---------------------------------------------------------------------------
class BoundaryCore {
public:
    // BoundaryCore::DefaultPhysicalCopy (src/5_boundary/Boundary_Physical.cpp:288-315)
    static void default_copy(const Field &F, int b, int loc, int dir, std::vector<int> &pairs) {
        const Box in = inner_slab(F, b, loc, dir);
        const Box gs = ghost_slab(in, dir, F.blk[b].g);
        const int ax = std::abs(dir) - 1;
        const int ref = dir < 0 ? in.lo[ax] : in.hi[ax] - 1;
        for_box(gs, [&](int i, int j, int k) {
            int s[3] = {i, j, k};
            s[ax] = ref;
            pairs.push_back(F.gidx(b, i, j, k));
            pairs.push_back(F.gidx(b, s[0], s[1], s[2]));
        });
    }
    static void ghost_list(const Field &F, int b, int loc, int dir, std::vector<int> &dst) {
        const Box gs = ghost_slab(inner_slab(F, b, loc, dir), dir, F.blk[b].g);
        for_box(gs, [&](int i, int j, int k) { dst.push_back(F.gidx(b, i, j, k)); });
    }
    // fill_physical_corner_ghost on the 2D / 3D corner ghost boxes of physical
    // edges and vertices (src/5_boundary/Boundary_Physical.cpp:39-59, 206-264)
    static void corner_default(const Field &F, int b, int loc, std::vector<int> &pairs) {
        const MercBlk &k = F.blk[b];
        int ohi[3];
        F.owned_hi(b, loc, ohi);
        const int G = k.g;
        auto strip = [&](int ax, int side, int &lo, int &hi) {
            if (side == 0) { lo = -G; hi = 0; }
            else { lo = ohi[ax]; hi = ohi[ax] + G; }
        };
        auto add = [&](const Box &x) {
            for_box(x, [&](int i, int j, int kk) {
                const int ii = std::min(std::max(i, 0), ohi[0] - 1), jj = std::min(std::max(j, 0), ohi[1] - 1),
                          k2 = std::min(std::max(kk, 0), ohi[2] - 1);
                pairs.push_back(F.gidx(b, i, j, kk));
                pairs.push_back(F.gidx(b, ii, jj, k2));
            });
        };
        // an edge or vertex with a periodic or interface face is an inner patch
        // (src/2_topology/TopologyEdge.cpp:161-166): no corner default there
        auto phys = [&](int f) { return k.bc[f] == MBC_FARFIELD || k.bc[f] == MBC_WALL; };
        for (int a = 0; a < 3; ++a)
            for (int c = a + 1; c < 3; ++c)
                for (int sa = 0; sa < 2; ++sa)
                    for (int sc = 0; sc < 2; ++sc) {
                        if (!phys(2 * a + sa) || !phys(2 * c + sc)) continue;
                        const int t = 3 - a - c;
                        Box x;
                        strip(a, sa, x.lo[a], x.hi[a]);
                        strip(c, sc, x.lo[c], x.hi[c]);
                        x.lo[t] = 0;
                        x.hi[t] = ohi[t];
                        add(x);
                    }
        for (int s0 = 0; s0 < 2; ++s0)
            for (int s1 = 0; s1 < 2; ++s1)
                for (int s2 = 0; s2 < 2; ++s2) {
                    if (!phys(s0) || !phys(2 + s1) || !phys(4 + s2)) continue;
                    Box x;
                    strip(0, s0, x.lo[0], x.hi[0]);
                    strip(1, s1, x.lo[1], x.hi[1]);
                    strip(2, s2, x.lo[2], x.hi[2]);
                    add(x);
                }
    }
};

// ---------------------------------------------------------------------------
// Halo: periodic self-connections and 1:1 block interfaces (translations).
// The reference fills ghosts stage by stage (MercuryBoundary::Sync_,
// Z4_Mercury/src/boundary/Boundary_Sync_.cpp:11-99): face strips
// (build_inner_1DCorner_pattern, src/4_halo/build/HaloBuildFace.cpp:6-57; strips
// of src/0_basic/LayoutTraits.cpp:93-130), then edge and vertex corner boxes
// through the face connection of the patch (src/2_topology/TopologyEdge.cpp:
// 132-217), each read at the translated index of the neighbour
// (src/4_halo/component_copy/HaloExchangeInner.cpp:90-247).  A stage-s ghost is
// read one hop away, where it has s-1 out-of-range axes and was written by the
// previous stage, a physical handler or the corner default.  The plan follows
// those hops on the host and copies straight from where the chain ends: an owned
// DOF, or a ghost of physical faces only (written by the handlers and corner
// defaults, which run in earlier phases).
// ---------------------------------------------------------------------------
class Halo {
public:
    static void connected_pairs(const Field &F, int b, int loc, int level, std::vector<int> &pairs) {
        const MercBlk &k = F.blk[b];
        int ohi[3];
        F.owned_hi(b, loc, ohi);
        const int G = k.g;
        for (int i = -G; i < ohi[0] + G; ++i)
            for (int j = -G; j < ohi[1] + G; ++j)
                for (int kk = -G; kk < ohi[2] + G; ++kk) {
                    int idx[3] = {i, j, kk}, nout = 0;
                    for (int a = 0; a < 3; ++a) nout += (idx[a] < 0 || idx[a] >= ohi[a]);
                    if (nout == 0 || nout > level) continue;
                    int cb = b, q[3] = {i, j, kk}, hops = 0;
                    for (;;) {
                        int ch[3];
                        F.owned_hi(cb, loc, ch);
                        int f = -1;
                        for (int a = 0; a < 3 && f < 0; ++a) {
                            const int side = q[a] < 0 ? 0 : (q[a] >= ch[a] ? 1 : -1);
                            if (side >= 0 && F.cnb[6 * cb + 2 * a + side] >= 0) f = 2 * a + side;
                        }
                        if (f < 0) break;
                        const auto &o = F.coff[6 * cb + f];
                        for (int a = 0; a < 3; ++a) q[a] += o[a];
                        cb = F.cnb[6 * cb + f];
                        ++hops;
                    }
                    if (hops == 0) continue;
                    pairs.push_back(F.gidx(b, i, j, kk));
                    pairs.push_back(F.gidx(cb, q[0], q[1], q[2]));
                }
    }
};

// Equivalence classes of the DOFs shared through periodic and interface faces
// (topology equivalence, src/2_topology/TopologyFaceEquiv.cpp,
// TopologyEdgeEquiv.cpp): every owned DOF on a connected face is united with its
// translate in the neighbour; members sorted by arena index, i.e. by EntityKey
// (block, i, j, k); the first member is the owner.
struct Classes {
    std::vector<int> start[8], member[8];
    void build(const Field &F) {
        for (int L = 0; L < 8; ++L) {
            start[L].assign(1, 0);
            member[L].clear();
        }
        if (!F.connected()) return;
        std::vector<int> par(F.Vtot);
        std::vector<char> hit(F.Vtot);
        auto find = [&](int x) {
            while (par[x] != x) x = par[x] = par[par[x]];
            return x;
        };
        for (int L = LOC_FX; L <= LOC_NODE; ++L) {
            for (long long x = 0; x < F.Vtot; ++x) par[x] = (int)x;
            std::fill(hit.begin(), hit.end(), 0);
            for (int b = 0; b < F.nb(); ++b) {
                const MercBlk &k = F.blk[b];
                int ohi[3];
                F.owned_hi(b, L, ohi);
                for (int f = 0; f < 6; ++f) {
                    const int a = f / 2, nb = F.cnb[6 * b + f];
                    if (DELTA[L][a] != 0 || nb < 0) continue;
                    const auto &o = F.coff[6 * b + f];
                    Box x;
                    for (int c = 0; c < 3; ++c) {
                        x.lo[c] = 0;
                        x.hi[c] = ohi[c];
                    }
                    x.lo[a] = (f & 1) ? k.m[a] : 0;
                    x.hi[a] = x.lo[a] + 1;
                    for_box(x, [&](int i, int j, int kk) {
                        const int u = F.gidx(b, i, j, kk), v = F.gidx(nb, i + o[0], j + o[1], kk + o[2]);
                        const int ru = find(u), rv = find(v);
                        if (ru != rv) par[std::max(ru, rv)] = std::min(ru, rv);
                        hit[u] = hit[v] = 1;
                    });
                }
            }
            std::vector<std::pair<int, int>> mem;
            for (long long x = 0; x < F.Vtot; ++x)
                if (hit[x]) mem.push_back({find((int)x), (int)x});
            std::sort(mem.begin(), mem.end());
            for (size_t t = 0; t < mem.size(); ++t) {
                if (t > 0 && mem[t].first != mem[t - 1].first) start[L].push_back((int)t);
                member[L].push_back(mem[t].second);
            }
            if (!mem.empty()) start[L].push_back((int)mem.size());
        }
    }
    int ncls(int L) const { return (int)start[L].size() - 1; }
    // owner -> alias copy pairs (Halo::sync_owner_alias_group,
    // src/4_halo/core/HaloSyncFramework.cpp:695-712), orientation sign +1
    void owner_pairs(int L, std::vector<int> &pairs) const {
        for (int c = 0; c < ncls(L); ++c)
            for (int t = start[L][c] + 1; t < start[L][c + 1]; ++t) {
                pairs.push_back(member[L][t]);
                pairs.push_back(member[L][start[L][c]]);
            }
    }
};

// edge alias classes (ReduceEdgeAliasCandidatesToOwners_, MercurySolver_Current.cpp:3-57):
// per member its arena index, orientation sign and axis (the triplet field it
// lives in); per class the owner's index, sign and axis
struct AliasClasses {
    DBuf<int> start, member, msign, maxis, owner, osign, oaxis;
    long long n = 0;
    // host copies (arena indices), for the export and the resistive unknowns
    std::vector<int> hst{0}, hmem, hms, hmax, how, hos, hoax;
    void set(std::vector<int> st, std::vector<int> mem, std::vector<int> ms, std::vector<int> max, std::vector<int> ow,
             std::vector<int> os, std::vector<int> oax) {
        n = (long long)ow.size();
        start.upload(st);
        member.upload(mem);
        msign.upload(ms);
        maxis.upload(max);
        owner.upload(ow);
        osign.upload(os);
        oaxis.upload(oax);
        hst = std::move(st);
        hmem = std::move(mem);
        hms = std::move(ms);
        hmax = std::move(max);
        how = std::move(ow);
        hos = std::move(os);
        hoax = std::move(oax);
    }
    void build(const Classes &C) {
        std::vector<int> st{0}, mem, ms, max, ow, os, oax;
        for (int a = 0; a < 3; ++a) {
            const int L = LOC_EX + a;
            for (int c = 0; c < C.ncls(L); ++c) {
                for (int t = C.start[L][c]; t < C.start[L][c + 1]; ++t) {
                    mem.push_back(C.member[L][t]);
                    ms.push_back(1);
                    max.push_back(a);
                }
                st.push_back((int)mem.size());
                ow.push_back(C.member[L][C.start[L][c]]);
                os.push_back(1);
                oax.push_back(a);
            }
        }
        set(std::move(st), std::move(mem), std::move(ms), std::move(max), std::move(ow), std::move(os), std::move(oax));
    }
    // non-owner members of each axis (arena index flags)
    void alias_flags(long long V, std::vector<char> flags[3]) const {
        for (int a = 0; a < 3; ++a) flags[a].assign(V, 0);
        for (size_t c = 0; c + 1 < hst.size(); ++c)
            for (int t = hst[c]; t < hst[c + 1]; ++t)
                if (!(hmem[t] == how[c] && hmax[t] == hoax[c])) flags[hmax[t]][hmem[t]] = 1;
    }
    void launch(hipStream_t st, double *const *F, int slot0) const {
        launch_alias_reduce(st, start.p, member.p, msign.p, maxis.p, owner.p, osign.p, oaxis.p, n, F, slot0);
    }
};

// Ends a stream capture that an exception interrupted: the capture is ended on
// every stream that is still capturing (the origin and any stream forked into
// it) and the partial graph is dropped, so the streams are usable again.
struct CaptureGuard {
    std::vector<hipStream_t> streams;
    bool active = false;
    explicit CaptureGuard(std::vector<hipStream_t> s) : streams(std::move(s)) {}
    ~CaptureGuard() {
        if (!active) return;
        for (hipStream_t s : streams) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs == hipStreamCaptureStatusNone) continue;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
    }
};

// ---------------------------------------------------------------------------
// MercurySolver
// ---------------------------------------------------------------------------
class MercurySolver {
public:
    std::unique_ptr<Field> F;
    MercPhys P{};  // the physics of run 0; every run shares all of it but the per-run fields below
    // Ensemble context: independent runs, each a set of blocks with its own
    // physics (hall_coef, inflow state, IMF), dt and abort status; one launch
    // per pass covers every block of every run.  A plain context has one run.
    int nruns = 1;
    std::vector<int> brun;        // run of each block
    std::vector<MercPhys> Prun;   // physics of each run
    const MercPhys &Pb(int b) const { return Prun[brun[b]]; }
    // ion species of the run: 2 (H+, Na+) or 3 (He++ added)
    int nspecies() const { return P.n_species >= 3.0 ? 3 : 2; }
    std::vector<std::unique_ptr<BlockGeom>> geom;
    RangeOwner R_cell, R_box, R_node, R_eface;
    RangeOwner R_res;  // node-box points around the resistive unknowns (the CG's curls)
    RangeOwner R_march;  // flux pass: ceil(m0 / march) x m1 x m2 per block (column kernel)
    int march = 4;       // owned cells per thread along i in the flux pass
    RangeOwner R_tile;   // tiled flux pass: ceil(m0 / tile_c) x ceil(m1 / 16) x ceil(m2 / 16) tiles per block
    int tile_c = 0;      // planes per tile along i; 0: the column kernel runs
    GroupProg prog[G_NGROUP];
    std::map<unsigned, MergedProg> mprog;  // keyed by the bit set of groups
    // caller sync plans: per group, stage 0 stands in for the context's own
    // halo copies and stage 1 for its owner -> alias copies; entries
    // (dst slot, dst, src slot, src, sign) on arena indices
    bool caller_plan[G_NGROUP] = {};
    std::vector<int> plan_h[G_NGROUP][2];
    DBuf<int> plan_d[G_NGROUP][2];
    // a plan whose sources include its own destinations runs staged (every
    // read before every write) through plan_tmp
    bool plan_staged[G_NGROUP][2] = {};
    DBuf<double> plan_tmp[G_NGROUP][2];
    // patch connections (rotated, mirrored, partial-face interfaces): their
    // halo and owner -> alias copies are the groups' plans
    std::unique_ptr<patch::Topo> ptopo;
    bool any_caller_plan() const {
        for (bool b : caller_plan)
            if (b) return true;
        return false;
    }
    Classes cls;          // shared-DOF classes (owner / aliases) of every location
    AliasClasses alias;
    // singular edges of a patch-connected context (none otherwise)
    struct SingularEdges {
        DBuf<MercSingEntry> e;
        DBuf<MercSingCell> c;
        DBuf<MercSingFace> f;
        int n = 0;
        singular::Registry R;
        // Tecplot: stencils of the singular end nodes and their owned aliases
        DBuf<int> nstart, ncell;
        int nnode = 0;
        std::vector<patch::Dof> nalias;
    } sing;
    hipStream_t st = nullptr;      // the stream all of the context's work goes to
    hipStream_t own_st = nullptr;  // the context's own stream (st unless a caller's stream is set)
    // side stream of the fused stage: the fluid RHS runs there while the
    // induction EMF chain runs on st (they read the same state and write
    // disjoint fields); both streams keep the default priority
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fork2 = nullptr, ev_join2 = nullptr;
    bool overlap = true;
    unsigned long long *d_dt = nullptr;       // dt word (double bits)
    unsigned long long *d_dt_next = nullptr;  // next step's dt, reduced at the end of the derived fields
    // run status written by k_dt_commit: [0] stop code (1 = dt underflow, sticky),
    // [1] steps committed since the status was cleared, [2] bits of the refused dt
    unsigned long long *d_status = nullptr;
    // bumped whenever something a captured graph has baked in changes (form,
    // launch shape): graphs captured at an older generation must not replay
    unsigned long long generation = 0;
    int form = 1;                             // 0: one kernel per reference function, 1: fused FP64, 2: mixed
    float *farena = nullptr;                  // FP32 copies of the static arrays (mixed form)
    float **fptr = nullptr;                   // device table indexed like the FP64 slots (null where absent)
    std::vector<int> fslots;                  // FP32 static copies
    std::vector<int> islots;                  // FP32 derived intermediates of the mixed form
    unsigned long long *d_red = nullptr;      // reduction result (2 words)
    double *d_part = nullptr;                 // per-workgroup partial maxima of the div B diagnostic
    // the div B monitor of the fused derived pass: per-workgroup maxima and
    // (last normalised div B, its maximum since the last reset, last max |d2 B|)
    double *d_mpart = nullptr, *d_mon = nullptr;
    bool divb_mon_on = false;  // off by default: it costs the step about 10 us on shell6
    double *mon_part() const { return divb_mon_on ? d_mpart : nullptr; }
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_integrator = -1;
    // implicit resistive correction: unknown edges and the CG work vectors
    bool resist = false;
    ResDofs rd;
    DBuf<int> r_q, r_ax, r_sI, r_sJ;
    DBuf<double> r_eta, r_vec, r_part, r_S;
    double *rwd = nullptr;  // w / dg, the weight of r.z
    double *rx = nullptr, *rr = nullptr, *rp = nullptr, *rAp = nullptr, *rb = nullptr, *rdg = nullptr, *rw = nullptr,
           *rt = nullptr;
    int res_its = 0;
    double res_rel = 0.0;
    static constexpr int RES_CHECK = 4;  // the residual is read back every RES_CHECK iterations

    // Interfaces: the face's IndexTransform (neighbour node index[perm[a]] =
    // sign[a] * index[a] + offset[a], include/2_topology/TopologyTypes.h:93-98)
    // must be a translation gluing the face to the opposite face of the
    // neighbour over its whole extent, declared from both sides, with at least
    // as many cells as ghost layers along the connection.
    static void check_topology(int nb, const int *dims, int g, const int *bc, const int *nbr, const int *trans) {
        for (int b = 0; b < nb; ++b) {
            for (int a = 0; a < 3; ++a)
                if (dims[3 * b + a] < 2) throw std::runtime_error("every block needs at least 2 cells per axis");
            for (int f = 0; f < 6; ++f) {
                const int c = bc[6 * b + f];
                if (c < 0 || c > MBC_EXTERNAL) throw std::runtime_error("unknown boundary code");
                if (c == MBC_PERIODIC && bc[6 * b + (f ^ 1)] != MBC_PERIODIC) throw std::runtime_error("periodic faces must be paired");
                if (c == MBC_PERIODIC && dims[3 * b + f / 2] < g) throw std::runtime_error("a periodic axis needs at least nghost cells");
                if (c != MBC_INTERFACE) continue;
                if (!nbr || !trans) throw std::runtime_error("interface faces need neighbour blocks and index transforms");
                const int n = nbr[6 * b + f], a = f / 2, fo = f ^ 1;
                if (n < 0 || n >= nb || bc[6 * n + fo] != MBC_INTERFACE || nbr[6 * n + fo] != b)
                    throw std::runtime_error("an interface must join opposite faces and be declared from both sides");
                const int *t = trans + 9 * (6 * b + f), *u = trans + 9 * (6 * n + fo);
                for (int d = 0; d < 3; ++d) {
                    if (t[d] != d || t[3 + d] != 1)
                        throw std::runtime_error("only translation interfaces (identity permutation, sign +1) are supported");
                    if (t[6 + d] != -u[6 + d]) throw std::runtime_error("interface transforms of the two sides are not inverse");
                    if (d != a && (t[6 + d] != 0 || dims[3 * b + d] != dims[3 * n + d]))
                        throw std::runtime_error("an interface must match the whole face of the neighbour");
                }
                if (t[6 + a] != ((f & 1) ? -dims[3 * b + a] : dims[3 * n + a]))
                    throw std::runtime_error("interface offset does not glue the face to the neighbour's opposite face");
                if (dims[3 * b + a] < g || dims[3 * n + a] < g) throw std::runtime_error("an interface axis needs at least nghost cells");
            }
        }
    }

    // The physics fields an ensemble's runs may differ in: the Hall coefficient,
    // the far-field (inflow) states and the IMF.  Everything else is shared.
    static void check_run_physics(int nruns, const MercPhys *prun) {
        MercPhys a = prun[0];
        auto blank = [](MercPhys &q) {
            q.hall_coef = 0.0;
            for (int c = 0; c < 5; ++c) q.qinf[c] = q.qinfs[c] = q.q_pv_inf[c] = q.qinf_He[c] = 0.0;
            for (int c = 0; c < 3; ++c) q.B_imf[c] = 0.0;
        };
        blank(a);
        for (int r = 1; r < nruns; ++r) {
            MercPhys b = prun[r];
            blank(b);
            if (std::memcmp(&a, &b, sizeof(MercPhys)) != 0)
                throw std::runtime_error("ensemble run " + std::to_string(r) +
                                         ": the runs of one context may differ only in hall_coef, the far-field states and B_imf");
        }
        if (nruns > 1 && prun[0].resist_enabled != 0.0)
            throw std::runtime_error("an ensemble context cannot run the implicit resistive correction");
    }

    MercurySolver(int nb, const int *dims, int g, const int *bc, const int *nbr, const int *trans, const double *nodes,
                  const MercPhys &phys)
        : MercurySolver(nb, dims, g, bc, nbr, trans, nodes, 1, nullptr, &phys) {}

    // block_run: run of each block (NULL with one run); prun: physics of each run
    MercurySolver(int nb, const int *dims, int g, const int *bc, const int *nbr, const int *trans, const double *nodes, int nruns_,
                  const int *block_run, const MercPhys *prun, const patch::Topo *pt = nullptr)
        : P(prun[0]), nruns(nruns_) {
        if (g < 2) throw std::runtime_error("nghost must be >= 2");
        if (nruns < 1) throw std::runtime_error("an ensemble needs at least one run");
        if (pt) {
            if (nruns != 1) throw std::runtime_error("patch connections need a one-run context");
            ptopo.reset(new patch::Topo(*pt));
            nbr = trans = nullptr;
        } else {
            check_topology(nb, dims, g, bc, nbr, trans);
        }
        brun.assign(nb, 0);
        if (block_run)
            for (int b = 0; b < nb; ++b) {
                if (block_run[b] < 0 || block_run[b] >= nruns) throw std::runtime_error("block run id out of range");
                brun[b] = block_run[b];
            }
        for (int b = 0; b < nb; ++b)
            for (int f = 0; f < 6; ++f)
                if (nbr && bc[6 * b + f] == MBC_INTERFACE && brun[nbr[6 * b + f]] != brun[b])
                    throw std::runtime_error("block " + std::to_string(b) + ": an interface joins two runs of the ensemble");
        Prun.assign(prun, prun + nruns);
        check_run_physics(nruns, prun);
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        own_st = st;
        F.reset(new Field(nb, dims, g, bc, nbr, trans));
        for (int b = 0; b < nb; ++b) {
            F->blk[b].run = brun[b];
            F->blk[b].hall_coef = Pb(b).hall_coef;
        }
        F->upload_blocks();
        HIPCHK(hipMalloc(&d_dt, sizeof(unsigned long long) * nruns));
        HIPCHK(hipMalloc(&d_dt_next, sizeof(unsigned long long) * nruns));
        HIPCHK(hipMalloc(&d_status, 4 * sizeof(unsigned long long) * nruns));
        HIPCHK(hipMemset(d_status, 0, 4 * sizeof(unsigned long long) * nruns));
        HIPCHK(hipMalloc(&d_red, 2 * sizeof(unsigned long long)));
        HIPCHK(hipMalloc(&d_part, 2 * sizeof(double) * divb_partials()));
        HIPCHK(hipMalloc(&d_mon, 3 * sizeof(double)));
        HIPCHK(hipMemset(d_mon, 0, 3 * sizeof(double)));
        // static geometry per block
        long long noff = 0;
        for (int b = 0; b < nb; ++b) {
            const int m[3] = {dims[3 * b], dims[3 * b + 1], dims[3 * b + 2]};
            geom.emplace_back(new BlockGeom(m, g, nodes + noff));
            noff += 3LL * (m[0] + 1) * (m[1] + 1) * (m[2] + 1);
            struct W : SlotWriter {
                Field *F;
                int b;
                void put(int slot, const std::vector<double> &v) override { F->upload_block(slot, b, v.data()); }
            } w;
            w.F = F.get();
            w.b = b;
            geom[b]->compute(Pb(b), w);
        }
        // FP32 static copies: reconstruction weights, Hodge stars, edge geometry, taper, eta, photo rate
        for (int c = 0; c < 18; ++c) fslots.push_back(SL_WB + c);
        for (int c = 0; c < 36; ++c) fslots.push_back(SL_WJ + c);
        for (int c = 0; c < 3; ++c) {
            fslots.push_back(SL_BETA + c);
            fslots.push_back(SL_ALPHA + c);
            fslots.push_back(SL_DL + c);
            fslots.push_back(SL_AREA + c);
            fslots.push_back(SL_TAPER + c);
            fslots.push_back(SL_LETA + c);
        }
        for (int c = 0; c < 9; ++c) fslots.push_back(SL_DR + c);
        fslots.push_back(SL_PHOTO);
        // FP32 derived intermediates of the mixed form (written by its passes):
        // PV u, v, w, T (the pressure stays FP64), U+, Hall cell terms,
        // contravariant U+, face EMFs, fluid RHS
        for (int s = 0; s < nspecies(); ++s)
            for (int c : {0, 1, 2, 4}) islots.push_back(SPEC_PV(s) + c);
        for (int c = 0; c < 3; ++c) islots.push_back(SL_UP + c);
        for (int c = 0; c < 3; ++c) islots.push_back(SL_EHC + c);
        islots.push_back(SL_HB);
        for (int c = 0; c < 3; ++c) islots.push_back(SL_UCON + c);
        for (int c = 0; c < 9; ++c) islots.push_back(SL_EF + c);
        for (int c = 0; c < 5; ++c) islots.push_back(SL_RH + c);
        for (int c = 0; c < 5; ++c) islots.push_back(SL_RN + c);
        if (nspecies() > 2)
            for (int c = 0; c < 5; ++c) islots.push_back(SL_RHE + c);
        islots.push_back(SL_PEN);
        const size_t nf = fslots.size() + islots.size();
        HIPCHK(hipMalloc(&farena, sizeof(float) * F->Vtot * nf));
        HIPCHK(hipMemset(farena, 0, sizeof(float) * F->Vtot * nf));
        std::vector<float *> hf(SL_NSLOT, nullptr);
        for (size_t n = 0; n < fslots.size(); ++n) hf[fslots[n]] = farena + (long long)n * F->Vtot;
        for (size_t n = 0; n < islots.size(); ++n) hf[islots[n]] = farena + (long long)(fslots.size() + n) * F->Vtot;
        HIPCHK(hipMalloc(&fptr, sizeof(float *) * SL_NSLOT));
        HIPCHK(hipMemcpy(fptr, hf.data(), sizeof(float *) * SL_NSLOT, hipMemcpyHostToDevice));
        fptr_h = hf;
        make_range(*F, 0, 0, R_cell);
        make_range(*F, -g, g + 1, R_box);
        make_range(*F, 0, 1, R_node);
        HIPCHK(hipMalloc(&d_mpart, 2 * sizeof(double) * (size_t)std::max<long long>((R_node.h.total + 255) / 256, 1)));
        make_range(*F, -1, 2, R_eface);
        if (const char *e = std::getenv("MERCURY_RHS_MARCH")) march = std::max(1, std::atoi(e));
        if (const char *e = std::getenv("MERCURY_RHS_TILE")) tile_c = std::max(0, std::atoi(e));
        if (const char *e = std::getenv("MERCURY_OVERLAP")) overlap = std::atoi(e) != 0;
        HIPCHK(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_fork2, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_join2, hipEventDisableTiming));
        set_flux_shape(tile_c > 0 ? 1 : 0, tile_c > 0 ? tile_c : march);
        cls.build(*F);
        build_programs();
        alias.build(cls);
        if (ptopo) apply_patch_plans();
        build_resist();
    }
    ~MercurySolver() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_dt) (void)hipFree(d_dt);
        if (d_dt_next) (void)hipFree(d_dt_next);
        if (d_status) (void)hipFree(d_status);
        if (farena) (void)hipFree(farena);
        if (fptr) (void)hipFree(fptr);
        if (d_red) (void)hipFree(d_red);
        if (d_part) (void)hipFree(d_part);
        if (d_mpart) (void)hipFree(d_mpart);
        if (d_mon) (void)hipFree(d_mon);
        F.reset();
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_fork2) (void)hipEventDestroy(ev_fork2);
        if (ev_join2) (void)hipEventDestroy(ev_join2);
        if (st2) (void)hipStreamDestroy(st2);
        if (own_st) (void)hipStreamDestroy(own_st);
    }

    // Moves the context's work to a caller's stream (nullptr: back to its own).
    // Work already queued on the current stream and on the side stream
    // completes first, so the order of the caller's calls is kept.
    void set_stream(hipStream_t s) {
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipStreamSynchronize(st2));
        st = s ? s : own_st;
    }

    double *const *Fd() const { return F->dptr; }
    const float *const *Gd() const { return fptr; }
    // the FP64 copy of an intermediate the mixed form keeps in FP32 (downloads)
    void materialize(int slot) {
        if (form != 2) return;
        for (int s : islots)
            if (s == slot) {
                launch_to_double(st, fptr_h[s], F->hptr[s], F->Vtot);
                HIPCHK(hipStreamSynchronize(st));
            }
    }
    std::vector<float *> fptr_h;
    bool mix() const { return form == 2; }
    // Launch-shape autotuner: times the flux pass (HIP events on the solver
    // stream) as the column kernel with march lengths 1, 2, 4, 8 and as the
    // tiled kernel with 4, 8, 16 planes per tile, in the current form, and
    // keeps the fastest; the choice never changes a result.  Returns the
    // chosen length (cells per thread, or planes per tile when tiled).
    int autotune() {
        int best_len = tile_c > 0 ? tile_c : march, best_kind = flux_kind();
        double tbest = 1e300;
        for (int kind = 0; kind < 2; ++kind)
            for (int L : kind == 1 ? std::vector<int>{4, 8, 16} : std::vector<int>{1, 2, 4, 8}) {
                set_flux_shape(kind, L);
                const double t = time_op("rhs_cell", 5);
                if (t < tbest) tbest = t, best_len = L, best_kind = kind;
            }
        set_flux_shape(best_kind, best_len);
        return best_len;
    }
    // 0: column kernel, 1: tiled kernel
    int flux_kind() const { return tile_c > 0 ? 1 : 0; }

    // the flux pass's launch shape (kind as flux_kind): L cells per thread
    // along i for the column kernel, L planes per 16 x 16 tile for the tiled
    // one; drops a captured step graph (it holds the old shape)
    void set_flux_shape(int kind, int L) {
        if (L < 1) throw std::runtime_error("mercury_set_flux_shape: length must be >= 1");
        if (kind < 0 || kind > 1) throw std::runtime_error("mercury_set_flux_shape: kind must be 0 (column) or 1 (tiled)");
        ++generation;
        if (gexec) {
            (void)hipGraphExecDestroy(gexec);
            (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
        }
        std::vector<int> lo(3 * F->nb(), 0), n(3 * F->nb());
        if (kind == 1) {
            tile_c = L;
            for (int b = 0; b < F->nb(); ++b) {
                n[3 * b] = (F->blk[b].m[0] + L - 1) / L;
                n[3 * b + 1] = (F->blk[b].m[1] + 15) / 16;
                n[3 * b + 2] = (F->blk[b].m[2] + 15) / 16;
            }
            make_range_boxes(*F, lo, n, R_tile);
        } else {
            tile_c = 0;
            march = L;
            for (int b = 0; b < F->nb(); ++b) {
                n[3 * b] = (F->blk[b].m[0] + L - 1) / L;
                n[3 * b + 1] = F->blk[b].m[1];
                n[3 * b + 2] = F->blk[b].m[2];
            }
            make_range_boxes(*F, lo, n, R_march);
        }
    }
    void launch_flux(hipStream_t s) {
        if (tile_c > 0) launch_rhs_tile(s, R_tile.h, Bd(), Fd(), Gd(), P, tile_c, mix());
        else launch_rhs_cell(s, R_march.h, Bd(), Fd(), Gd(), P, march, mix());
    }

    // refresh the FP32 static copies from the FP64 arrays
    void refresh_float_statics() {
        for (int s : fslots) launch_to_float(st, F->hptr[s], fptr_h[s], F->Vtot);
    }

    // Sync programs: for every group the physical handlers (far field before
    // the wall, the order of the sorted physical patch list), the corner
    // defaults, the staged periodic halo and the owner -> alias copies
    // (MercuryBoundary::Sync_, Z4_Mercury/src/boundary/Boundary_Sync_.cpp:4-110;
    // handlers Boundary_Build.cpp:5-300)
    void build_programs() {
        const Field &Fl = *F;
        for (int gid = 0; gid < G_NGROUP; ++gid) {
            std::vector<std::unique_ptr<SyncOp>> A, Bc, C, D, H, E;
            auto add = [](std::vector<std::unique_ptr<SyncOp>> &ph, OpType t, const std::vector<int> &slots, std::vector<int> &&idx,
                          std::vector<double> &&val = {}, const double *fill = nullptr, int is_na = 0) {
                if (idx.empty()) return;
                if (t == OP_COPY)
                    for (auto &o : ph)
                        if (o->type == OP_COPY && o->hidx == idx && o->slots.size() + slots.size() <= MERC_MAX_SYNC_SLOTS) {
                            o->slots.insert(o->slots.end(), slots.begin(), slots.end());
                            return;
                        }
                std::unique_ptr<SyncOp> o(new SyncOp());
                o->type = t;
                o->slots = slots;
                o->hidx = std::move(idx);
                o->hval = std::move(val);
                o->is_na = is_na;
                if (fill) o->fillv.assign(fill, fill + slots.size());
                ph.push_back(std::move(o));
            };
            for (const GField &gf : group_fields(gid, nspecies())) {
                const auto sl = slots_of(gf);
                std::vector<int> fillA, copyB, setd, wallC, zeroC, copyD, haloH, ownE;
                std::vector<double> setv;
                const double *fill = nullptr;
                static const double zero3[3] = {0, 0, 0};
                // per-run far-field values: with one run a constant fill, in an
                // ensemble explicit per-entry values (one SETV list per component)
                std::vector<std::vector<int>> fill_d(gf.ncomp);
                std::vector<std::vector<double>> fill_v(gf.ncomp);
                auto farfill = [&](int b, int dir, const double *val) {
                    if (nruns == 1) {
                        BoundaryCore::ghost_list(Fl, b, gf.loc, dir, fillA);
                        fill = val;
                        return;
                    }
                    std::vector<int> d;
                    BoundaryCore::ghost_list(Fl, b, gf.loc, dir, d);
                    for (int c = 0; c < gf.ncomp; ++c)
                        for (int x : d) {
                            fill_d[c].push_back(x);
                            fill_v[c].push_back(val[c]);
                        }
                };
                for (int b = 0; b < Fl.nb(); ++b) {
                    const MercBlk &k = Fl.blk[b];
                    const MercPhys &Q = Pb(b);
                    for (int f = 0; f < 6; ++f) {
                        if (k.bc[f] != MBC_FARFIELD) continue;
                        const int dir = face_dir(f);
                        switch (gf.kind) {
                        case K_UH: farfill(b, dir, Q.qinf); break;
                        case K_UNA: farfill(b, dir, Q.qinfs); break;
                        case K_UHE: farfill(b, dir, Q.qinf_He); break;
                        case K_BFACE: break;
                        case K_BCELL: farfill(b, dir, Q.B_imf); break;
                        case K_BIND: farfill(b, dir, zero3); break;
                        case K_EEDGE:
                        case K_ERES:
                        case K_ERES0: {
                            // BC_Farfield_Eedge_set_zerocurl (Boundary_Condition.cpp:573-623), also the
                            // Eres far-field handler (Boundary_Build.cpp:198-200); K_ERES0 sets 0
                            const double *u = Q.q_pv_inf, *Bi = Q.B_imf;
                            const bool zero = gf.kind == K_ERES0;
                            const double Ei[3] = {Bi[1] * u[2] - Bi[2] * u[1], Bi[2] * u[0] - Bi[0] * u[2], Bi[0] * u[1] - Bi[1] * u[0]};
                            const int a = gf.loc - LOC_EX;
                            const BlockGeom &G = *geom[b];
                            for_box(inner_slab(Fl, b, gf.loc, dir), [&](int i, int j, int kk) {
                                const auto r0 = G.node(i, j, kk), r1 = G.node(i + (a == 0), j + (a == 1), kk + (a == 2));
                                setd.push_back(Fl.gidx(b, i, j, kk));
                                setv.push_back(zero ? 0.0 : Ei[0] * (r1[0] - r0[0]) + Ei[1] * (r1[1] - r0[1]) + Ei[2] * (r1[2] - r0[2]));
                            });
                            BoundaryCore::default_copy(Fl, b, gf.loc, dir, copyB);
                            break;
                        }
                        default: BoundaryCore::default_copy(Fl, b, gf.loc, dir, copyB); break;
                        }
                    }
                    for (int f = 0; f < 6; ++f) {
                        if (k.bc[f] != MBC_WALL) continue;
                        const int dir = face_dir(f);
                        const int ax = std::abs(dir) - 1, sgn = dir > 0 ? 1 : -1;
                        if (gf.kind == K_UH || gf.kind == K_UNA || gf.kind == K_UHE) {
                            // BC_Solid_Surface_ (Boundary_Condition_Fluid.cpp:36-260): (first, deeper, ghosts)
                            int lo[3], hi[3];
                            for (int a = 0; a < 3; ++a) { lo[a] = -k.g; hi[a] = k.m[a] + k.g; }
                            auto cl = [&](int v, int a) { return std::min(std::max(v, lo[a]), hi[a] - 1); };
                            for_box(inner_slab(Fl, b, LOC_CELL, dir), [&](int i, int j, int kk) {
                                int c[3] = {i, j, kk};
                                wallC.push_back(Fl.gidx(b, i, j, kk));
                                int r[3] = {c[0], c[1], c[2]};
                                r[ax] -= sgn;
                                wallC.push_back(Fl.gidx(b, cl(r[0], 0), cl(r[1], 1), cl(r[2], 2)));
                                for (int l = 1; l <= k.g; ++l) {
                                    int q[3] = {c[0], c[1], c[2]};
                                    q[ax] += l * sgn;
                                    wallC.push_back(Fl.gidx(b, cl(q[0], 0), cl(q[1], 1), cl(q[2], 2)));
                                }
                            });
                        } else if (gf.kind == K_JCELL) {
                            // BC_Solid_Surface_Jcell (Boundary_Condition_EM.cpp:382-427)
                            for_box(inner_slab(Fl, b, LOC_CELL, dir), [&](int i, int j, int kk) {
                                int c[3] = {i, j, kk};
                                zeroC.push_back(Fl.gidx(b, i, j, kk));
                                c[ax] -= sgn;
                                zeroC.push_back(Fl.gidx(b, c[0], c[1], c[2]));
                            });
                            BoundaryCore::default_copy(Fl, b, gf.loc, dir, copyD);
                        }
                    }
                    BoundaryCore::corner_default(Fl, b, gf.loc, copyD);
                    Halo::connected_pairs(Fl, b, gf.loc, gf.level, haloH);
                }
                if (gf.owner_sync) cls.owner_pairs(gf.loc, ownE);
                if (!fillA.empty()) add(A, OP_FILL, sl, std::move(fillA), {}, fill);
                for (int c = 0; c < gf.ncomp; ++c)
                    if (!fill_d[c].empty()) add(A, OP_SETV, {sl[c]}, std::move(fill_d[c]), std::move(fill_v[c]));
                if (!setd.empty()) add(A, OP_SETV, sl, std::move(setd), std::move(setv));
                add(Bc, OP_COPY, sl, std::move(copyB));
                if (!wallC.empty()) add(C, OP_WALLF, sl, std::move(wallC), {}, nullptr, gf.kind == K_UNA ? 1 : (gf.kind == K_UHE ? 2 : 0));
                if (!zeroC.empty()) {
                    static const double z[10] = {0};
                    add(C, OP_FILL, sl, std::move(zeroC), {}, z);
                }
                add(D, OP_COPY, sl, std::move(copyD));
                add(H, OP_COPY, sl, std::move(haloH));
                add(E, OP_COPY, sl, std::move(ownE));
            }
            int phase = 0;
            for (auto *ph : {&A, &Bc, &C, &D, &H, &E}) {
                for (auto &o : *ph) {
                    o->phase = phase;
                    prog[gid].ops.push_back(std::move(o));
                }
                ++phase;
            }
        }
        build_all_merged();
    }
    void build_all_merged() {
        mprog.clear();
        for (int gid = 0; gid < G_NGROUP; ++gid) build_merged(1u << gid);
        build_merged(MG_EDGE_PAIR);
        build_merged(MG_STATE_PAIR);
        build_merged(MG_DERIVED_PAIR);
        build_merged(MG_PREPARE_PAIR);
    }

    // the group's copies of one stage (0: halo, 1: owner -> alias) as
    // (dst slot, dst, src slot, src, sign): the caller's plan when one is set,
    // else the context's own
    std::vector<int> plan_entries(int gid, int stage) const {
        if (caller_plan[gid]) return plan_h[gid][stage];
        std::vector<int> out;
        for (const auto &o : prog[gid].ops) {
            if (o->type != OP_COPY || o->phase != (stage == 0 ? kPhaseHalo : kPhaseOwner)) continue;
            for (size_t e = 0; e + 1 < o->hidx.size(); e += 2)
                for (int sl : o->slots) out.insert(out.end(), {sl, o->hidx[e], sl, o->hidx[e + 1], 1});
        }
        return out;
    }
    void set_sync_plan(int gid, int stage, std::vector<int> ent) {
        if (gid < 0 || gid >= G_NGROUP || stage < 0 || stage > 1) throw std::runtime_error("mercury_set_sync_plan: unknown group or stage");
        if (!caller_plan[gid]) {
            // the other stage starts as the context's own copies
            for (int t = 0; t < 2; ++t) plan_h[gid][t] = plan_entries(gid, t);
            caller_plan[gid] = true;
        }
        plan_h[gid][stage] = std::move(ent);
        upload_plans(gid);
        build_all_merged();
        drop_graph();
    }
    void upload_plans(int gid) {
        for (int t = 0; t < 2; ++t) {
            const std::vector<int> &e = plan_h[gid][t];
            plan_d[gid][t].upload(e);
            std::vector<std::pair<int, int>> dst;
            for (size_t q = 0; q + 4 < e.size(); q += 5) dst.push_back({e[q], e[q + 1]});
            std::sort(dst.begin(), dst.end());
            bool staged = false;
            for (size_t q = 0; q + 4 < e.size() && !staged; q += 5) staged = std::binary_search(dst.begin(), dst.end(), std::make_pair(e[q + 2], e[q + 3]));
            plan_staged[gid][t] = staged;
            plan_tmp[gid][t].upload(std::vector<double>(staged ? e.size() / 5 : 0, 0.0));
        }
    }
    // Plans of every group from the patch connections (rule and regions as
    // merc::patch states): stage 0 the halo, stage 1 owner -> alias; and the
    // edge alias classes of the current and EMF means
    void apply_patch_plans() {
        const patch::Topo &T = *ptopo;
        const Field &Fl = *F;
        for (int gid = 0; gid < G_NGROUP; ++gid) {
            const auto gfs = group_fields(gid, nspecies());
            const int loc = gfs[0].loc;
            const int kind = loc == LOC_CELL ? patch::KIND_COMPONENT : (loc >= LOC_FX && loc <= LOC_FZ ? patch::KIND_FACE : patch::KIND_EDGE);
            for (int t = 0; t < 2; ++t) {
                const std::vector<patch::Entry> ent = t == 0 ? patch::halo_entries(T, kind, gfs[0].level) : patch::owner_entries(T, kind);
                std::vector<int> out;
                for (const patch::Entry &e : ent) {
                    const int d = Fl.gidx(e.dst.b, e.dst.i, e.dst.j, e.dst.k), s = Fl.gidx(e.src.b, e.src.i, e.src.j, e.src.k);
                    if (kind == patch::KIND_COMPONENT) {
                        for (const GField &gf : gfs)
                            for (int c = 0; c < gf.ncomp; ++c) out.insert(out.end(), {gf.slot + c, d, gf.slot + c, s, e.sign});
                    } else {
                        out.insert(out.end(), {gfs[e.dst.f].slot, d, gfs[e.src.f].slot, s, e.sign});
                    }
                }
                plan_h[gid][t] = std::move(out);
            }
            caller_plan[gid] = true;
            upload_plans(gid);
        }
        const patch::EdgeClasses E = patch::edge_classes(T);
        std::vector<int> st{0}, mem, ms, max, ow, os, oax;
        for (size_t c = 0; c + 1 < E.start.size(); ++c) {
            for (int t = E.start[c]; t < E.start[c + 1]; ++t) {
                const patch::Dof &m = E.mem[t];
                mem.push_back(Fl.gidx(m.b, m.i, m.j, m.k));
                ms.push_back(E.sign[t]);
                max.push_back(m.f);
            }
            st.push_back((int)mem.size());
            const patch::Dof &o = E.mem[E.start[c]];
            ow.push_back(Fl.gidx(o.b, o.i, o.j, o.k));
            os.push_back(E.sign[E.start[c]]);
            oax.push_back(o.f);
        }
        alias.set(std::move(st), std::move(mem), std::move(ms), std::move(max), std::move(ow), std::move(os), std::move(oax));
        build_all_merged();
        // a singular edge has three or more aliases: without such a class there
        // is no registry to build (and no geometry to download)
        bool candidates = false;
        for (size_t c = 0; c + 1 < E.start.size(); ++c) candidates |= E.start[c + 1] - E.start[c] >= 3;
        if (candidates) build_singular();
    }

    // The singular-edge registry from the static geometry on the device; the
    // entries' inverse Hodge scale replaces alpha on their owned aliases
    // (SingularEdgeRegistry::build, src/7_metric/SingularEdgeRegistry.cpp:253-266)
    void build_singular() {
        const patch::Topo &T = *ptopo;
        const Field &Fl = *F;
        const int nb = T.nb();
        static const int kSlots[19] = {SL_DL, SL_DL + 1, SL_DL + 2, SL_DR, SL_DR + 1, SL_DR + 2, SL_DR + 3, SL_DR + 4,
                                       SL_DR + 5, SL_DR + 6, SL_DR + 7, SL_DR + 8, SL_JAC, SL_CC, SL_CC + 1, SL_CC + 2,
                                       SL_AREA, SL_AREA + 1, SL_AREA + 2};
        std::vector<std::vector<double>> store((size_t)nb * 19);
        std::vector<singular::BlockGeo> G(nb);
        for (int b = 0; b < nb; ++b) {
            for (int u = 0; u < 19; ++u) {
                std::vector<double> &v = store[(size_t)b * 19 + u];
                v.resize(Fl.vol(b));
                Fl.download_block(kSlots[u], b, v.data());
            }
            singular::set_block_geo(G[b], T.m[b], T.g, &store[(size_t)b * 19]);
        }
        sing.R = singular::build(T, G);
        set_singular(G);
        const singular::NodeStencils NS = singular::node_stencils(T, sing.R);
        std::vector<int> cq;
        for (const patch::Dof &c : NS.cells) cq.push_back(Fl.gidx(c.b, c.i, c.j, c.k));
        sing.nnode = (int)NS.start.size() - 1;
        sing.nstart.upload(NS.start);
        sing.ncell.upload(cq);
        sing.nalias = NS.alias;
    }
    void set_singular(const std::vector<singular::BlockGeo> &G) {
        const singular::Registry &R = sing.R;
        const Field &Fl = *F;
        std::vector<MercSingEntry> he;
        std::vector<MercSingCell> hc;
        std::vector<MercSingFace> hf;
        for (const singular::Owner &o : R.e) {
            MercSingEntry x{};
            x.oq = Fl.gidx(o.dof.b, o.dof.i, o.dof.j, o.dof.k);
            x.oax = o.dof.f;
            x.c0 = o.c0, x.c1 = o.c1, x.f0 = o.f0, x.f1 = o.f1;
            x.osign = o.sign;
            for (int u = 0; u < 3; ++u) x.cv[u] = o.cv[u];
            he.push_back(x);
        }
        for (const singular::Cell &c : R.c) {
            const MercBlk &k = Fl.blk[c.dof.b];
            const MercPhys &Pc = Pb(c.dof.b);
            const singular::BlockGeo &g = G[c.dof.b];
            const long long q = g.idx(c.dof.i, c.dof.j, c.dof.k);
            MercSingCell y{};
            y.q = Fl.gidx(c.dof.b, c.dof.i, c.dof.j, c.dof.k);
            y.d = c.dof.f == 0 ? k.sI : (c.dof.f == 1 ? k.sJ : 1);
            y.blk = c.dof.b;
            y.w = c.w;
            y.orient = c.orient;
            const double x0 = g.cc[0][q], x1 = g.cc[1][q], x2 = g.cc[2][q];
            y.taper = singular::hall_taper(Pc, x0, x1, x2);
            y.leta_on = singular::local_eta(Pc, x0, x1, x2, y.leta) ? 1 : 0;
            hc.push_back(y);
        }
        for (const singular::Face &f : R.f) hf.push_back(MercSingFace{Fl.gidx(f.dof.b, f.dof.i, f.dof.j, f.dof.k), f.dof.f, f.coef});
        sing.n = (int)he.size();
        sing.e.upload(he);
        sing.c.upload(hc);
        sing.f.upload(hf);
        for (size_t t = 0; t < R.alpha_dof.size(); ++t) {
            const patch::Dof &d = R.alpha_dof[t];
            HIPCHK(hipMemcpy(F->hptr[SL_ALPHA + d.f] + Fl.gidx(d.b, d.i, d.j, d.k), &R.alpha_val[t], sizeof(double), hipMemcpyHostToDevice));
        }
    }
    // E and E_hall (after the alias mean of E and the Hall edge EMF) and the
    // current or its increment (after the alias mean) on the singular owners
    void singular_emf(bool mixed) {
        if (sing.n) launch_singular_emf(st, sing.n, sing.e.p, sing.c.p, Bd(), Fd(), Gd(), P, mixed);
    }
    void singular_current(int sB, int sJ) {
        if (sing.n) launch_singular_current(st, sing.n, sing.e.p, sing.f.p, Fd(), sB, sJ);
    }
    void drop_graph() {
        ++generation;
        if (gexec) {
            (void)hipGraphExecDestroy(gexec);
            (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
        }
    }

    // Groups that read and write disjoint slots and that the step syncs back
    // to back run as one program: phase by phase, one launch per op type.
    static constexpr unsigned MG_EDGE_PAIR = (1u << G_EHALL) | (1u << G_EEDGE);
    static constexpr unsigned MG_STATE_PAIR = (1u << G_UCELL) | (1u << G_BFACE);
    static constexpr unsigned MG_DERIVED_PAIR = (1u << G_BCELL) | (1u << G_JEDGE);
    static constexpr unsigned MG_PREPARE_PAIR = (1u << G_BADD) | (1u << G_BFACE);

    static constexpr size_t kWave = 64;  // wavefront width of gfx950
    static constexpr int kPhaseHalo = 4, kPhaseOwner = 5;  // phases of a group's program (A, B, C, D, H, E)
    void build_merged(unsigned gset) {
        MergedProg &mp = mprog[gset];
        const int g = F->blk[0].g;
        for (int phase = 0; phase < 6; ++phase)
            for (OpType t : {OP_FILL, OP_SETV, OP_WALLF, OP_COPY}) {
                std::unique_ptr<MergedOp> m(new MergedOp());
                m->type = t;
                int nset = 0;
                std::vector<int> ent;
                std::vector<double> val;
                for (int gid = 0; gid < G_NGROUP; ++gid) {
                    if (!(gset & (1u << gid))) continue;
                    // a caller plan stands in for the halo (phase 4) and owner (phase 5) copies
                    if (caller_plan[gid] && phase >= kPhaseHalo) continue;
                    for (auto &o : prog[gid].ops) {
                        if (o->phase != phase || o->type != t) continue;
                        if (t == OP_COPY || t == OP_FILL) {
                            if (nset >= MERC_MAX_SETS) throw std::runtime_error("merged sync: too many slot sets");
                            m->S.n[nset] = (int)o->slots.size();
                            for (size_t c = 0; c < o->slots.size(); ++c) {
                                m->S.s[nset][c] = o->slots[c];
                                m->S.v[nset][c] = t == OP_FILL ? o->fillv[c] : 0.0;
                            }
                        }
                        // each op's entries are padded to whole wavefronts with
                        // entries that do nothing (dst < 0), so that the kernels can
                        // take the op's slot set as uniform over a wave
                        const size_t width = t == OP_COPY ? 3 : (t == OP_WALLF ? 4 + g : 2);
                        if (t == OP_COPY) {
                            for (size_t e = 0; e + 1 < o->hidx.size(); e += 2) ent.insert(ent.end(), {o->hidx[e], o->hidx[e + 1], nset});
                            while ((ent.size() / width) % kWave) ent.insert(ent.end(), {-1, -1, nset});
                        } else if (t == OP_FILL) {
                            for (int d : o->hidx) ent.insert(ent.end(), {d, nset});
                            while ((ent.size() / width) % kWave) ent.insert(ent.end(), {-1, nset});
                        } else if (t == OP_SETV) {
                            for (size_t e = 0; e < o->hidx.size(); ++e) {
                                ent.insert(ent.end(), {o->hidx[e], o->slots[0]});
                                val.push_back(o->hval[e]);
                            }
                            while ((ent.size() / width) % kWave) {
                                ent.insert(ent.end(), {-1, o->slots[0]});
                                val.push_back(0.0);
                            }
                        } else {
                            const size_t w = 2 + g;
                            for (size_t e = 0; e + w <= o->hidx.size(); e += w) {
                                ent.push_back(o->slots[0]);
                                ent.push_back(o->is_na);
                                ent.insert(ent.end(), o->hidx.begin() + e, o->hidx.begin() + e + w);
                            }
                            while ((ent.size() / width) % kWave) {
                                ent.push_back(o->slots[0]);
                                ent.push_back(o->is_na);
                                ent.insert(ent.end(), w, -1);
                            }
                        }
                        ++nset;
                    }
                }
                if (ent.empty()) continue;
                const int width = t == OP_COPY ? 3 : (t == OP_WALLF ? 4 + g : 2);
                m->n = (long long)ent.size() / width;
                m->ent.upload(ent);
                if (!val.empty()) m->val.upload(val);
                mp.ops.push_back(std::move(m));
            }
    }

    void run_merged(unsigned gset) { run_merged(gset, st); }
    void run_merged(unsigned gset, hipStream_t s) {
        auto it = mprog.find(gset);
        if (it == mprog.end()) throw std::runtime_error("merged sync program missing");
        const int g = F->blk[0].g;
        for (auto &m : it->second.ops) {
            switch (m->type) {
            case OP_COPY: launch_copy_multi(s, m->ent.p, m->n, Fd(), m->S); break;
            case OP_FILL: launch_fill_multi(s, m->ent.p, m->n, Fd(), m->S); break;
            case OP_SETV: launch_setv_multi(s, m->ent.p, m->val.p, m->n, Fd()); break;
            case OP_WALLF: launch_wall_multi(s, m->ent.p, m->n, g, Fd(), P); break;
            default: break;
            }
        }
        // caller plans: the halo stage of every group of the set, then the owner stage
        for (int t = 0; t < 2; ++t)
            for (int gid = 0; gid < G_NGROUP; ++gid)
                if ((gset & (1u << gid)) && caller_plan[gid] && plan_d[gid][t].n > 0) {
                    if (plan_staged[gid][t]) launch_copy_plan_staged(s, plan_d[gid][t].p, plan_d[gid][t].n / 5, Fd(), plan_tmp[gid][t].p);
                    else launch_copy_plan(s, plan_d[gid][t].p, plan_d[gid][t].n / 5, Fd());
                }
    }

    void sync(int gid) { run_merged(1u << gid); }

    // Unknowns of the implicit resistive correction (BuildImplicitResistiveEdgeDofMap_,
    // Mercury_ImplicitResistive.cpp:165-217; ImplicitResistiveEtaAtEdge_ :219-252;
    // MercuryResistivityShape_, MercurySolver_Tools.cpp:50-69): owned edges that own
    // their periodic class, |dl| > 1e-12, eta = inver_Rem * shape(|edge midpoint|) > 0.
    // Edges with alpha = 0 (identity row, zero right-hand side: they solve to 0)
    // and far-field boundary-layer edges (overwritten by the Eres far-field
    // handler before reaching a face) are left out: the solution on the other
    // edges and the updated B do not depend on them.
    void build_resist() {
        resist = P.resist_enabled != 0.0;
        if (!resist) return;
        if (F->has_external())
            throw std::runtime_error("the implicit resistive correction needs the context's own topology: no EXTERNAL faces");
        std::vector<int> q, ax, si, sj;
        std::vector<double> eta;
        const Field &Fl = *F;
        std::vector<char> alias_ax[3];  // non-owner members of the edge classes, per axis
        alias.alias_flags(Fl.Vtot, alias_ax);
        for (int b = 0; b < Fl.nb(); ++b) {
            const MercBlk &k = Fl.blk[b];
            const BlockGeom &G = *geom[b];
            std::vector<double> dl(Fl.vol(b)), al(Fl.vol(b)), etab(Fl.vol(b));
            for (int a = 0; a < 3; ++a) {
                const int loc = LOC_EX + a;
                Fl.download_block(SL_DL + a, b, dl.data());
                Fl.download_block(SL_ALPHA + a, b, al.data());
                std::fill(etab.begin(), etab.end(), 0.0);
                std::vector<char> ff(Fl.vol(b), 0);
                for (int f = 0; f < 6; ++f)
                    if (k.bc[f] == MBC_FARFIELD)
                        for_box(inner_slab(Fl, b, loc, face_dir(f)), [&](int i, int j, int kk) { ff[Fl.gidx(b, i, j, kk) - k.off] = 1; });
                int hi[3];
                Fl.owned_hi(b, loc, hi);
                for (int i = 0; i < hi[0]; ++i)
                    for (int j = 0; j < hi[1]; ++j)
                        for (int kk = 0; kk < hi[2]; ++kk) {
                            if (alias_ax[a][Fl.gidx(b, i, j, kk)]) continue;
                            const long long l = Fl.gidx(b, i, j, kk) - k.off;
                            if (std::fabs(dl[l]) <= 1.0e-12 || al[l] == 0.0 || ff[l]) continue;
                            const auto r0 = G.node(i, j, kk), r1 = G.node(i + (a == 0), j + (a == 1), kk + (a == 2));
                            const double xm = 0.5 * (r0[0] + r1[0]), ym = 0.5 * (r0[1] + r1[1]), zm = 0.5 * (r0[2] + r1[2]);
                            const double r = std::sqrt(xm * xm + ym * ym + zm * zm);
                            double sh;
                            if (P.resist_profile == 1.0) {  // exponential profile (no reference counterpart)
                                sh = std::exp(-(r - P.resist_r_inner) / P.resist_width);
                            } else {
                                const double shape = 0.5 * (std::tanh((r - P.resist_r_inner) / P.resist_width) -
                                                            std::tanh((r - P.resist_r_outer) / P.resist_width));
                                sh = (!std::isfinite(shape) || shape <= 1.0e-10) ? 0.0 : std::min(1.0, shape);
                            }
                            const double e = P.inver_Rem * sh;
                            if (!(e > 0.0)) continue;
                            q.push_back(Fl.gidx(b, i, j, kk));
                            ax.push_back(a);
                            si.push_back(k.sI);
                            sj.push_back(k.sJ);
                            eta.push_back(e);
                            etab[l] = e;
                        }
                F->upload_block(SL_ETAR + a, b, etab.data());
            }
        }
        // the CG applies the linear operator, whose faces and edges lie within two
        // points of an unknown (and of its periodic images): per block the bounding
        // box of the unknowns grown by 2, the whole axis where the block is periodic.
        // With interfaces, owner -> alias copies reach into neighbour blocks, so
        // every block takes its whole owned range.
        if (Fl.has_interface()) {
            make_range(Fl, 0, 1, R_res);
        } else {
            std::vector<int> lo(3 * Fl.nb(), 0), nn(3 * Fl.nb(), 0);
            std::vector<int> bmin(3 * Fl.nb(), 1 << 30), bmax(3 * Fl.nb(), -1);
            for (size_t e = 0; e < q.size(); ++e) {
                int b = 0;
                while (b + 1 < Fl.nb() && Fl.blk[b + 1].off <= q[e]) ++b;
                const MercBlk &k = Fl.blk[b];
                long long r = q[e] - k.off;
                const int ii = (int)(r / k.sI) - k.g, jj = (int)((r % k.sI) / k.sJ) - k.g, kq = (int)(r % k.sJ) - k.g;
                const int id[3] = {ii, jj, kq};
                for (int c = 0; c < 3; ++c) {
                    bmin[3 * b + c] = std::min(bmin[3 * b + c], id[c]);
                    bmax[3 * b + c] = std::max(bmax[3 * b + c], id[c]);
                }
            }
            for (int b = 0; b < Fl.nb(); ++b) {
                if (bmax[3 * b] < 0) continue;
                const MercBlk &k = Fl.blk[b];
                for (int c = 0; c < 3; ++c) {
                    int l = std::max(0, bmin[3 * b + c] - 2), h = std::min(k.m[c] + 1, bmax[3 * b + c] + 3);
                    if (k.bc[2 * c] == MBC_PERIODIC) l = 0, h = k.m[c] + 1;
                    lo[3 * b + c] = l;
                    nn[3 * b + c] = h - l;
                }
            }
            make_range_boxes(Fl, lo, nn, R_res);
        }
        const int n = (int)q.size();
        rd.n = n;
        r_q.upload(q);
        r_ax.upload(ax);
        r_sI.upload(si);
        r_sJ.upload(sj);
        r_eta.upload(eta);
        rd.q = r_q.p;
        rd.ax = r_ax.p;
        rd.sI = r_sI.p;
        rd.sJ = r_sJ.p;
        rd.eta = r_eta.p;
        r_vec.upload(std::vector<double>(9 * (size_t)std::max(n, 1), 0.0));
        double *v = r_vec.p;
        const size_t m = (size_t)std::max(n, 1);
        rx = v; rr = v + m; rp = v + 2 * m; rAp = v + 3 * m; rb = v + 4 * m; rdg = v + 5 * m; rw = v + 6 * m; rt = v + 7 * m;
        rwd = v + 8 * m;
        r_part.upload(std::vector<double>(3 * (size_t)std::max(res_blocks(n), 1), 0.0));
        r_S.upload(std::vector<double>(CG_NSCALAR, 0.0));
    }

    // MatMultImplicitResistive_ (Mercury_ImplicitResistive.cpp:406-465): Eres <- x on
    // the unknowns, Sync(Eres), dB = -dt * CurlEdgeToFace(Eres), Sync(dB), dJ =
    // alpha * CurlAdjFaceToEdge(beta * dB), alias mean, y = Eres - eta * dJ.  The
    // reference's Sync(dJ) touches no unknown and is left out.  lin: the far-field
    // Eres handler with a zero field (the linear part).  With w, the partial sums of
    // x.y.w go to r_part.  scattered: x is already in the Eres slots of the unknowns.
    void res_apply(const double *x, double *y, double dt, bool lin, const double *w, bool scattered = false) {
        if (!scattered) launch_res_scatter(st, rd, x, Fd());
        run_merged(1u << (lin ? G_ERES0 : G_ERES));
        launch_dec_d1(st, R_res.h, Bd(), Fd(), SL_ERES, SL_DB, -dt);
        run_merged(1u << G_DB);
        launch_jedge(st, R_res.h, Bd(), Fd(), SL_DB, SL_DJ);
        alias.launch(st, Fd(), SL_DJ);
        singular_current(SL_DB, SL_DJ);
        launch_res_gather(st, rd, Fd(), y, x, w, r_part.p);
    }

    // SolveImplicitResistiveDiffusion_ (Mercury_ImplicitResistive.cpp:467-547) and
    // ApplyImplicitResistiveUpdate_ (:549-575) on the synchronised state B*:
    // b = eta * J(B*); conjugate gradients on D^-1 A_lin x = D^-1 (b - A(0)),
    // D = diag(eta*alpha), with the Jacobi preconditioner (D^-1 A_lin = D^-1 +
    // dt C^T diag(beta) C is symmetric positive definite); stop when |b - A(x)|_2
    // <= max(rtol |b|_2, atol), tested every RES_CHECK iterations; the previous
    // solution is the initial guess.  Then Eres = Sync(x), B += dt * (-curl Eres)
    // on owned faces, Sync(Bface).  The reference runs PETSc GMRES without a
    // preconditioner on the same equation.
    void res_solve(double dt) {
        const int n = rd.n;
        // Calc_J_Edge on B* around the unknowns (its Sync(Jedge) touches no
        // unknown; the derived pass recomputes J everywhere afterwards)
        launch_jedge(st, R_res.h, Bd(), Fd());
        alias.launch(st, Fd(), SL_J);
        singular_current(SL_B, SL_J);
        launch_res_rhs(st, rd, Fd(), dt, rb, rdg, rw, rwd);
        res_apply(rx, rAp, dt, false, nullptr);
        launch_cg_init(st, n, rb, rAp, rdg, rwd, rr, rp, r_part.p, r_S.p);
        launch_res_scatter(st, rd, rp, Fd());
        int its = 0;
        double S[CG_NSCALAR];
        for (;;) {
            if (its % RES_CHECK == 0) {
                HIPCHK(hipMemcpyAsync(S, r_S.p, sizeof(S), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                const double rn = std::sqrt(S[CG_RR]), bn = std::sqrt(S[CG_BB]);
                res_rel = bn > 0.0 ? rn / bn : rn;
                if (!std::isfinite(rn)) throw std::runtime_error("implicit resistive CG: residual is not finite");
                if (!(rn > std::max(P.resist_rtol * bn, P.resist_atol)) || n == 0) break;
                if (its >= (int)P.resist_max_it)
                    throw std::runtime_error("implicit resistive CG did not converge within " + std::to_string(its) + " iterations");
            }
            // p is in the Eres slots already (scattered by the init or by the
            // previous iteration's k_cg_pxs)
            res_apply(rp, rAp, dt, true, rw, true);
            launch_cg_pap(st, n, r_part.p, r_S.p);
            launch_cg_r(st, n, r_S.p, rAp, rwd, rr, r_part.p, r_S.p);
            launch_cg_pxs(st, rd, r_S.p, rr, rdg, rp, rx, Fd());
            ++its;
        }
        res_its = its;
        launch_res_scatter(st, rd, rx, Fd());
        run_merged(1u << G_ERES);
        launch_dec_d1(st, R_node.h, Bd(), Fd(), SL_ERES, SL_RHSBR, -1.0);
        launch_bres_update(st, R_node.h, Bd(), Fd(), dt);
        run_merged(1u << G_BFACE);
    }

    const MercBlk *Bd() const { return F->dblk; }
    const double *dtp() const { return reinterpret_cast<const double *>(d_dt); }

    // UpdateDerivedFields_ (MercurySolver_Tools.cpp:303-321), fused passes; the
    // last derived pass of a step also reduces the next step's dt
    void derived_fused(bool final) {
        launch_bcell_jedge(st, R_node.h, Bd(), Fd(), Gd(), mix(), mon_part(), d_mon);
        alias.launch(st, Fd(), SL_J);
        singular_current(SL_B, SL_J);
        run_merged(MG_DERIVED_PAIR);
        launch_jcell_t(st, R_cell.h, Bd(), Fd(), Gd(), mix());
        sync(G_JCELL);
        launch_cell_end(st, R_box.h, Bd(), Fd(), Gd(), P, d_dt_next, final ? 1 : 0, mix(), nruns);
    }

    // After an update: Sync(U), Sync(Bface), the implicit resistive correction
    // when it is due, and the derived passes.  With overlap (and no correction)
    // the U sync runs on st2 beside the B sync and the B-cell / J-edge pass, and,
    // when no DOF is shared between blocks or across periodic seams (so no
    // owner -> alias copy writes an owned J edge), the B_cell / J-edge sync also
    // runs on st2 beside the J_cell pass; the last derived pass joins both.
    void after_update(bool final) {
        if (!overlap || res_now) {
            run_merged(MG_STATE_PAIR);
            if (res_now) res_solve(res_dt);
            derived_fused(final);
            return;
        }
        HIPCHK(hipEventRecord(ev_fork2, st));
        HIPCHK(hipStreamWaitEvent(st2, ev_fork2, 0));
        run_merged(1u << G_UCELL, st2);
        run_merged(1u << G_BFACE, st);
        launch_bcell_jedge(st, R_node.h, Bd(), Fd(), Gd(), mix(), mon_part(), d_mon);
        alias.launch(st, Fd(), SL_J);
        singular_current(SL_B, SL_J);
        if (alias.n == 0 && !F->connected() && !any_caller_plan()) {
            HIPCHK(hipEventRecord(ev_fork2, st));
            HIPCHK(hipStreamWaitEvent(st2, ev_fork2, 0));
            run_merged(MG_DERIVED_PAIR, st2);
        } else {
            run_merged(MG_DERIVED_PAIR, st);
        }
        launch_jcell_t(st, R_cell.h, Bd(), Fd(), Gd(), mix());
        sync(G_JCELL);
        HIPCHK(hipEventRecord(ev_join2, st2));
        HIPCHK(hipStreamWaitEvent(st, ev_join2, 0));
        launch_cell_end(st, R_box.h, Bd(), Fd(), Gd(), P, d_dt_next, final ? 1 : 0, mix(), nruns);
    }

    // StepOnce without the time step, fused passes
    void stage_fused(bool final) {
        // Scheme_U_ + sources (writes RHS_H/RHS_Na) and the induction EMF chain
        // (writes face EMFs, E, E_hall) only read what the previous stage left:
        // with overlap the RHS pass runs on st2, forked from and joined to st by
        // events (also under graph capture)
        if (overlap) {
            HIPCHK(hipEventRecord(ev_fork, st));
            HIPCHK(hipStreamWaitEvent(st2, ev_fork, 0));
            launch_flux(st2);
            HIPCHK(hipEventRecord(ev_join, st2));
        } else {
            launch_flux(st);
        }
        launch_eface_u(st, R_eface.h, Bd(), Fd(), Gd(), mix(), P.ambipolar_enabled != 0.0, nspecies());
        if (alias.n == 0) {
            launch_edge_total(st, R_node.h, Bd(), Fd(), Gd(), P, 1, mix());
            run_merged(MG_EDGE_PAIR);
            if (overlap) HIPCHK(hipStreamWaitEvent(st, ev_join, 0));
            launch_update_all(st, R_node.h, Bd(), Fd(), Gd(), dtp(), mix(), nspecies());
            after_update(final);
            return;
        } else {
            // shared edges (periodic seams, interfaces): the alias mean of the
            // non-Hall EMF before the Hall part is added; one pass writes both
            launch_edge_total(st, R_node.h, Bd(), Fd(), Gd(), P, 2, mix());
            alias.launch(st, Fd(), SL_E);
            singular_emf(mix());
            sync(G_EHALL);
            launch_add_hall_wall(st, R_node.h, Bd(), Fd());
        }
        sync(G_EEDGE);
        if (overlap) HIPCHK(hipStreamWaitEvent(st, ev_join, 0));
        launch_update_all(st, R_node.h, Bd(), Fd(), Gd(), dtp(), mix(), nspecies());
        after_update(final);
    }

    // UpdateDerivedFields_ (MercurySolver_Tools.cpp:303-321)
    void derived() {
        launch_bcell(st, R_cell.h, Bd(), Fd());
        sync(G_BCELL);
        launch_jedge(st, R_node.h, Bd(), Fd());
        alias.launch(st, Fd(), SL_J);
        singular_current(SL_B, SL_J);
        sync(G_JEDGE);
        launch_jcell(st, R_cell.h, Bd(), Fd());
        sync(G_JCELL);
        launch_pv_uplus(st, R_box.h, Bd(), Fd(), P);
    }

    // AssembleRHS_Induction_CT_ (MercurySolver_MagneticRHS.cpp:90-171) up to Sync(Eedge)
    void induction_emf() {
        const int es[3] = {SL_E, SL_E + 1, SL_E + 2};
        launch_zero_span(st, Fd(), es, 3, F->Vtot);
        launch_eface(st, R_eface.h, Bd(), Fd());
        launch_edge_emf(st, R_node.h, Bd(), Fd(), P);
        alias.launch(st, Fd(), SL_E);
        launch_hall_cell(st, R_box.h, Bd(), Fd(), P);
        launch_hall_edge(st, R_node.h, Bd(), Fd());
        singular_emf(false);
        sync(G_EHALL);
        launch_add_hall_wall(st, R_node.h, Bd(), Fd());
        sync(G_EEDGE);
    }

    // StepOnce (MercurySolver_Driver.cpp:21-61) without the time-step computation
    void stage() {
        launch_flux_div(st, R_cell.h, Bd(), Fd(), P);
        launch_sources(st, R_cell.h, Bd(), Fd(), P);
        induction_emf();
        launch_curl_update(st, R_node.h, Bd(), Fd(), dtp(), 1);
        launch_update_u(st, R_cell.h, Bd(), Fd(), dtp(), nspecies());
        run_merged(MG_STATE_PAIR);
        if (res_now) res_solve(res_dt);
        derived();
    }

    // With the implicit resistive correction the step reads dt back (the CG
    // solve is host driven); Euler: after the update and Sync(Bface) as in
    // StepOnce (MercurySolver_Driver.cpp:31-38); SSP-RK2: once, after the average.
    bool res_now = false;
    double res_dt = 0.0;

    void enqueue_step(int integrator) {
        // the fused forms reduced this step's dt at the end of the previous one;
        // the plain form reduces it here.  The commit applies the abort rule.
        if (form == 0) launch_timestep(st, R_cell.h, Bd(), Fd(), P, d_dt_next, nruns);
        launch_dt_commit(st, d_dt, d_dt_next, d_status, nruns);
        if (resist) {
            res_dt = read_dt();
            check_status();
        }
        res_now = resist && integrator == 0;
        if (form >= 1) {
            if (integrator == 0) {
                stage_fused(true);
            } else {
                launch_rk2(st, R_node.h, Bd(), Fd(), 1, nspecies());
                stage_fused(false);
                stage_fused(false);
                launch_rk2(st, R_node.h, Bd(), Fd(), 0, nspecies());
                res_now = resist;
                after_update(true);
            }
            res_now = false;
            return;
        }
        if (integrator == 0) {
            stage();
        } else {
            launch_rk2(st, R_node.h, Bd(), Fd(), 1, nspecies());
            stage();
            stage();
            launch_rk2(st, R_node.h, Bd(), Fd(), 0, nspecies());
            run_merged(MG_STATE_PAIR);
            if (resist) res_solve(res_dt);
            derived();
        }
        res_now = false;
    }

    double read_dt() {
        unsigned long long bits;
        HIPCHK(hipMemcpyAsync(&bits, d_dt, sizeof(bits), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        double dt;
        std::memcpy(&dt, &bits, sizeof(dt));
        return dt;
    }

    // Compute_Timestep abort rule (MercurySolver_Timestep.cpp:129-140), applied on
    // the device by k_dt_commit: after a stop the state is that of the last
    // completed step and every further step is refused until prepare().
    void check_status() {
        std::vector<unsigned long long> h(4 * (size_t)nruns);
        HIPCHK(hipMemcpyAsync(h.data(), d_status, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int r = 0; r < nruns; ++r) {
            const unsigned long long *w = &h[4 * (size_t)r];
            if (w[0] == 0ULL) continue;
            double bad;
            std::memcpy(&bad, &w[2], sizeof(bad));
            char msg[320];
            std::snprintf(msg, sizeof(msg),
                          "%sdt too small/NaN (dt underflow): step %llu since prepare refused dt = %.17g; the state is that of "
                          "step %llu and the context refuses further steps until prepare",
                          nruns > 1 ? ("run " + std::to_string(r) + ": ").c_str() : "", w[1] + 1ULL, bad, w[1]);
            throw std::runtime_error(msg);
        }
    }
    // the dt each run committed for its last step (0 for a stopped run)
    void run_dts(double *out) {
        std::vector<unsigned long long> bits(nruns);
        HIPCHK(hipMemcpyAsync(bits.data(), d_dt, sizeof(unsigned long long) * nruns, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::memcpy(out, bits.data(), sizeof(double) * nruns);
    }
    void clear_status() { HIPCHK(hipMemsetAsync(d_status, 0, 4 * sizeof(unsigned long long) * nruns, st)); }

    void step(int n, int integrator, double *dt_out) {
        for (int s = 0; s < n; ++s) {
            enqueue_step(integrator);
            if (dt_out) {
                dt_out[s] = read_dt();
                check_status();
            }
        }
        HIPCHK(hipGetLastError());
        if (n > 0) check_status();
    }

    void capture(int integrator) {
        if (resist) throw std::runtime_error("a step with the implicit resistive correction cannot be captured (its CG iteration count is data dependent)");
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        {
            CaptureGuard guard({st, st2});
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            guard.active = true;
            enqueue_step(integrator);
            guard.active = false;
            HIPCHK(hipStreamEndCapture(st, &graph));
        }
        HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        graph_integrator = integrator;
        graph_generation = generation;
    }
    unsigned long long graph_generation = 0;
    // n replays of the captured step, then the abort rule's status (one
    // synchronisation per call)
    void graph_replay(int n) {
        if (!gexec) throw std::runtime_error("no captured step graph");
        if (graph_generation != generation) throw std::runtime_error("the captured step graph is stale (form or launch shape changed)");
        for (int s = 0; s < n; ++s) HIPCHK(hipGraphLaunch(gexec, st));
    }
    void graph_launch(int n) {
        graph_replay(n);
        if (n > 0) check_status();
    }

    void init_bface(const double Bu[3], const double *const *A, int ndip = 0, const double *dip = nullptr) {
        if (ndip < 0 || (ndip > 0 && !dip)) throw std::runtime_error("init_bface: invalid dipole list");
        if (ndip > 0 && F->nb() > 0 && geom[0]->g < 1) throw std::runtime_error("init_bface: dipoles need at least one ghost layer");
        for (int b = 0; b < F->nb(); ++b) {
            const BlockGeom &G = *geom[b];
            const long long V = G.V;
            const double *Ax = A ? A[3 * b] : nullptr, *Ay = A ? A[3 * b + 1] : nullptr, *Az = A ? A[3 * b + 2] : nullptr;
            auto AA = [&](const double *p, int i, int j, int k) { return p ? p[G.idx(i, j, k)] : 0.0; };
            std::array<std::vector<double>, 3> Badd;
            for (int a = 0; a < 3; ++a) {
                std::vector<double> Bf(V, 0.0), &Ba = Badd[a];
                Ba.assign(V, 0.0);
                int lo[3], hi[3];
                for (int c = 0; c < 3; ++c) {
                    lo[c] = -G.g;
                    hi[c] = G.m[c] + 1 - DELTA[LOC_FX + a][c] + G.g;
                }
                for (int i = lo[0]; i < hi[0]; ++i)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int k = lo[2]; k < hi[2]; ++k) {
                            const long long q = G.idx(i, j, k);
                            const double j0 = G.JD_[3 * a][q], j1 = G.JD_[3 * a + 1][q], j2 = G.JD_[3 * a + 2][q];
                            double c;
                            if (a == 0) c = (AA(Ay, i, j, k) - AA(Ay, i, j, k + 1)) + (AA(Az, i, j + 1, k) - AA(Az, i, j, k));
                            else if (a == 1) c = (AA(Ax, i, j, k + 1) - AA(Ax, i, j, k)) + (AA(Az, i, j, k) - AA(Az, i + 1, j, k));
                            else c = (AA(Ax, i, j, k) - AA(Ax, i, j + 1, k)) + (AA(Ay, i + 1, j, k) - AA(Ay, i, j, k));
                            Bf[q] = (Bu[0] * j0 + Bu[1] * j1 + Bu[2] * j2) + c;
                            Ba[q] = Pb(b).B_imf[0] * j0 + Pb(b).B_imf[1] * j1 + Pb(b).B_imf[2] * j2;
                        }
                F->upload_block(SL_B + a, b, Bf.data());
            }
            // the dipoles' part of Badd on the face boxes of ghost depth 1
            // (DipoleField::Build_Badd_FaceFlux, Z4_Mercury/include/operators/Dipole.h:89-374)
            G.add_dipole_flux(ndip, dip, Badd);
            for (int a = 0; a < 3; ++a) F->upload_block(SL_BADD + a, b, Badd[a].data());
        }
    }

    // MercurySolver ctor tail + Advance head (MercurySolver.cpp:141-145, MercurySolver_Driver.cpp:6-19)
    // the derived fields (and the dt candidate) of the current form are a
    // function of U and B only: a fused form recomputes the whole derived chain
    // with its own passes, so a state read from a restart continues exactly as
    // the run that wrote it
    // Syncs Badd and Bface and runs UpdateDerivedFields_ once, as the
    // reference's set-up does; the cell-end pass adds the fused forms'
    // intermediates and the first dt.  A sync whose halo reads ghosts of its own
    // level (the node-index rule of patch connections) is not idempotent, so
    // neither this nor set_form syncs anything twice.
    bool prepared = false;
    void prepare() {
        clear_status();
        run_merged(MG_PREPARE_PAIR);
        if (form == 2) {
            // the mixed form's own derived fields, as its steps leave them
            refresh_float_statics();
            derived_fused(true);
        } else {
            derived();
            refresh_float_statics();
            launch_cell_end(st, R_box.h, Bd(), Fd(), Gd(), P, d_dt_next, 1, false, nruns);
        }
        prepared = true;
        HIPCHK(hipStreamSynchronize(st));
    }

    void set_form(int f) {
        if (f < 0 || f > 2) throw std::runtime_error("form must be 0 (per-function FP64), 1 (fused FP64) or 2 (mixed)");
        // the same form, or a context not prepared yet: nothing to recompute
        // (a context of patch connections under the node-index rule should get
        // its form before prepare: switching after it syncs the derived fields
        // a second time)
        if (f == form) return;
        form = f;
        ++generation;
        if (f >= 1 && prepared) {
            refresh_float_statics();
            derived_fused(true);
        }
        if (gexec) {
            (void)hipGraphExecDestroy(gexec);
            (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
        }
    }

    void run_op(const std::string &op, double arg) {
        if (op == "pv_uplus") launch_pv_uplus(st, R_box.h, Bd(), Fd(), P);
        else if (op == "bcell") launch_bcell(st, R_cell.h, Bd(), Fd());
        else if (op == "jedge") launch_jedge(st, R_node.h, Bd(), Fd());
        else if (op == "jcell") launch_jcell(st, R_cell.h, Bd(), Fd());
        else if (op == "derived") derived();
        else if (op == "flux") launch_flux_div(st, R_cell.h, Bd(), Fd(), P);
        else if (op == "sources") launch_sources(st, R_cell.h, Bd(), Fd(), P);
        else if (op == "eface") launch_eface(st, R_eface.h, Bd(), Fd());
        else if (op == "edge_emf") launch_edge_emf(st, R_node.h, Bd(), Fd(), P);
        else if (op == "hall_cell") launch_hall_cell(st, R_box.h, Bd(), Fd(), P);
        else if (op == "hall_edge") launch_hall_edge(st, R_node.h, Bd(), Fd());
        else if (op == "induction") {
            induction_emf();
            launch_curl_update(st, R_node.h, Bd(), Fd(), dtp(), 0);
        } else if (op == "alias_e")
            alias.launch(st, Fd(), SL_E);
        else if (op == "singular_emf")
            singular_emf(false);
        else if (op == "singular_j")
            singular_current(SL_B, SL_J);
        else if (op == "alias_j")
            alias.launch(st, Fd(), SL_J);
        else if (op == "timestep") launch_timestep(st, R_cell.h, Bd(), Fd(), P, d_dt, nruns);
        else if (op == "stage") stage();
        else if (op == "rhs_cell") launch_flux(st);
        else if (op == "cell_end") launch_cell_end(st, R_box.h, Bd(), Fd(), Gd(), P, d_dt_next, 1, mix(), nruns);
        else if (op == "eface_u") launch_eface_u(st, R_eface.h, Bd(), Fd(), Gd(), mix(), P.ambipolar_enabled != 0.0, nspecies());
        else if (op == "edge_total") launch_edge_total(st, R_node.h, Bd(), Fd(), Gd(), P, alias.n == 0 ? 1 : 0, mix());
        else if (op == "update_all") launch_update_all(st, R_node.h, Bd(), Fd(), Gd(), dtp(), mix(), nspecies());
        else if (op == "bcell_jedge") launch_bcell_jedge(st, R_node.h, Bd(), Fd(), Gd(), mix());
        else if (op == "bcell_jedge_monitor") launch_bcell_jedge(st, R_node.h, Bd(), Fd(), Gd(), mix(), d_mpart, d_mon);
        else if (op == "jcell_t") launch_jcell_t(st, R_cell.h, Bd(), Fd(), Gd(), mix());
        else if (op == "sync_all") {
            for (int gid = 0; gid < G_NGROUP; ++gid) sync(gid);
        }
        else if (op == "dec_d0") launch_dec_d0(st, R_node.h, Bd(), Fd(), SL_NODEF, SL_EDGEF);
        else if (op == "dec_d1") launch_dec_d1(st, R_node.h, Bd(), Fd(), SL_EDGEF, SL_FACEF, arg == 0.0 ? 1.0 : arg);
        else if (op == "res_matvec" || op == "res_matvec_lin") {
            // the resistive operator with time step arg on the generic edge 1-form:
            // x taken on the unknowns, y written back there (0 elsewhere)
            if (!resist) throw std::runtime_error("the implicit resistive correction is off");
            launch_res_pick(st, rd, Fd(), SL_EDGEF, rt);
            res_apply(rt, rb, arg, op == "res_matvec_lin", nullptr);
            const int es[3] = {SL_EDGEF, SL_EDGEF + 1, SL_EDGEF + 2};
            launch_zero_span(st, Fd(), es, 3, F->Vtot);
            launch_res_scatter(st, rd, rb, Fd(), SL_EDGEF);
        } else if (op == "res_solve") {
            if (!resist) throw std::runtime_error("the implicit resistive correction is off");
            res_solve(arg);
        } else if (op == "sync_eres") sync(G_ERES);
        else if (op == "sync_db") sync(G_DB);
        else if (op == "dec_d2") launch_dec_d2(st, R_cell.h, Bd(), Fd(), SL_FACEF, SL_CELLS);
        else if (op == "sync_ucell") sync(G_UCELL);
        else if (op == "sync_bface") sync(G_BFACE);
        else if (op == "sync_eedge") sync(G_EEDGE);
        else if (op == "sync_ehall") sync(G_EHALL);
        else if (op == "sync_jedge") sync(G_JEDGE);
        else if (op == "sync_bcell") sync(G_BCELL);
        else if (op == "sync_jcell") sync(G_JCELL);
        else if (op == "sync_badd") sync(G_BADD);
        else if (op == "divb") launch_divb(st, R_node.h, Bd(), Fd(), d_part, d_red, 1);  // calc_divB on the device, no read-back
        else throw std::runtime_error("unknown op: " + op);
        HIPCHK(hipGetLastError());
    }

    // the monitor: last normalised div B of the fused derived pass and its
    // maximum since the last reset (optionally reset now)
    void divb_monitor(double *last, double *maxn, double *last_abs, bool reset) {
        double h[3];
        HIPCHK(hipMemcpyAsync(h, d_mon, sizeof(h), hipMemcpyDeviceToHost, st));
        if (reset) launch_divb_monitor_reset(st, d_mon);
        HIPCHK(hipStreamSynchronize(st));
        if (last) *last = h[0];
        if (maxn) *maxn = h[1];
        if (last_abs) *last_abs = h[2];
    }
    void divb(double *norm, double *maxabs) {
        launch_divb(st, R_node.h, Bd(), Fd(), d_part, d_red, 1);
        unsigned long long h[2];
        HIPCHK(hipMemcpyAsync(h, d_red, sizeof(h), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        double dm, bm;
        std::memcpy(&dm, &h[0], 8);
        std::memcpy(&bm, &h[1], 8);
        if (maxabs) *maxabs = dm;
        if (norm) *norm = bm > 0.0 ? dm / bm : dm;
    }

    // average device time of one op over reps launches, HIP events on the solver stream
    double time_op(const std::string &op, int reps) {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        if (op == "step_graph") graph_replay(1);
        else run_op(op, 0.0);
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) {
            if (op == "step_graph") graph_replay(1);
            else run_op(op, 0.0);
        }
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return (double)ms / reps;
    }
};

}  // namespace merc

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct MercuryContext {
    std::unique_ptr<merc::MercurySolver> s;
};

#define ABI_TRY(...)                                   \
    try {                                              \
        __VA_ARGS__;                                   \
        return 0;                                      \
    } catch (const std::exception &e) {                \
        merc::g_last_error = e.what();                 \
        return -1;                                     \
    } catch (...) {                                    \
        merc::g_last_error = "unknown error";          \
        return -1;                                     \
    }

#define CHECK_CTX(c)                                                   \
    if (!(c) || !(c)->s) {                                             \
        merc::g_last_error = "null or destroyed context";              \
        return -1;                                                     \
    }

extern "C" {

const char *mercury_last_error(void) { return merc::g_last_error.c_str(); }

int mercury_version(void) { return MERCURY_HIP_ABI_VERSION; }

int mercury_create_connected(MercuryContext **out, int nblocks, const int *dims, int nghost, const int *bc, const int *nbr,
                             const int *transform, const double *owned_nodes, const MercuryPhysics *phys) {
    if (!out || nblocks <= 0 || !dims || !bc || !owned_nodes || !phys) {
        merc::g_last_error = "mercury_create: invalid argument";
        return -1;
    }
    ABI_TRY({
        MercPhys P;
        std::memcpy(&P, phys, sizeof(P));
        std::unique_ptr<MercuryContext> c(new MercuryContext());
        c->s.reset(new merc::MercurySolver(nblocks, dims, nghost, bc, nbr, transform, owned_nodes, P));
        *out = c.release();
    })
}

int mercury_create_patched(MercuryContext **out, int nblocks, const int *dims, int nghost, const int *bc, int npatch, const int *patches,
                           int rule, const double *owned_nodes, const MercuryPhysics *phys) {
    if (!out || nblocks <= 0 || !dims || !bc || !owned_nodes || !phys) {
        merc::g_last_error = "mercury_create_patched: invalid argument";
        return -1;
    }
    ABI_TRY({
        const merc::patch::Topo T = merc::patch::make_topo(nblocks, dims, nghost, bc, npatch, patches, rule);
        MercPhys P;
        std::memcpy(&P, phys, sizeof(P));
        std::unique_ptr<MercuryContext> c(new MercuryContext());
        c->s.reset(new merc::MercurySolver(nblocks, dims, nghost, bc, nullptr, nullptr, owned_nodes, 1, nullptr, &P, &T));
        *out = c.release();
    })
}

int mercury_patch_plan(int nblocks, const int *dims, int nghost, const int *bc, int npatch, const int *patches, int rule, int kind,
                       int level, int stage, long long capacity, long long *n, int *entries) {
    if (!n || (stage != 0 && stage != 1)) {
        merc::g_last_error = "mercury_patch_plan: invalid argument";
        return -1;
    }
    ABI_TRY({
        const merc::patch::Topo T = merc::patch::make_topo(nblocks, dims, nghost, bc, npatch, patches, rule);
        const std::vector<merc::patch::Entry> e = stage == 0 ? merc::patch::halo_entries(T, kind, level) : merc::patch::owner_entries(T, kind);
        *n = (long long)e.size();
        if (!entries || capacity < *n) return 0;
        for (size_t t = 0; t < e.size(); ++t) {
            int *o = entries + 11 * t;
            const merc::patch::Dof &d = e[t].dst, &s = e[t].src;
            const int v[11] = {d.f, d.b, d.i, d.j, d.k, s.f, s.b, s.i, s.j, s.k, e[t].sign};
            std::copy(v, v + 11, o);
        }
    })
}

int mercury_singular_edges(int nblocks, const int *dims, int nghost, const int *bc, int npatch, const int *patches, int rule,
                           const double *owned_nodes, const MercuryPhysics *phys, int capacity, int *n, int *ints,
                           double *values) {
    if (!n || !dims || !owned_nodes || !phys || nblocks <= 0) {
        merc::g_last_error = "mercury_singular_edges: invalid argument";
        return -1;
    }
    ABI_TRY({
        const merc::patch::Topo T = merc::patch::make_topo(nblocks, dims, nghost, bc, npatch, patches, rule);
        T.validate();
        MercPhys P;
        std::memcpy(&P, phys, sizeof(MercPhys));
        // the static geometry each block's context computes, kept on the host
        std::vector<std::vector<double>> store((size_t)nblocks * 19);
        std::vector<merc::singular::BlockGeo> G(nblocks);
        long long noff = 0;
        for (int b = 0; b < nblocks; ++b) {
            const int m[3] = {dims[3 * b], dims[3 * b + 1], dims[3 * b + 2]};
            merc::BlockGeom geo(m, nghost, owned_nodes + noff);
            noff += 3LL * (m[0] + 1) * (m[1] + 1) * (m[2] + 1);
            struct W : merc::SlotWriter {
                std::vector<double> *out;
                void put(int slot, const std::vector<double> &v) override {
                    int u = -1;
                    if (slot >= SL_DL && slot < SL_DL + 3) u = slot - SL_DL;
                    else if (slot >= SL_DR && slot < SL_DR + 9) u = 3 + slot - SL_DR;
                    else if (slot == SL_JAC) u = 12;
                    else if (slot >= SL_CC && slot < SL_CC + 3) u = 13 + slot - SL_CC;
                    else if (slot >= SL_AREA && slot < SL_AREA + 3) u = 16 + slot - SL_AREA;
                    if (u >= 0) out[u] = v;
                }
            } w;
            w.out = &store[(size_t)b * 19];
            geo.compute(P, w);
            merc::singular::set_block_geo(G[b], T.m[b], nghost, &store[(size_t)b * 19]);
        }
        const merc::singular::Registry R = merc::singular::build(T, G);
        *n = (int)R.e.size();
        if (!ints || !values || capacity < *n) return 0;
        for (size_t t = 0; t < R.e.size(); ++t) {
            const merc::singular::Owner &o = R.e[t];
            const int v[8] = {o.dof.b, o.dof.i, o.dof.j, o.dof.k, o.dof.f, o.sign, o.c1 - o.c0, o.f1 - o.f0};
            std::copy(v, v + 8, ints + 8 * t);
            const double d[4] = {o.ih, o.cv[0], o.cv[1], o.cv[2]};
            std::copy(d, d + 4, values + 4 * t);
        }
    })
}

int mercury_create(MercuryContext **out, int nblocks, const int *dims, int nghost, const int *bc, const double *owned_nodes,
                   const MercuryPhysics *phys) {
    return mercury_create_connected(out, nblocks, dims, nghost, bc, nullptr, nullptr, owned_nodes, phys);
}

int mercury_create_ensemble(MercuryContext **out, int nruns, const int *run_nblocks, const int *dims, int nghost, const int *bc,
                            const int *nbr, const int *transform, const double *owned_nodes, const MercuryPhysics *phys) {
    if (!out || nruns <= 0 || !run_nblocks || !dims || !bc || !owned_nodes || !phys) {
        merc::g_last_error = "mercury_create_ensemble: invalid argument";
        return -1;
    }
    ABI_TRY({
        std::vector<int> block_run;
        for (int r = 0; r < nruns; ++r) {
            if (run_nblocks[r] <= 0) throw std::runtime_error("mercury_create_ensemble: every run needs at least one block");
            block_run.insert(block_run.end(), run_nblocks[r], r);
        }
        std::vector<MercPhys> P(nruns);
        std::memcpy(P.data(), phys, sizeof(MercPhys) * nruns);
        std::unique_ptr<MercuryContext> c(new MercuryContext());
        c->s.reset(new merc::MercurySolver((int)block_run.size(), dims, nghost, bc, nbr, transform, owned_nodes, nruns, block_run.data(),
                                           P.data()));
        *out = c.release();
    })
}

int mercury_run_dt(MercuryContext *ctx, int *nruns, double *dts) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (nruns) *nruns = ctx->s->nruns;
        if (dts) ctx->s->run_dts(dts);
    })
}

int mercury_destroy(MercuryContext *ctx) {
    ABI_TRY({ delete ctx; })
}

int mercury_field_info(const char *name, int *slot, int *ncomp, int *location) {
    const merc::FieldSpec *f = name ? merc::find_field(name) : nullptr;
    if (!f) {
        merc::g_last_error = std::string("unknown field: ") + (name ? name : "(null)");
        return -1;
    }
    if (slot) *slot = f->slot;
    if (ncomp) *ncomp = f->ncomp;
    if (location) *location = f->loc;
    return 0;
}

long long mercury_block_volume(MercuryContext *ctx, int block) {
    if (!ctx || !ctx->s || block < 0 || block >= ctx->s->F->nb()) return -1;
    return ctx->s->F->vol(block);
}

static int field_comp_slot(const char *name, int comp) {
    const merc::FieldSpec *f = name ? merc::find_field(name) : nullptr;
    if (!f) throw std::runtime_error(std::string("unknown field: ") + (name ? name : "(null)"));
    if (comp < 0 || comp >= f->ncomp) throw std::runtime_error("component out of range");
    return f->slot + comp;
}

int mercury_block_shape(MercuryContext *ctx, int block, int *dims, int *nghost) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (block < 0 || block >= ctx->s->F->nb()) throw std::runtime_error("invalid block");
        const MercBlk &k = ctx->s->F->blk[block];
        if (dims)
            for (int a = 0; a < 3; ++a) dims[a] = k.m[a];
        if (nghost) *nghost = k.g;
    })
}

int mercury_field_device(MercuryContext *ctx, const char *name, int block, int comp, double **ptr) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (block < 0 || block >= ctx->s->F->nb() || !ptr) throw std::runtime_error("invalid block or output pointer");
        const int slot = field_comp_slot(name, comp);
        HIPCHK(hipStreamSynchronize(ctx->s->st));
        ctx->s->materialize(slot);
        *ptr = ctx->s->F->hptr[slot] + ctx->s->F->blk[block].off;
    })
}

int mercury_stream(MercuryContext *ctx, void **stream) {
    CHECK_CTX(ctx);
    if (!stream) {
        merc::g_last_error = "mercury_stream: stream is NULL";
        return -1;
    }
    *stream = (void *)ctx->s->st;
    return 0;
}

// Sync groups by the names of MercuryBoundary::Sync (Boundary_Build.cpp:564-703),
// in GroupId order; "Eres0" is Eres with a zero far field, "dB" the resistive dB
static int sync_group_id(const char *name) {
    static const char *const names[] = {"Ucell", "Bface", "Eedge", "Ehall", "Jedge", "B_cell", "J_cell", "Badd", "Eres", "Eres0", "dB"};
    static_assert(sizeof(names) / sizeof(names[0]) == merc::G_NGROUP, "one name per sync group");
    if (name)
        for (int g = 0; g < merc::G_NGROUP; ++g)
            if (!std::strcmp(name, names[g])) return g;
    throw std::runtime_error(std::string("unknown sync group ") + (name ? name : "(null)"));
}

int mercury_set_sync_plan(MercuryContext *ctx, const char *group, int stage, long long n, const int *entries) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (n < 0 || (n > 0 && !entries)) throw std::runtime_error("mercury_set_sync_plan: invalid entries");
        merc::MercurySolver &S = *ctx->s;
        const int gid = sync_group_id(group);
        std::vector<int> ent;
        ent.reserve(5 * (size_t)n);
        for (long long e = 0; e < n; ++e) {
            const int *x = entries + 11 * e;
            if (x[0] < 0 || x[0] >= SL_NSLOT || x[5] < 0 || x[5] >= SL_NSLOT)
                throw std::runtime_error("mercury_set_sync_plan: slot out of range in entry " + std::to_string(e));
            if (x[10] != 1 && x[10] != -1) throw std::runtime_error("mercury_set_sync_plan: sign must be +1 or -1");
            const int d = S.F->gidx_checked(x[1], x[2], x[3], x[4], "mercury_set_sync_plan destination");
            const int src = S.F->gidx_checked(x[6], x[7], x[8], x[9], "mercury_set_sync_plan source");
            ent.insert(ent.end(), {x[0], d, x[5], src, x[10]});
        }
        HIPCHK(hipStreamSynchronize(S.st));
        S.set_sync_plan(gid, stage, std::move(ent));
    })
}

int mercury_sync_plan(MercuryContext *ctx, const char *group, int stage, long long capacity, long long *n, int *entries) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (!n || stage < 0 || stage > 1) throw std::runtime_error("mercury_sync_plan: invalid argument");
        merc::MercurySolver &S = *ctx->s;
        const std::vector<int> ent = S.plan_entries(sync_group_id(group), stage);
        *n = (long long)ent.size() / 5;
        if (!entries || capacity < *n) return 0;
        for (long long e = 0; e < *n; ++e) {
            const int *x = ent.data() + 5 * e;
            int d[4], src[4];
            S.F->point_of(x[1], d);
            S.F->point_of(x[3], src);
            int *y = entries + 11 * e;
            y[0] = x[0];
            for (int c = 0; c < 4; ++c) y[1 + c] = d[c];
            y[5] = x[2];
            for (int c = 0; c < 4; ++c) y[6 + c] = src[c];
            y[10] = x[4];
        }
    })
}

int mercury_set_alias_classes(MercuryContext *ctx, int ncls, const int *start, const int *axis, const int *members, const int *owners) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (ncls < 0 || (ncls > 0 && (!start || !axis || !members || !owners)))
            throw std::runtime_error("mercury_set_alias_classes: invalid argument");
        merc::MercurySolver &S = *ctx->s;
        std::vector<int> st{0}, mem, ms, max, ow, os, oax;
        if (ncls > 0 && start[0] != 0) throw std::runtime_error("mercury_set_alias_classes: start[0] must be 0");
        for (int c = 0; c < ncls; ++c) {
            if (start[c + 1] <= start[c]) throw std::runtime_error("mercury_set_alias_classes: empty class " + std::to_string(c));
            if (axis[c] < 0 || axis[c] > 2) throw std::runtime_error("mercury_set_alias_classes: axis must be 0, 1 or 2");
            for (int t = start[c]; t < start[c + 1]; ++t) {
                const int *m = members + 5 * (long long)t;
                if (m[4] != 1 && m[4] != -1) throw std::runtime_error("mercury_set_alias_classes: sign must be +1 or -1");
                mem.push_back(S.F->gidx_checked(m[0], m[1], m[2], m[3], "mercury_set_alias_classes member"));
                ms.push_back(m[4]);
                max.push_back(axis[c]);
            }
            st.push_back((int)mem.size());
            const int *o = owners + 5 * (long long)c;
            if (o[4] != 1 && o[4] != -1) throw std::runtime_error("mercury_set_alias_classes: sign must be +1 or -1");
            ow.push_back(S.F->gidx_checked(o[0], o[1], o[2], o[3], "mercury_set_alias_classes owner"));
            os.push_back(o[4]);
            oax.push_back(axis[c]);
        }
        HIPCHK(hipStreamSynchronize(S.st));
        S.alias.set(std::move(st), std::move(mem), std::move(ms), std::move(max), std::move(ow), std::move(os), std::move(oax));
        S.drop_graph();
    })
}

int mercury_alias_classes(MercuryContext *ctx, int cls_capacity, long long member_capacity, int *ncls, long long *nmembers, int *start,
                          int *axis, int *members, int *owners) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (!ncls || !nmembers) throw std::runtime_error("mercury_alias_classes: invalid argument");
        const merc::AliasClasses &A = ctx->s->alias;
        *ncls = (int)A.how.size();
        *nmembers = (long long)A.hmem.size();
        if (!start || !axis || !members || !owners || cls_capacity < *ncls || member_capacity < *nmembers) return 0;
        for (int c = 0; c <= *ncls; ++c) start[c] = A.hst[c];
        for (int c = 0; c < *ncls; ++c)
            for (int t = A.hst[c]; t < A.hst[c + 1]; ++t)
                if (A.hmax[t] != A.hoax[c])
                    throw std::runtime_error("mercury_alias_classes: class " + std::to_string(c) +
                                             " joins edges of different axes (a patch-connected context); one axis per class is exported");
        for (int c = 0; c < *ncls; ++c) {
            int p[4];
            axis[c] = A.hoax[c];
            ctx->s->F->point_of(A.how[c], p);
            for (int q = 0; q < 4; ++q) owners[5 * c + q] = p[q];
            owners[5 * c + 4] = A.hos[c];
        }
        for (long long t = 0; t < *nmembers; ++t) {
            int p[4];
            ctx->s->F->point_of(A.hmem[t], p);
            for (int q = 0; q < 4; ++q) members[5 * t + q] = p[q];
            members[5 * t + 4] = A.hms[t];
        }
    })
}

int mercury_set_stream(MercuryContext *ctx, void *stream) {
    CHECK_CTX(ctx);
    ABI_TRY({ ctx->s->set_stream((hipStream_t)stream); })
}

int mercury_upload(MercuryContext *ctx, const char *name, int block, int comp, const double *host) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (block < 0 || block >= ctx->s->F->nb() || !host) throw std::runtime_error("invalid block or buffer");
        HIPCHK(hipStreamSynchronize(ctx->s->st));
        ctx->s->F->upload_block(field_comp_slot(name, comp), block, host);
    })
}

int mercury_download(MercuryContext *ctx, const char *name, int block, int comp, double *host) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (block < 0 || block >= ctx->s->F->nb() || !host) throw std::runtime_error("invalid block or buffer");
        HIPCHK(hipStreamSynchronize(ctx->s->st));
        const int slot = field_comp_slot(name, comp);
        ctx->s->materialize(slot);
        ctx->s->F->download_block(slot, block, host);
    })
}

// ---------------------------------------------------------------------------
// restart files (MPCNSRST v1) and run data (MPCNSRUN v1) on device-resident fields
// ---------------------------------------------------------------------------
extern "C++" {
namespace {
// the reference's restart whitelist (Z4_Mercury/src/MercurySolver.cpp:31-44)
const char *const kRestartFields[] = {"U_H", "U_Na", "B_xi", "B_eta", "B_zeta"};
const char *const kJedgeFields[] = {"J_xi", "J_eta", "J_zeta"};

// allocated box of a location (FieldBlock get_lo / get_hi): [-g, m + 1 - delta + g)
std::array<int, 6> alloc_box(const MercBlk &k, int loc) {
    std::array<int, 6> b;
    for (int a = 0; a < 3; ++a) {
        b[a] = -k.g;
        b[3 + a] = k.m[a] + 1 - merc::DELTA[loc][a] + k.g;
    }
    return b;
}
}  // namespace
}

int mercury_write_restart(MercuryContext *ctx, const char *path, int step, double time, int with_jedge) {
    CHECK_CTX(ctx);
    if (!path) {
        merc::g_last_error = "mercury_write_restart: path is null";
        return -1;
    }
    ABI_TRY({
        merc::MercurySolver &S = *ctx->s;
        HIPCHK(hipStreamSynchronize(S.st));
        const merc::Field &F = *S.F;
        std::vector<std::string> names(std::begin(kRestartFields), std::end(kRestartFields));
        // a three-species run adds its He++ state after the whitelist (the reader
        // takes any field it knows by name)
        if (S.nspecies() > 2) names.insert(names.begin() + 2, "U_He");
        if (with_jedge) names.insert(names.end(), std::begin(kJedgeFields), std::end(kJedgeFields));
        std::vector<merc::io::RestartField> out;
        for (const std::string &nm : names) {
            const merc::FieldSpec *f = merc::find_field(nm.c_str());
            merc::io::RestartField rf;
            rf.meta.name = nm;
            rf.meta.loc = f->loc;
            rf.meta.ncomp = f->ncomp;
            rf.meta.nghost = F.blk[0].g;
            std::vector<std::vector<double>> comp(f->ncomp);
            for (int b = 0; b < F.nb(); ++b) {
                const MercBlk &k = F.blk[b];
                const auto box = alloc_box(k, f->loc);
                for (int c = 0; c < f->ncomp; ++c) {
                    S.materialize(f->slot + c);
                    comp[c].resize(F.vol(b));
                    F.download_block(f->slot + c, b, comp[c].data());
                }
                std::vector<double> v;
                v.reserve((size_t)(box[3] - box[0]) * (box[4] - box[1]) * (box[5] - box[2]) * f->ncomp);
                for (int i = box[0]; i < box[3]; ++i)
                    for (int j = box[1]; j < box[4]; ++j)
                        for (int kk = box[2]; kk < box[5]; ++kk) {
                            const long long q = (long long)(i + k.g) * k.sI + (long long)(j + k.g) * k.sJ + (kk + k.g);
                            for (int c = 0; c < f->ncomp; ++c) v.push_back(comp[c][q]);
                        }
                rf.box.push_back(box);
                rf.data.push_back(std::move(v));
            }
            out.push_back(std::move(rf));
        }
        merc::io::write_restart(path, step, time, F.nb(), out);
    })
}

int mercury_read_restart(MercuryContext *ctx, const char *path, int *step, double *time) {
    CHECK_CTX(ctx);
    if (!path) {
        merc::g_last_error = "mercury_read_restart: path is null";
        return -1;
    }
    ABI_TRY({
        merc::MercurySolver &S = *ctx->s;
        HIPCHK(hipStreamSynchronize(S.st));
        merc::Field &F = *S.F;
        std::map<std::string, std::vector<char>> delivered;
        int njedge = 0;
        auto want = [&](const merc::io::FieldMeta &m) {
            const merc::FieldSpec *f = merc::find_field(m.name.c_str());
            if (!f) return false;  // fields this build does not hold are stepped over
            if (m.loc != f->loc) throw std::runtime_error("restart: location mismatch for " + m.name);
            if (m.ncomp != f->ncomp) throw std::runtime_error("restart: ncomp mismatch for " + m.name);
            if (m.nghost != F.blk[0].g) throw std::runtime_error("restart: nghost mismatch for " + m.name);
            if (delivered.count(m.name)) throw std::runtime_error("restart: duplicate field " + m.name);
            for (const char *j : kJedgeFields)
                if (m.name == j) ++njedge;
            delivered[m.name].assign(F.nb(), 0);
            return true;
        };
        auto check = [&](const merc::io::FieldMeta &m, int b, const std::array<int, 6> &box) {
            const merc::FieldSpec *f = merc::find_field(m.name.c_str());
            if (box != alloc_box(F.blk[b], f->loc)) throw std::runtime_error("restart: lo/hi mismatch for " + m.name + " block " + std::to_string(b));
        };
        // every value is staged on the host first; the device fields change only
        // after the whole file has been read and validated
        std::map<std::pair<int, int>, std::vector<double>> staged;  // (slot, block) -> box array
        auto put = [&](const merc::io::FieldMeta &m, int b, const std::array<int, 6> &box, const double *vals) {
            const merc::FieldSpec *f = merc::find_field(m.name.c_str());
            const MercBlk &k = F.blk[b];
            for (int c = 0; c < f->ncomp; ++c) {
                std::vector<double> &comp = staged[{f->slot + c, b}];
                if (comp.empty()) {
                    comp.resize(F.vol(b));
                    F.download_block(f->slot + c, b, comp.data());  // keeps the box points outside the allocation
                }
                size_t t = (size_t)c;
                for (int i = box[0]; i < box[3]; ++i)
                    for (int j = box[1]; j < box[4]; ++j)
                        for (int kk = box[2]; kk < box[5]; ++kk, t += f->ncomp)
                            comp[(long long)(i + k.g) * k.sI + (long long)(j + k.g) * k.sJ + (kk + k.g)] = vals[t];
            }
            delivered[m.name][b] = 1;
        };
        merc::io::read_restart(path, F.nb(), want, check, put, step, time);
        if (njedge != 0 && njedge != 3) throw std::runtime_error("restart: incomplete J edge triplet (J_xi, J_eta, J_zeta go together)");
        for (const auto &kv : delivered)
            for (int b = 0; b < F.nb(); ++b)
                if (!kv.second[b])
                    throw std::runtime_error("restart: block " + std::to_string(b) + " of " + kv.first + " is inactive in the file");
        for (const auto &kv : staged) F.upload_block(kv.first.first, kv.first.second, kv.second.data());
    })
}

int mercury_write_rundata(const char *path, int step, double time, double dt, int nref, const double *residual_ref,
                          double residual_max) {
    if (!path || nref < 0 || (nref > 0 && !residual_ref)) {
        merc::g_last_error = "mercury_write_rundata: invalid argument";
        return -1;
    }
    ABI_TRY({
        merc::io::RunData r;
        r.step = step;
        r.time = time;
        r.dt = dt;
        r.residual_ref.assign(residual_ref, residual_ref + nref);
        r.residual_max = residual_max;
        merc::io::write_rundata(path, r);
    })
}

int mercury_read_rundata(const char *path, int *step, double *time, double *dt, int max_ref, int *nref, double *residual_ref,
                         double *residual_max) {
    if (!path) {
        merc::g_last_error = "mercury_read_rundata: path is null";
        return -1;
    }
    ABI_TRY({
        const merc::io::RunData r = merc::io::read_rundata(path);
        if (step) *step = r.step;
        if (time) *time = r.time;
        if (dt) *dt = r.dt;
        if (nref) *nref = (int)r.residual_ref.size();
        if ((int)r.residual_ref.size() > max_ref && residual_ref)
            throw std::runtime_error("mercury_read_rundata: residual_ref holds " + std::to_string(r.residual_ref.size()) + " values");
        if (residual_ref)
            for (size_t i = 0; i < r.residual_ref.size(); ++i) residual_ref[i] = r.residual_ref[i];
        if (residual_max) *residual_max = r.residual_max;
    })
}

// ---------------------------------------------------------------------------
// visualisation output: Tecplot binary with every variable at the nodes
// ---------------------------------------------------------------------------
extern "C++" {
namespace {
// the Mercury output list and its component names (Z4_Mercury/src/MercurySolver.cpp:53-86)
struct TecOutput {
    const char *field;
    std::vector<std::string> comps;
};
const TecOutput kTecOutputs[] = {
    {"PV_H", {"u_H", "v_H", "w_H", "p_H", "T_H"}},
    {"PV_Na", {"u_Na", "v_Na", "w_Na", "p_Na", "T_Na"}},
    {"B_cell", {"Bx", "By", "Bz"}},
    {"Bind_cell", {"bx_ind", "by_ind", "bz_ind"}},
    {"Na", {"n_Na"}},
    {"J_cell", {"Jx", "Jy", "Jz"}},
};

// names of a field's components: the Mercury names, else name_c
// (IOModule::AddTecFieldOrGroup_, src/6_io/IOModule_Tecplot.cpp:119-226)
std::string tec_component_name(const std::string &field, int c) {
    for (const TecOutput &t : kTecOutputs)
        if (field == t.field && c < (int)t.comps.size()) return t.comps[c];
    return field + "_" + std::to_string(c);
}

struct DeviceScratch {
    double *p = nullptr;
    explicit DeviceScratch(size_t n) { HIPCHK(hipMalloc(&p, sizeof(double) * (n ? n : 1))); }
    ~DeviceScratch() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace
}

int mercury_write_tecplot(MercuryContext *ctx, const char *path, double time, const char *const *fields, int nfields,
                          const char *const *block_names) {
    CHECK_CTX(ctx);
    if (!path) {
        merc::g_last_error = "mercury_write_tecplot: path is null";
        return -1;
    }
    if (nfields < 0 || (nfields > 0 && !fields)) {
        merc::g_last_error = "mercury_write_tecplot: bad field list";
        return -1;
    }
    ABI_TRY({
        merc::MercurySolver &S = *ctx->s;
        HIPCHK(hipStreamSynchronize(S.st));
        const merc::Field &F = *S.F;
        std::vector<std::string> names;
        if (nfields > 0)
            for (int t = 0; t < nfields; ++t) names.push_back(fields[t] ? fields[t] : "");
        else
            for (const auto &t : kTecOutputs) names.push_back(t.field);
        // variables: coordinates, then the components of every cell or node field
        struct Var {
            int slot, loc;
        };
        std::vector<std::string> vnames = {"x", "y", "z"};
        std::vector<Var> vars;
        std::vector<std::string> seen;
        for (const std::string &nm : names) {
            if (nm.empty()) continue;
            const merc::FieldSpec *f = merc::find_field(nm.c_str());
            if (!f || (f->loc != LOC_CELL && f->loc != LOC_NODE)) {
                std::fprintf(stderr, "[mercury_write_tecplot] skip missing or non-cell/node field: %s\n", nm.c_str());
                continue;
            }
            if (std::find(seen.begin(), seen.end(), nm) != seen.end()) continue;
            seen.push_back(nm);
            for (int c = 0; c < f->ncomp; ++c) {
                vnames.push_back(tec_component_name(nm, c));
                vars.push_back({f->slot + c, f->loc});
            }
        }
        for (const Var &v : vars) S.materialize(v.slot);
        std::vector<merc::io::TecplotWriter::Zone> zones;
        long long nmax = 0;
        for (int b = 0; b < F.nb(); ++b) {
            const MercBlk &k = F.blk[b];
            merc::io::TecplotWriter::Zone z;
            z.name = "Zone" + std::to_string(b) + "_" + ((block_names && block_names[b]) ? block_names[b] : "Fluid");
            z.imax = k.m[0] + 1;
            z.jmax = k.m[1] + 1;
            z.kmax = k.m[2] + 1;
            zones.push_back(z);
            nmax = std::max(nmax, (long long)z.imax * z.jmax * z.kmax);
        }
        merc::io::TecplotWriter w(path, "flow_field", vnames, zones, time);
        DeviceScratch dnode((size_t)nmax);
        // singular nodes (the end nodes of singular edges): the |Jac|-weighted mean
        // over the cells of the physical node instead of the block's own cells
        // (EvalTecplotSingularNodeAverage_, src/6_io/IOModule_Tecplot.cpp:611-631, 358-365)
        std::vector<std::vector<double>> smean(vars.size());
        if (S.sing.nnode > 0) {
            DeviceScratch dm((size_t)S.sing.nnode);
            for (size_t v = 0; v < vars.size(); ++v) {
                if (vars[v].loc != LOC_CELL) continue;
                launch_node_stencil_mean(S.st, S.sing.nnode, S.sing.nstart.p, S.sing.ncell.p, F.hptr[vars[v].slot], F.hptr[SL_JAC], dm.p);
                smean[v].resize(S.sing.nnode);
                HIPCHK(hipMemcpyAsync(smean[v].data(), dm.p, sizeof(double) * S.sing.nnode, hipMemcpyDeviceToHost, S.st));
            }
            HIPCHK(hipStreamSynchronize(S.st));
        }
        std::vector<double> tmp, box;
        for (int b = 0; b < F.nb(); ++b) {
            const MercBlk &k = F.blk[b];
            const int ni = k.m[0] + 1, nj = k.m[1] + 1, nk = k.m[2] + 1;
            const size_t n = (size_t)ni * nj * nk;
            auto file_index = [&](int i, int j, int kk) { return ((size_t)kk * nj + j) * ni + i; };
            std::vector<std::vector<double>> values(vnames.size(), std::vector<double>(n));
            const merc::BlockGeom &G = *S.geom[b];
            for (int i = 0; i < ni; ++i)
                for (int j = 0; j < nj; ++j)
                    for (int kk = 0; kk < nk; ++kk) {
                        const auto x = G.node(i, j, kk);
                        for (int c = 0; c < 3; ++c) values[c][file_index(i, j, kk)] = x[c];
                    }
            for (size_t v = 0; v < vars.size(); ++v) {
                std::vector<double> &dst = values[3 + v];
                if (vars[v].loc == LOC_CELL) {
                    merc::launch_node_average(S.st, F.hptr[vars[v].slot] + k.off, F.hptr[SL_JAC] + k.off, k, dnode.p);
                    tmp.resize(n);
                    HIPCHK(hipMemcpyAsync(tmp.data(), dnode.p, sizeof(double) * n, hipMemcpyDeviceToHost, S.st));
                    HIPCHK(hipStreamSynchronize(S.st));
                    for (int i = 0; i < ni; ++i)
                        for (int j = 0; j < nj; ++j)
                            for (int kk = 0; kk < nk; ++kk) dst[file_index(i, j, kk)] = tmp[((size_t)i * nj + j) * nk + kk];
                    if (!smean[v].empty())
                        for (const merc::patch::Dof &a : S.sing.nalias)
                            if (a.b == b) dst[file_index(a.i, a.j, a.k)] = smean[v][a.f];
                } else {
                    box.resize(F.vol(b));
                    F.download_block(vars[v].slot, b, box.data());
                    for (int i = 0; i < ni; ++i)
                        for (int j = 0; j < nj; ++j)
                            for (int kk = 0; kk < nk; ++kk)
                                dst[file_index(i, j, kk)] = box[(size_t)(i + k.g) * k.sI + (size_t)(j + k.g) * k.sJ + (kk + k.g)];
                }
            }
            w.zone_data(values);
        }
        w.close();
    })
}

// ---------------------------------------------------------------------------
// CASE geometry directories (Grid::ReadGeometry + TOPO::build_topology)
// ---------------------------------------------------------------------------
struct MercuryCase {
    merc::casefile::Geometry G;
};

int mercury_case_read(MercuryCase **out, const char *geometry_dir, int nrank, int binary) {
    if (!out || !geometry_dir) {
        merc::g_last_error = "mercury_case_read: null argument";
        return -1;
    }
    *out = nullptr;
    ABI_TRY({
        std::unique_ptr<MercuryCase> c(new MercuryCase);
        c->G = merc::casefile::Geometry::read(geometry_dir, nrank, binary != 0);
        *out = c.release();
    })
}

void mercury_case_destroy(MercuryCase *c) { delete c; }

int mercury_case_blocks(const MercuryCase *c, int *nblocks, int *dims) {
    if (!c || !nblocks) {
        merc::g_last_error = "mercury_case_blocks: null argument";
        return -1;
    }
    *nblocks = (int)c->G.blocks.size();
    if (dims)
        for (size_t b = 0; b < c->G.blocks.size(); ++b)
            for (int a = 0; a < 3; ++a) dims[3 * b + a] = c->G.blocks[b].nn[a] - 1;
    return 0;
}

int mercury_case_block_name(const MercuryCase *c, int block, char *buf, int len) {
    if (!c || !buf || len <= 0 || block < 0 || block >= (int)c->G.blocks.size()) {
        merc::g_last_error = "mercury_case_block_name: bad argument";
        return -1;
    }
    const std::string &n = c->G.blocks[block].name;
    if ((int)n.size() >= len) {
        merc::g_last_error = "mercury_case_block_name: buffer too small";
        return -1;
    }
    std::memcpy(buf, n.c_str(), n.size() + 1);
    return 0;
}

int mercury_case_nodes(const MercuryCase *c, double *owned_nodes) {
    if (!c || !owned_nodes) {
        merc::g_last_error = "mercury_case_nodes: null argument";
        return -1;
    }
    double *o = owned_nodes;
    for (const auto &B : c->G.blocks) {
        const int ni = B.nn[0], nj = B.nn[1], nk = B.nn[2];
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < ni; ++i)
                for (int j = 0; j < nj; ++j)
                    for (int k = 0; k < nk; ++k) *o++ = B.xyz[a][((size_t)k * nj + j) * ni + i];
    }
    return 0;
}

int mercury_case_topology(const MercuryCase *c, int *bc, int *nbr, int *transform) {
    if (!c || !bc || !nbr || !transform) {
        merc::g_last_error = "mercury_case_topology: null argument";
        return -1;
    }
    ABI_TRY({ c->G.face_topology(bc, nbr, transform); })
}

int mercury_case_patch_topology(const MercuryCase *c, int *bc, int capacity, int *npatch, int *patches) {
    if (!c || !bc || !npatch) {
        merc::g_last_error = "mercury_case_patch_topology: null argument";
        return -1;
    }
    ABI_TRY({
        std::vector<int> pt;
        c->G.patch_topology(bc, pt);
        *npatch = (int)pt.size() / 18;
        if (patches && capacity >= *npatch) std::copy(pt.begin(), pt.end(), patches);
    })
}

int mercury_case_interfaces(const MercuryCase *c, int *n, int *records) {
    if (!c || !n) {
        merc::g_last_error = "mercury_case_interfaces: null argument";
        return -1;
    }
    *n = (int)c->G.interfaces.size();
    if (records)
        for (const auto &p : c->G.interfaces) {
            const int head[6] = {p.kind, p.periodic ? 1 : 0, p.block, p.nb_block, p.direction, p.nb_direction};
            for (int v : head) *records++ = v;
            for (const int *a : {p.lo, p.hi, p.nb_lo, p.nb_hi, p.perm, p.sign, p.offset})
                for (int d = 0; d < 3; ++d) *records++ = a[d];
        }
    return 0;
}

int mercury_case_physical(const MercuryCase *c, int *n, int *records) {
    if (!c || !n) {
        merc::g_last_error = "mercury_case_physical: null argument";
        return -1;
    }
    *n = (int)c->G.physical.size();
    if (records)
        for (const auto &p : c->G.physical) {
            *records++ = p.block;
            *records++ = p.direction;
            *records++ = p.bc_id;
            for (int d = 0; d < 3; ++d) *records++ = p.sub[d];
            for (int d = 0; d < 3; ++d) *records++ = p.sup[d] + 1;
        }
    return 0;
}

int mercury_init_bface_dipoles(MercuryContext *ctx, const double *Bu, const double *const *A, int ndipoles, const double *dipoles) {
    CHECK_CTX(ctx);
    if (!Bu) {
        merc::g_last_error = "mercury_init_bface: Bu is null";
        return -1;
    }
    ABI_TRY({
        HIPCHK(hipStreamSynchronize(ctx->s->st));
        ctx->s->init_bface(Bu, A, ndipoles, dipoles);
    })
}

int mercury_init_bface(MercuryContext *ctx, const double *Bu, const double *const *A) {
    return mercury_init_bface_dipoles(ctx, Bu, A, 0, nullptr);
}

int mercury_prepare(MercuryContext *ctx) {
    CHECK_CTX(ctx);
    ABI_TRY({ ctx->s->prepare(); })
}

int mercury_step(MercuryContext *ctx, int nsteps, int integrator, double *dt_out) {
    CHECK_CTX(ctx);
    if (nsteps < 0 || (integrator != 0 && integrator != 1)) {
        merc::g_last_error = "mercury_step: invalid nsteps or integrator";
        return -1;
    }
    ABI_TRY({ ctx->s->step(nsteps, integrator, dt_out); })
}

int mercury_op(MercuryContext *ctx, const char *op, double arg) {
    CHECK_CTX(ctx);
    if (!op) {
        merc::g_last_error = "mercury_op: null op";
        return -1;
    }
    ABI_TRY({ ctx->s->run_op(op, arg); })
}

int mercury_timestep(MercuryContext *ctx, double *dt) {
    CHECK_CTX(ctx);
    ABI_TRY({
        ctx->s->run_op("timestep", 0.0);
        const double v = ctx->s->read_dt();
        if (dt) *dt = v;
    })
}

int mercury_set_divb_monitor(MercuryContext *ctx, int on) {
    CHECK_CTX(ctx);
    ABI_TRY({
        HIPCHK(hipStreamSynchronize(ctx->s->st));
        if (ctx->s->divb_mon_on != (on != 0)) {
            ctx->s->divb_mon_on = on != 0;
            ctx->s->drop_graph();
        }
    })
}

int mercury_divb_monitor(MercuryContext *ctx, double *last, double *max_since_reset, double *last_maxabs, int reset) {
    CHECK_CTX(ctx);
    ABI_TRY({ ctx->s->divb_monitor(last, max_since_reset, last_maxabs, reset != 0); })
}

int mercury_divb(MercuryContext *ctx, double *normalised, double *maxabs) {
    CHECK_CTX(ctx);
    ABI_TRY({ ctx->s->divb(normalised, maxabs); })
}

// ---------------------------------------------------------------------------
// DEC operators on caller device arrays (one block)
// ---------------------------------------------------------------------------
static void dec_args(const int *dims, int nghost, std::initializer_list<const void *> ptrs) {
    merc::dec_check_box(dims, nghost);
    for (const void *p : ptrs)
        if (!p) throw std::runtime_error("DEC: a device array is NULL");
}

int mercury_dec_d0(const int *dims, int nghost, const double *node, double *const *edge, void *stream) {
    ABI_TRY({
        if (!edge) throw std::runtime_error("DEC: edge is NULL");
        dec_args(dims, nghost, {node, edge[0], edge[1], edge[2]});
        merc::dec_d0(dims, nghost, node, edge, (hipStream_t)stream);
    })
}

int mercury_dec_d1(const int *dims, int nghost, const double *const *edge, double *const *face, void *stream) {
    ABI_TRY({
        if (!edge || !face) throw std::runtime_error("DEC: edge or face is NULL");
        dec_args(dims, nghost, {edge[0], edge[1], edge[2], face[0], face[1], face[2]});
        merc::dec_d1(dims, nghost, edge, face, 0, 0.0, (hipStream_t)stream);
    })
}

int mercury_dec_d2(const int *dims, int nghost, const double *const *face, double *cell, void *stream) {
    ABI_TRY({
        if (!face) throw std::runtime_error("DEC: face is NULL");
        dec_args(dims, nghost, {face[0], face[1], face[2], cell});
        merc::dec_d2(dims, nghost, face, cell, (hipStream_t)stream);
    })
}

int mercury_dec_faraday_update(const int *dims, int nghost, const double *const *edge, double *const *face, double dt, void *stream) {
    ABI_TRY({
        if (!edge || !face) throw std::runtime_error("DEC: edge or face is NULL");
        dec_args(dims, nghost, {edge[0], edge[1], edge[2], face[0], face[1], face[2]});
        merc::dec_d1(dims, nghost, edge, face, 1, dt, (hipStream_t)stream);
    })
}

// partial maxima in a stream-ordered allocation released on the same stream
int mercury_dec_max_abs_div_curl(const int *dims, int nghost, const double *const *edge, double *result, void *stream) {
    ABI_TRY({
        if (!edge) throw std::runtime_error("DEC: edge is NULL");
        dec_args(dims, nghost, {edge[0], edge[1], edge[2], result});
        const hipStream_t st = (hipStream_t)stream;
        double *part = nullptr;
        HIPCHK(hipMallocAsync((void **)&part, sizeof(double) * merc::dec_div_curl_groups(dims), st));
        merc::dec_div_curl(dims, nghost, edge, part, result, st);
        HIPCHK(hipFreeAsync(part, st));
    })
}

int mercury_dec_max_abs_curl_grad(const int *dims, int nghost, const double *node, double *result, void *stream) {
    ABI_TRY({
        dec_args(dims, nghost, {node, result});
        const hipStream_t st = (hipStream_t)stream;
        double *part = nullptr;
        HIPCHK(hipMallocAsync((void **)&part, sizeof(double) * merc::dec_curl_grad_groups(dims), st));
        merc::dec_curl_grad(dims, nghost, node, part, result, st);
        HIPCHK(hipFreeAsync(part, st));
    })
}

// the same diagnostic with its partial maxima in the library's static scratch
// array: no allocation, at most MERC_DEC_SCRATCH_DOUBLES workgroups
int mercury_dec_max_abs_div_curl_noalloc(const int *dims, int nghost, const double *const *edge, double *result, void *stream) {
    ABI_TRY({
        if (!edge) throw std::runtime_error("DEC: edge is NULL");
        dec_args(dims, nghost, {edge[0], edge[1], edge[2], result});
        const long long ng = merc::dec_div_curl_groups(dims);
        if (ng > MERC_DEC_SCRATCH_DOUBLES)
            throw std::runtime_error("mercury_dec_max_abs_div_curl_noalloc: " + std::to_string((long long)dims[0] * dims[1] * dims[2]) +
                                     " cells exceed the scratch limit of " + std::to_string(256LL * MERC_DEC_SCRATCH_DOUBLES) +
                                     " cells per block");
        merc::dec_div_curl(dims, nghost, edge, merc::dec_noalloc_scratch(), result, (hipStream_t)stream);
    })
}

int mercury_synchronize(MercuryContext *ctx) {
    CHECK_CTX(ctx);
    ABI_TRY({ HIPCHK(hipStreamSynchronize(ctx->s->st)); })
}

int mercury_set_form(MercuryContext *ctx, int form) {
    CHECK_CTX(ctx);
    ABI_TRY({ ctx->s->set_form(form); })
}

// ---------------------------------------------------------------------------
// ensembles: independent contexts stepped by one graph
// ---------------------------------------------------------------------------
struct MercuryEnsemble {
    std::vector<MercuryContext *> ctx;
    std::vector<unsigned long long> gen;  // each member's generation at capture
    hipStream_t root = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~MercuryEnsemble() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : join)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (root) (void)hipStreamDestroy(root);
    }
};

int mercury_ensemble_create(MercuryContext *const *ctxs, int n, int integrator, MercuryEnsemble **out) {
    if (!ctxs || n <= 0 || !out) {
        merc::g_last_error = "mercury_ensemble_create: invalid argument";
        return -1;
    }
    for (int i = 0; i < n; ++i) CHECK_CTX(ctxs[i]);
    std::unique_ptr<MercuryEnsemble> e(new MercuryEnsemble());
    ABI_TRY({
        for (int i = 0; i < n; ++i)
            if (ctxs[i]->s->resist) throw std::runtime_error("an ensemble cannot hold a context with the implicit resistive correction");
        e->ctx.assign(ctxs, ctxs + n);
        HIPCHK(hipStreamCreateWithFlags(&e->root, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
        e->join.assign(n, nullptr);
        for (int i = 0; i < n; ++i) HIPCHK(hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming));
        // fork from the root stream into every context's stream, join back
        std::vector<hipStream_t> all{e->root};
        for (int i = 0; i < n; ++i) {
            all.push_back(e->ctx[i]->s->st);
            all.push_back(e->ctx[i]->s->st2);
            e->gen.push_back(e->ctx[i]->s->generation);
        }
        {
            merc::CaptureGuard guard(all);
            HIPCHK(hipStreamBeginCapture(e->root, hipStreamCaptureModeThreadLocal));
            guard.active = true;
            HIPCHK(hipEventRecord(e->fork, e->root));
            for (int i = 0; i < n; ++i) {
                hipStream_t cs = e->ctx[i]->s->st;
                HIPCHK(hipStreamWaitEvent(cs, e->fork, 0));
                e->ctx[i]->s->enqueue_step(integrator);
                HIPCHK(hipEventRecord(e->join[i], cs));
                HIPCHK(hipStreamWaitEvent(e->root, e->join[i], 0));
            }
            guard.active = false;
            HIPCHK(hipStreamEndCapture(e->root, &e->graph));
        }
        HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        *out = e.release();
    })
}

int mercury_ensemble_launch(MercuryEnsemble *e, int nsteps) {
    if (!e || !e->exec || nsteps < 0) {
        merc::g_last_error = "mercury_ensemble_launch: invalid ensemble";
        return -1;
    }
    ABI_TRY({
        // a member whose form or launch shape changed since the capture would
        // replay kernels with arguments that no longer hold: refuse before launching
        for (size_t i = 0; i < e->ctx.size(); ++i)
            if (e->ctx[i]->s->generation != e->gen[i])
                throw std::runtime_error("ensemble member " + std::to_string(i) +
                                         " changed form or launch shape after the ensemble was captured; create the ensemble again");
        for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(e->exec, e->root));
        HIPCHK(hipStreamSynchronize(e->root));
        for (size_t i = 0; i < e->ctx.size(); ++i) {
            try {
                e->ctx[i]->s->check_status();
            } catch (const std::exception &x) {
                throw std::runtime_error("ensemble member " + std::to_string(i) + ": " + x.what());
            }
        }
    })
}

int mercury_ensemble_destroy(MercuryEnsemble *e) {
    if (!e) return 0;
    ABI_TRY({ delete e; })
}

int mercury_graph_capture(MercuryContext *ctx, int integrator) {
    CHECK_CTX(ctx);
    ABI_TRY({ ctx->s->capture(integrator); })
}

int mercury_graph_launch(MercuryContext *ctx, int nsteps) {
    CHECK_CTX(ctx);
    ABI_TRY({
        ctx->s->graph_launch(nsteps);
        HIPCHK(hipGetLastError());
    })
}

int mercury_resist_stats(MercuryContext *ctx, int *n_unknowns, int *iterations, double *rel_residual) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (n_unknowns) *n_unknowns = ctx->s->resist ? ctx->s->rd.n : 0;
        if (iterations) *iterations = ctx->s->res_its;
        if (rel_residual) *rel_residual = ctx->s->res_rel;
    })
}

int mercury_autotune(MercuryContext *ctx, int *march_out) {
    CHECK_CTX(ctx);
    ABI_TRY({
        const int L = ctx->s->autotune();
        if (march_out) *march_out = L;
    })
}

int mercury_set_flux_shape(MercuryContext *ctx, int kind, int length) {
    CHECK_CTX(ctx);
    if (length < 1) {
        merc::g_last_error = "mercury_set_flux_shape: length must be >= 1";
        return -1;
    }
    ABI_TRY({ ctx->s->set_flux_shape(kind, length); })
}

int mercury_flux_shape(MercuryContext *ctx, int *kind, int *length) {
    CHECK_CTX(ctx);
    ABI_TRY({
        if (kind) *kind = ctx->s->flux_kind();
        if (length) *length = ctx->s->tile_c > 0 ? ctx->s->tile_c : ctx->s->march;
    })
}

int mercury_time_op(MercuryContext *ctx, const char *op, int reps, double *ms_avg) {
    CHECK_CTX(ctx);
    if (!op || reps <= 0 || !ms_avg) {
        merc::g_last_error = "mercury_time_op: invalid argument";
        return -1;
    }
    ABI_TRY({ *ms_avg = ctx->s->time_op(op, reps); })
}

int mercury_geometry_block(const int *dims, int nghost, const double *owned_nodes, const MercuryPhysics *phys, const char *name,
                           int comp, double *out) {
    if (!dims || !owned_nodes || !phys || !name || !out || nghost < 2) {
        merc::g_last_error = "mercury_geometry_block: invalid argument";
        return -1;
    }
    ABI_TRY({
        MercPhys P;
        std::memcpy(&P, phys, sizeof(P));
        const int slot = field_comp_slot(name, comp);
        merc::BlockGeom G(dims, nghost, owned_nodes);
        struct W : merc::SlotWriter {
            int want;
            double *out;
            long long V;
            void put(int s, const std::vector<double> &v) override {
                if (s == want) std::memcpy(out, v.data(), sizeof(double) * V);
            }
        } w;
        w.want = slot;
        w.out = out;
        w.V = G.V;
        G.compute(P, w);
    })
}

long long mercury_total_volume(MercuryContext *ctx) {
    if (!ctx || !ctx->s) return -1;
    return ctx->s->F->Vtot;
}

}  // extern "C"
