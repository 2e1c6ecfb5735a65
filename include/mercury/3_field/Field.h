// Field storage of the drop-in C++ interface: every field of the Mercury
// catalogue lives in the device context, one FP64 array per component over
// each block box (the storage convention of mercury_hip.h).
// Replaces Field (include/3_field/Field.h:16-96), FieldBlock
// (include/3_field/FieldBlock.h:12-67) and FieldDescriptor
// (include/3_field/FieldDescriptor.h:31-78).  Differences, each forced by
// device residency: FieldBlock is a device view (data() is a device address;
// element access goes through download/upload), one FieldBlock per component,
// and the catalogue is the fixed Mercury one (MercurySolver::RegisterFields),
// allocated when the Field is built from the grid, topology and physics.
#pragma once

#include <string>
#include <vector>

#include "../1_grid/1_MPCNS_Grid.h"
#include "../2_topology/Topology.h"

// include/0_basic/StaggerLocation.h:3-15
enum class StaggerLocation { Cell, Node, FaceXi, FaceEt, FaceZe, EdgeXi, EdgeEt, EdgeZe };

// the library's location codes (mercury_field_info): cell, faces xi/eta/zeta, edges xi/eta/zeta, node
inline StaggerLocation stagger_from_library(int loc) {
    static const StaggerLocation L[8] = {StaggerLocation::Cell,   StaggerLocation::FaceXi, StaggerLocation::FaceEt,
                                         StaggerLocation::FaceZe, StaggerLocation::EdgeXi, StaggerLocation::EdgeEt,
                                         StaggerLocation::EdgeZe, StaggerLocation::Node};
    if (loc < 0 || loc > 7) throw std::runtime_error("unknown location code");
    return L[loc];
}

struct Int3 {
    int i, j, k;
};

struct FieldDescriptor {
    std::string name;
    StaggerLocation location;
    int ncomp;
    int nghost;
};

class FieldBlock {
public:
    FieldBlock() = default;
    FieldBlock(MercuryContext *ctx, const FieldDescriptor &d, int block, int comp) : ctx_(ctx), desc_(d), block_(block), comp_(comp) {
        using mercury_hip_detail::check;
        check(mercury_block_shape(ctx, block, m_, &g_), "FieldBlock");
        check(mercury_field_device(ctx, d.name.c_str(), block, comp, &data_), "FieldBlock");
    }
    bool is_allocated() const { return data_ != nullptr; }
    const FieldDescriptor &descriptor() const { return desc_; }
    int block() const { return block_; }
    int component() const { return comp_; }
    const int *cells() const { return m_; }
    int nghost() const { return g_; }
    // allocated range [lo, hi) of the location and its owned range [0, inner_hi)
    Int3 get_lo() const { return {-g_, -g_, -g_}; }
    Int3 get_hi() const {
        const int *d = delta();
        return {m_[0] + 1 - d[0] + g_, m_[1] + 1 - d[1] + g_, m_[2] + 1 - d[2] + g_};
    }
    Int3 inner_lo() const { return {0, 0, 0}; }
    Int3 inner_hi() const {
        const int *d = delta();
        return {m_[0] + 1 - d[0], m_[1] + 1 - d[1], m_[2] + 1 - d[2]};
    }
    // element offset of (i, j, k) in the block box (FieldBlock::operator(), :38-49)
    long long index(int i, int j, int k) const {
        const long long N1 = m_[1] + 1 + 2LL * g_, N2 = m_[2] + 1 + 2LL * g_;
        return ((long long)(i + g_) * N1 + (j + g_)) * N2 + (k + g_);
    }
    long long volume() const { return (long long)(m_[0] + 1 + 2 * g_) * (m_[1] + 1 + 2 * g_) * (m_[2] + 1 + 2 * g_); }
    // device address of the box (FieldBlock::raw_data, :52)
    double *data() const { return data_; }
    // host copies of the whole box
    std::vector<double> download() const {
        std::vector<double> h((size_t)volume());
        mercury_hip_detail::check(mercury_download(ctx_, desc_.name.c_str(), block_, comp_, h.data()), "FieldBlock::download");
        return h;
    }
    void upload(const std::vector<double> &h) {
        if ((long long)h.size() != volume()) throw std::runtime_error("FieldBlock::upload: size does not match the block box");
        mercury_hip_detail::check(mercury_upload(ctx_, desc_.name.c_str(), block_, comp_, h.data()), "FieldBlock::upload");
    }

private:
    const int *delta() const {
        // LayoutTraits dof_delta (src/0_basic/LayoutTraits.cpp:160-183)
        static const int D[8][3] = {{1, 1, 1}, {0, 0, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        return D[(int)desc_.location];
    }
    MercuryContext *ctx_ = nullptr;
    FieldDescriptor desc_{};
    int block_ = 0, comp_ = 0;
    int m_[3] = {0, 0, 0}, g_ = 0;
    double *data_ = nullptr;
};

class Field {
public:
    // Field(Grid*, Param*, ngg) + MercurySolver::RegisterFields + METRIC::build_field_geometry:
    // the device context with every Mercury field and the static geometry.
    Field(Grid *grid, const TOPO::Topology &topo, const MercuryPhysics &phys, int ngg) : grid_(grid), ngg_(ngg), phys_(phys) {
        if (!grid || grid->num_blocks() == 0) throw std::runtime_error("Field: empty grid");
        if (topo.num_blocks() != grid->num_blocks()) throw std::runtime_error("Field: topology and grid disagree on the block count");
        if (!topo.patches.empty())
            mercury_hip_detail::check(mercury_create_patched(&ctx_, grid->num_blocks(), grid->all_dims().data(), ngg, topo.bc.data(),
                                                             (int)topo.patches.size() / 18, topo.patches.data(), topo.rule,
                                                             grid->owned_nodes().data(), &phys),
                                      "Field");
        else
            mercury_hip_detail::check(mercury_create_connected(&ctx_, grid->num_blocks(), grid->all_dims().data(), ngg, topo.bc.data(),
                                                               topo.nbr.data(), topo.transform.data(), grid->owned_nodes().data(),
                                                               &phys),
                                      "Field");
    }
    ~Field() {
        if (ctx_) mercury_destroy(ctx_);
    }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    int num_blocks() const { return grid_->num_blocks(); }
    int field_id(const std::string &name) const {
        int slot = -1;
        mercury_hip_detail::check(mercury_field_info(name.c_str(), &slot, nullptr, nullptr), "Field::field_id");
        return slot;
    }
    bool has_field(const std::string &name) const { return mercury_field_info(name.c_str(), nullptr, nullptr, nullptr) == 0; }
    FieldDescriptor descriptor(const std::string &name) const {
        int nc = 0, loc = 0;
        mercury_hip_detail::check(mercury_field_info(name.c_str(), nullptr, &nc, &loc), "Field::descriptor");
        return {name, stagger_from_library(loc), nc, ngg_};
    }
    // one component of a field on one block (the reference's FieldBlock holds
    // all components; here component m of FieldBlock(i,j,k,m) is field(name, ib, m))
    FieldBlock field(const std::string &name, int iblock, int comp = 0) const {
        if (iblock < 0 || iblock >= num_blocks()) throw std::runtime_error("Field::field: block out of range");
        return FieldBlock(ctx_, descriptor(name), iblock, comp);
    }
    MercuryContext *context() const { return ctx_; }
    Grid *grid() const { return grid_; }
    int nghost() const { return ngg_; }
    const MercuryPhysics &physics() const { return phys_; }

private:
    Grid *grid_ = nullptr;
    int ngg_ = 0;
    MercuryPhysics phys_;
    MercuryContext *ctx_ = nullptr;
};
