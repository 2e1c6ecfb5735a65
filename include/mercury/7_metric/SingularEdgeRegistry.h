// Singular physical edges of the drop-in C++ interface: the shared edges whose
// aliases touch three, or more than four, owner faces (the corner lines of a
// cubed sphere).  Replaces METRIC::SingularEdgeRegistry for reading
// (include/7_metric/SingularEdgeRegistry.h, src/7_metric/SingularEdgeRegistry.cpp:
// 78-282): the context that Field builds on patch connections finds the same
// edges, replaces alpha on their aliases and assembles E, E_hall, J and dJ on
// them inside its step, so no contribution callbacks are needed here.
#pragma once

#include <array>
#include <cstddef>
#include <vector>

#include "../2_topology/Topology.h"
#include "../3_field/Field.h"

namespace METRIC {

struct SingularPhysicalEdge {
    int owner_block = -1, i = 0, j = 0, k = 0, axis = 0;  // the owner edge DOF
    int owner_orientation = +1;                         // the owner alias's orientation
    double inverse_hodge = 0.0;                         // |edge| / |dual face|
    std::array<double, 3> canonical_edge_vector{{0.0, 0.0, 0.0}};
    int global_valid_cell_count = 0;
    int global_valid_face_count = 0;
};

class SingularEdgeRegistry {
public:
    // rank is kept for the reference's signature (one process holds every block)
    void build(const TOPO::Topology &topology, Field &fields, Grid &grid, int rank = 0) {
        (void)rank;
        entries_.clear();
        if (topology.patches.empty()) return;  // per-face connections: no edge of valence 3
        const int nb = grid.num_blocks();
        int n = 0;
        mercury_hip_detail::check(mercury_singular_edges(nb, grid.all_dims().data(), fields.nghost(), topology.bc.data(),
                                                         (int)topology.patches.size() / 18, topology.patches.data(), topology.rule,
                                                         grid.owned_nodes().data(), &fields.physics(), 0, &n, nullptr, nullptr),
                                  "SingularEdgeRegistry::build");
        std::vector<int> iv(8 * (size_t)n);
        std::vector<double> dv(4 * (size_t)n);
        mercury_hip_detail::check(mercury_singular_edges(nb, grid.all_dims().data(), fields.nghost(), topology.bc.data(),
                                                         (int)topology.patches.size() / 18, topology.patches.data(), topology.rule,
                                                         grid.owned_nodes().data(), &fields.physics(), n, &n, iv.data(), dv.data()),
                                  "SingularEdgeRegistry::build");
        for (int t = 0; t < n; ++t) {
            SingularPhysicalEdge e;
            const int *v = &iv[8 * (size_t)t];
            e.owner_block = v[0], e.i = v[1], e.j = v[2], e.k = v[3], e.axis = v[4], e.owner_orientation = v[5];
            e.global_valid_cell_count = v[6], e.global_valid_face_count = v[7];
            e.inverse_hodge = dv[4 * (size_t)t];
            for (int q = 0; q < 3; ++q) e.canonical_edge_vector[q] = dv[4 * (size_t)t + 1 + q];
            entries_.push_back(e);
        }
    }
    // build() already throws where the reference aborts
    void validate_or_abort() const {}
    const std::vector<SingularPhysicalEdge> &entries() const { return entries_; }
    bool empty() const { return entries_.empty(); }
    std::size_t size() const { return entries_.size(); }

private:
    std::vector<SingularPhysicalEdge> entries_;
};

}  // namespace METRIC
