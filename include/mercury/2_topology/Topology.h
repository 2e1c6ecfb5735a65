// Block topology of the drop-in C++ interface: per block face a boundary code,
// and for connections the neighbour block and the IndexTransform
// (perm[3], sign[3], offset[3], include/2_topology/TopologyTypes.h:91-98).
// Replaces TOPO::Topology / TOPO::build_topology
// (include/2_topology/Topology.h, src/2_topology/TopologyBuilder.cpp:436-462).
#pragma once

#include <vector>

#include "../1_grid/1_MPCNS_Grid.h"

namespace TOPO {

struct Topology {
    std::vector<int> bc;         // 6 per block: MERCURY_BC_* codes, -xi,+xi,-eta,+eta,-zeta,+zeta
    std::vector<int> nbr;        // 6 per block: neighbour block of a MERCURY_BC_INTERFACE face, else -1
    std::vector<int> transform;  // 54 per block: IndexTransform of each face
    // patch connections instead of per-face ones (rotated, mirrored, split faces):
    // 18 ints per patch as mercury_create_patched takes them; rule 0 is the reference's halo
    std::vector<int> patches;
    int rule = 0;
    int num_blocks() const { return (int)bc.size() / 6; }
};

// From the grid's CASE files (myid and dimension are kept for the reference's
// signature: the files of every rank are already merged, and the path is 3-D).
inline Topology build_topology(const Grid &grid, int myid = 0, int dimension = 3) {
    (void)myid;
    (void)dimension;
    if (!grid.case_files()) throw std::runtime_error("build_topology: the grid was not read from a CASE directory");
    Topology t;
    const int nb = grid.num_blocks();
    t.bc.assign(6 * nb, 0);
    t.nbr.assign(6 * nb, -1);
    t.transform.assign(54 * nb, 0);
    mercury_hip_detail::check(mercury_case_topology(grid.case_files(), t.bc.data(), t.nbr.data(), t.transform.data()),
                              "build_topology");
    return t;
}

// Any CASE topology as patches (inner_patches of TOPO::build_topology,
// src/2_topology/TopologyBuilder.cpp:64-260): split, rotated and mirrored
// connections included.
inline Topology build_patch_topology(const Grid &grid, int rule = 0) {
    if (!grid.case_files()) throw std::runtime_error("build_patch_topology: the grid was not read from a CASE directory");
    Topology t;
    const int nb = grid.num_blocks();
    t.bc.assign(6 * nb, 0);
    t.nbr.assign(6 * nb, -1);
    t.transform.assign(54 * nb, 0);
    t.rule = rule;
    int np = 0;
    mercury_hip_detail::check(mercury_case_patch_topology(grid.case_files(), t.bc.data(), 0, &np, nullptr), "build_patch_topology");
    t.patches.assign(18 * (size_t)np, 0);
    mercury_hip_detail::check(mercury_case_patch_topology(grid.case_files(), t.bc.data(), np, &np, t.patches.data()),
                              "build_patch_topology");
    return t;
}

// One code for every face of every block (no connections), e.g. all periodic.
inline Topology uniform_topology(int nblocks, int code) {
    Topology t;
    t.bc.assign(6 * nblocks, code);
    t.nbr.assign(6 * nblocks, -1);
    t.transform.assign(54 * nblocks, 0);
    return t;
}

}  // namespace TOPO
