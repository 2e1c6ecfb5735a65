// Grid of the drop-in C++ interface: the blocks' owned node coordinates and
// cell counts, read from a CASE geometry directory or given directly.
// Replaces Grid::Grid_Preprocess (include/1_grid/1_MPCNS_Grid.h:152-226), whose
// reading of the grid, boundary and inner/parallel files runs in
// mercury_case_read; the ghost nodes and metrics are built by the device
// context (Field) from the owned nodes.
#pragma once

#include <string>
#include <vector>

#include "../mercury_check.h"

class Grid {
public:
    Grid() = default;
    ~Grid() {
        if (case_) mercury_case_destroy(case_);
    }
    Grid(const Grid &) = delete;
    Grid &operator=(const Grid &) = delete;

    // Grid_Preprocess: every rank's files of a CASE geometry directory, merged in
    // rank-major block order (nrank <= 0 counts the grid files present).
    void Grid_Preprocess(const std::string &geometry_dir, int nrank = 0, bool binary = true) {
        using mercury_hip_detail::check;
        MercuryCase *c = nullptr;
        check(mercury_case_read(&c, geometry_dir.c_str(), nrank, binary ? 1 : 0), "Grid_Preprocess");
        if (case_) mercury_case_destroy(case_);
        case_ = c;
        int nb = 0;
        check(mercury_case_blocks(c, &nb, nullptr), "Grid_Preprocess");
        dims_.assign(3 * nb, 0);
        check(mercury_case_blocks(c, &nb, dims_.data()), "Grid_Preprocess");
        nodes_.assign(node_count(), 0.0);
        check(mercury_case_nodes(c, nodes_.data()), "Grid_Preprocess");
    }

    // blocks given directly: dims 3 ints per block, owned nodes as mercury_create takes them
    void SetBlocks(const std::vector<int> &dims, const std::vector<double> &owned_nodes) {
        if (dims.size() % 3 != 0 || dims.empty()) throw std::runtime_error("Grid::SetBlocks: dims must hold 3 ints per block");
        dims_ = dims;
        if (owned_nodes.size() != node_count()) throw std::runtime_error("Grid::SetBlocks: owned node count does not match dims");
        nodes_ = owned_nodes;
    }

    int num_blocks() const { return (int)dims_.size() / 3; }
    const int *dims(int b) const { return &dims_[3 * b]; }
    const std::vector<int> &all_dims() const { return dims_; }
    const std::vector<double> &owned_nodes() const { return nodes_; }
    const MercuryCase *case_files() const { return case_; }

private:
    size_t node_count() const {
        size_t n = 0;
        for (size_t b = 0; b + 2 < dims_.size(); b += 3)
            n += 3 * (size_t)(dims_[b] + 1) * (dims_[b + 1] + 1) * (dims_[b + 2] + 1);
        return n;
    }
    MercuryCase *case_ = nullptr;
    std::vector<int> dims_;
    std::vector<double> nodes_;
};
