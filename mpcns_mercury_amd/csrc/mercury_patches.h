// Patch connections: rotated, mirrored and partial-face block interfaces.
//
// Host-only plan building.  A patch joins a node box on one block face to a
// node box of a neighbour face through an IndexTransform (neighbour node
// [perm[a]] = sign[a] * node[a] + offset[a], include/2_topology/
// TopologyTypes.h:93-98).  From the patches this module derives what one sync
// of a field group does to the ghosts (every ghost DOF written, with the
// start-of-halo value it ends up holding and the sign it carries) and the
// owner -> alias copies of shared face and edge DOFs, as flat entry lists the
// device copy kernels apply.
#pragma once
#include <array>
#include <vector>

namespace merc {
namespace patch {

struct Patch {
    int block, dir, lo[3], hi[3], nb, perm[3], sign[3], off[3];
};
constexpr int kPatchInts = 18;

// block sizes, ghost depth, face codes and the patch list of a context
struct Topo {
    int g = 0;
    std::vector<std::array<int, 3>> m;
    std::vector<int> bc;  // 6 per block (MercBC codes)
    std::vector<Patch> P;
    int rule = 0;  // 0: staggered DOFs mapped as node indices (the reference), 1: by position
    int nb() const { return (int)m.size(); }
    // throws std::runtime_error naming the first problem
    void validate() const;
};
Topo make_topo(int nb, const int *dims, int g, const int *bc, int npatch, const int *patches, int rule);

enum Kind { KIND_COMPONENT = 0, KIND_FACE = 1, KIND_EDGE = 2 };

// (field, block, i, j, k) of a DOF: field is the axis within a face or edge
// triplet (0 for a component-copy field)
struct Dof {
    int f, b, i, j, k;
};
struct Entry {
    Dof dst, src;
    int sign;
};

// Every ghost DOF one sync's halo writes for a group of `kind` up to `level`
// (1: face strips, 3: edges and vertices too), with the DOF whose value at the
// start of the halo it ends up holding.  A source may itself be a destination
// (the node-index rule reads ghosts of the same level): the entries must then
// be applied with every read before every write.
std::vector<Entry> halo_entries(const Topo &T, int kind, int level);
// owner -> alias copies of the shared DOFs of a face (KIND_FACE) or edge
// (KIND_EDGE) triplet group
std::vector<Entry> owner_entries(const Topo &T, int kind);

// classes of shared edge DOFs (two or more members): members in ascending
// (block, i, j, k, axis) order, the first is the owner; per member the axis and
// the orientation sign
struct EdgeClasses {
    std::vector<int> start{0};
    std::vector<Dof> mem;  // f = axis
    std::vector<int> sign;
};
EdgeClasses edge_classes(const Topo &T);
// the class representative of every owned node (nodes of one class are one
// physical node), per block in (i, j, k) order, k fastest; off[b] is block b's
// first entry (off has nb + 1 entries)
std::vector<long long> node_reps(const Topo &T, std::vector<long long> &off);
// the same for shared face DOFs (f = axis; sign +1 or -1 as the face's
// boundary-edge stencil equals or opposes the owner's)
EdgeClasses face_classes(const Topo &T);

}  // namespace patch
}  // namespace merc
