// Singular physical edges of a patch-connected context: the edges where three
// (or more than four) block faces meet around one physical edge, as along the
// eight corner lines of a cubed sphere.  Their EMF and current are assembled
// from the cells and faces around every alias instead of the block-local
// stencils (SingularEdgeRegistry, src/7_metric/SingularEdgeRegistry.cpp:78-332).
// Host-only registry building; the device assembly is in mercury_singular_kernels.hip.
#pragma once
#include <array>
#include <vector>

#include "mercury_common.h"
#include "mercury_patches.h"

namespace merc {
namespace singular {

// static geometry of one block on its box (ghost depth g, k fastest)
struct BlockGeo {
    int m[3], g;
    const double *dl[3], *dr[9], *jac, *cc[3], *area[3];
    long long idx(int i, int j, int k) const {
        return ((long long)(i + g) * (m[1] + 1 + 2 * g) + (j + g)) * (m[2] + 1 + 2 * g) + (k + g);
    }
};

struct Owner {
    patch::Dof dof;  // f = axis
    int sign;        // the owner alias's orientation
    double ih;       // inverse Hodge scale: mean alias length / dual area
    double cv[3];    // orientation-signed mean of the aliases' edge vectors
    int c0, c1, f0, f1;  // its cells and faces
};
struct Cell {
    patch::Dof dof;  // f = axis of the alias it was found from
    int orient;      // that alias's orientation
    double w;        // |Jac| / sum over the entry
};
struct Face {
    patch::Dof dof;  // f = face axis
    double coef;     // inverse Hodge * alias orientation * incidence
};
struct Registry {
    std::vector<Owner> e;
    std::vector<Cell> c;
    std::vector<Face> f;
    // owned aliases whose alpha (inverse Hodge 2-form -> 1-form) the entry replaces
    std::vector<patch::Dof> alpha_dof;
    std::vector<double> alpha_val;
};

// points a BlockGeo at 19 box arrays: dl (3), dr (9), Jac, cell centre (3), Area (3)
void set_block_geo(BlockGeo &g, const std::array<int, 3> &m, int ghost, const std::vector<double> *arrays);
// HallRadialTaper_ at a point (Z4_Mercury/src/physics/MercurySolver_MagExplicit_Hall.cpp:6-13)
double hall_taper(const MercPhys &P, double x, double y, double z);
// the local artificial resistivity of a singular edge's cell at its centre
// (MercurySolver_MagExplicit_Ideal.cpp:115-124): false where it does not apply
bool local_eta(const MercPhys &P, double x, double y, double z, double &eta);

// Tecplot node stencils of the singular edges' end nodes
// (IOModule::BuildTecplotSingularNodeStencils_, src/6_io/IOModule_Tecplot.cpp:430-526):
// per physical node the cells of every singular edge ending there, in
// (block, i, j, k) order, and every owned alias of the node
struct NodeStencils {
    std::vector<int> start{0};
    std::vector<patch::Dof> cells;
    std::vector<patch::Dof> alias;  // f = stencil index
};
NodeStencils node_stencils(const patch::Topo &T, const Registry &R);

// throws std::runtime_error for an entry the reference's validate_or_abort
// rejects (no length, no dual area, no cell)
Registry build(const patch::Topo &T, const std::vector<BlockGeo> &G);

}  // namespace singular
}  // namespace merc
