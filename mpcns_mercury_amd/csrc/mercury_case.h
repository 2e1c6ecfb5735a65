// Reader of a CASE geometry directory (host only): the grid, physical-boundary,
// inner-connection and parallel-connection files of every rank, merged into one
// list of blocks (global block = rank-major order), with the index transform of
// every connection patch built the way the reference's topology builder does.
#pragma once
#include <string>
#include <vector>

namespace merc {
namespace casefile {

struct Block {
    std::string name;
    int rank = 0, local = 0;
    int nn[3] = {0, 0, 0};           // node counts (imax, jmax, kmax)
    std::vector<double> xyz[3];      // file order: i fastest, then j, then k
};

// Physical_Boundary after Grid::Read_Phy_Boundary / Read_Phy_Name
// (src/1_grid/1_MPCNS_Grid.cpp:232-336): closed 0-based node range
struct Physical {
    int block = 0;
    int sub[3], sup[3];
    int direction = 0;  // +-(axis+1): high / low face
    int bc_id = 0;
    std::string bc_name;
};

// an inner or parallel connection patch (TOPO::InterfacePatch,
// src/2_topology/TopologyBuilder.cpp:67-160)
struct Interface {
    int kind = 0;  // 0 inner, 1 parallel
    bool periodic = false;
    int block = 0, nb_block = 0;
    int sub[3], sup[3], nb_sub[3], nb_sup[3];  // 0-based, signs kept (Pre_process)
    int lo[3], hi[3], nb_lo[3], nb_hi[3];      // half-open node boxes
    int direction = 0, nb_direction = 0;
    int perm[3], sign[3], offset[3];          // neighbour node[perm[a]] = sign[a] * node[a] + offset[a]
};

struct Geometry {
    std::vector<Block> blocks;
    std::vector<Physical> physical;
    std::vector<Interface> interfaces;
    std::vector<std::string> bc_names;  // by id (boundary_name.fvbnd)

    // nrank <= 0: as many ranks as consecutive grid files exist from rank 0
    static Geometry read(const std::string &dir, int nrank, bool binary);

    // per block face (6 per block, -xi,+xi,-eta,+eta,-zeta,+zeta): one physical
    // patch or one connection patch covering the whole face; codes of
    // include/mercury_hip.h, neighbour and IndexTransform (perm, sign, offset)
    void face_topology(int *bc, int *nbr, int *transform) const;
    // face codes (interface where connection patches lie, the physical code of a
    // whole-face Farfield / Solid_Surface patch) and every connection patch as 18
    // ints (block, direction, lo[3], hi[3], neighbour, perm[3], sign[3], offset[3])
    void patch_topology(int *bc, std::vector<int> &patches) const;
};

}  // namespace casefile
}  // namespace merc
