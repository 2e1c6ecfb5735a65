// Host-side static geometry of one block (see mercury_geometry.cpp).
#pragma once
#include <array>
#include <vector>

#include "mercury_common.h"

namespace merc {

// receives finished per-slot arrays (box layout of one block)
struct SlotWriter {
    virtual void put(int slot, const std::vector<double> &v) = 0;
    virtual ~SlotWriter() = default;
};

void jacobi_eigen3(const double M[3][3], double V[3][3], double d[3]);

struct BlockGeom {
    int m[3], g, N[3], Nn[3];
    long long V, Vn;
    std::array<std::vector<double>, 3> X;    // node coordinates, ghost depth g+1
    std::array<std::vector<double>, 9> JD_;  // face area vectors (kept for the initial face fluxes)

    BlockGeom(const int m_[3], int g_, const double *owned_nodes);
    long long idx(int i, int j, int k) const { return ((long long)(i + g) * N[1] + (j + g)) * N[2] + (k + g); }
    long long nidx(int i, int j, int k) const { return ((long long)(i + g + 1) * Nn[1] + (j + g + 1)) * Nn[2] + (k + g + 1); }
    std::array<double, 3> node(int i, int j, int k) const;
    void add_ghost_mesh();
    void compute(const MercPhys &P, SlotWriter &out);
    // dipoles' face flux curl(A.dr) on the face boxes of ghost depth 1, added to badd (3 face arrays)
    void add_dipole_flux(int nd, const double *dip, std::array<std::vector<double>, 3> &badd) const;
};

}  // namespace merc
