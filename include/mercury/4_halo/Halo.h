// Ghost synchronisation of the drop-in C++ interface.
// Replaces Halo (include/4_halo/Halo.h:15-68) as the Mercury solver drives it:
// sync_group(name) runs the whole staged Sync_ of that group on the device --
// face, edge and vertex halo copies across periodic seams and block
// connections, the physical handlers and corner defaults, and the owner ->
// alias copies (MercuryBoundary::Sync_, Z4_Mercury/src/boundary/Boundary_Sync_.cpp:4-110).
#pragma once

#include <string>

#include "../3_field/Field.h"

class Halo {
public:
    Halo(Field *field, TOPO::Topology *topo) : field_(field), topo_(topo) {
        if (!field) throw std::runtime_error("Halo: null field");
    }
    // group names of MercuryBoundary::Sync (Z4_Mercury/src/boundary/Boundary_Build.cpp:564-703)
    void sync_group(const std::string &group_name) {
        static const char *const names[][2] = {{"Ucell", "sync_ucell"}, {"Bface", "sync_bface"}, {"Eedge", "sync_eedge"},
                                               {"Ehall", "sync_ehall"}, {"Jedge", "sync_jedge"}, {"B_cell", "sync_bcell"},
                                               {"J_cell", "sync_jcell"}, {"Badd", "sync_badd"}};
        for (const auto &n : names)
            if (group_name == n[0]) {
                mercury_hip_detail::check(mercury_op(field_->context(), n[1], 0.0), "Halo::sync_group");
                return;
            }
        throw std::runtime_error("Halo::sync_group: unknown group " + group_name);
    }
    const TOPO::Topology *topology() const { return topo_; }

private:
    Field *field_ = nullptr;
    TOPO::Topology *topo_ = nullptr;
};
