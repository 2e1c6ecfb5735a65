// The Mercury explicit step of the drop-in C++ interface.
// Replaces MercurySolver (Z4_Mercury/include/MercurySolver.h:31-48): the
// constructor's setup (initial face fluxes, Badd with the dipoles, Sync(Badd),
// Sync(Bface), derived fields; MercurySolver.cpp:11-247) and Advance
// (MercurySolver_Driver.cpp:6-61), whose StepOnce runs on the device.  The
// reference's Param-driven stop rule (max_Nstep / max_Time) becomes an explicit
// step count; its dt abort rule (dt not finite or below 1e-7) throws.
#pragma once

#include <vector>

#include "3_field/Field.h"
#include "4_halo/Halo.h"

class MercurySolver {
public:
    // Bu: uniform induced field of the initial face flux (0 in the reference's
    // Common_Initial); dipoles: 7 nondimensional values each (x0 y0 z0 mx my mz a)
    MercurySolver(Grid *grd, TOPO::Topology *topo, Field *fld, Halo *halo, const double Bu[3] = nullptr,
                  const std::vector<double> &dipoles = {})
        : grd_(grd), topo_(topo), fld_(fld), halo_(halo) {
        if (!fld) throw std::runtime_error("MercurySolver: null field");
        if (dipoles.size() % 7 != 0) throw std::runtime_error("MercurySolver: dipoles must hold 7 values each");
        const double zero[3] = {0.0, 0.0, 0.0};
        mercury_hip_detail::check(mercury_init_bface_dipoles(fld->context(), Bu ? Bu : zero, nullptr, (int)(dipoles.size() / 7),
                                                             dipoles.empty() ? nullptr : dipoles.data()),
                                  "MercurySolver");
    }

    // The catalogue and the halo registrations are fixed in the library: these
    // keep the reference's call sequence (Z4_Mercury/main.cpp:53-66) compiling.
    static void RegisterFields(Field *, int) {}
    static void RegisterHaloFields(Field *, Halo *) {}

    // After the caller has uploaded its initial state (U_H, U_Na, B_*):
    // Sync(Badd), Sync(Bface) and UpdateDerivedFields_ (MercurySolver.cpp:141-145,
    // MercurySolver_Driver.cpp:6-19).
    void Prepare() { mercury_hip_detail::check(mercury_prepare(fld_->context()), "MercurySolver::Prepare"); }

    // nsteps explicit steps (integrator MERCURY_EULER, the reference's StepOnce,
    // or MERCURY_SSPRK2); each step's dt is appended to dt_history when given.
    void Advance(int nsteps = 1, int integrator = MERCURY_EULER, std::vector<double> *dt_history = nullptr) {
        std::vector<double> dts(dt_history ? (size_t)nsteps : 0);
        mercury_hip_detail::check(mercury_step(fld_->context(), nsteps, integrator, dt_history ? dts.data() : nullptr),
                                  "MercurySolver::Advance");
        if (dt_history) dt_history->insert(dt_history->end(), dts.begin(), dts.end());
    }

    // calc_divB (MercurySolver_MagneticCell.cpp:151-241): max|d2 B| / max|B_face|
    double NormalisedDivB() const {
        double n = 0.0, m = 0.0;
        mercury_hip_detail::check(mercury_divb(fld_->context(), &n, &m), "MercurySolver::NormalisedDivB");
        return n;
    }

    Field *field() const { return fld_; }
    Halo *halo() const { return halo_; }
    Grid *grid() const { return grd_; }
    TOPO::Topology *topology() const { return topo_; }

private:
    Grid *grd_ = nullptr;
    TOPO::Topology *topo_ = nullptr;
    Field *fld_ = nullptr;
    Halo *halo_ = nullptr;
};
