"""Physical constants of the Mercury explicit step.

The derivations follow ``MercurySolver::calc_physical_constant``
(Z4_Mercury/src/physics/MercurySolver_Tools.cpp:3-48), the coefficients of
``AddSourceToRHS_Fluid`` (Z4_Mercury/src/physics/MercurySolver_FluidRHS.cpp:206-231),
the ambipolar pressure coefficient (Z4_Mercury/src/physics/MercurySolver_MagExplicit_Ambipo.cpp:8-9)
and the far-field state of ``BuildMercuryBackgroundState``
(Z4_Mercury/include/0_BackgroundState.h:18-84).  Key names are the reference's
setup keys (Z4_Mercury/9999setup/*).

``MercuryPhysics`` is the flat, C-layout struct that crosses the C ABI
(``include/mercury_hip.h``); its field order is part of that ABI.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict

# rho_small of Z4_Mercury/include/00_Mercury_Const.h:3
RHO_SMALL = 1e-6


class MercuryPhysics(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double),
        ("M_H", ctypes.c_double), ("M_Na", ctypes.c_double), ("M_ref", ctypes.c_double),
        ("state_coeff_H", ctypes.c_double), ("state_coeff_Na", ctypes.c_double),
        ("T_ref", ctypes.c_double),
        ("inver_MA2", ctypes.c_double),
        ("hall_coef", ctypes.c_double),
        ("ne_floor", ctypes.c_double),
        ("c_ind", ctypes.c_double), ("c_hall", ctypes.c_double),
        ("a1_Na", ctypes.c_double), ("a4", ctypes.c_double), ("a5", ctypes.c_double), ("a6", ctypes.c_double),
        ("Tn0", ctypes.c_double),
        ("nu_lit", ctypes.c_double), ("nu_shadow", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("taper_rmin", ctypes.c_double), ("taper_rmax", ctypes.c_double),
        ("arti_eta_max", ctypes.c_double), ("J_start", ctypes.c_double), ("J_on", ctypes.c_double),
        ("local_enabled", ctypes.c_double), ("local_eta_max", ctypes.c_double),
        ("local_center", ctypes.c_double * 3), ("local_r_decay", ctypes.c_double),
        ("local_r_cutoff", ctypes.c_double),
        ("ambipolar_enabled", ctypes.c_double), ("pressure_coef", ctypes.c_double),
        ("qinf", ctypes.c_double * 5), ("qinfs", ctypes.c_double * 5),
        ("q_pv_inf", ctypes.c_double * 5), ("B_imf", ctypes.c_double * 3),
        ("wall_rho_floor_H", ctypes.c_double), ("wall_rho_floor_Na", ctypes.c_double),
        ("wall_p_floor", ctypes.c_double),
        # implicit resistive correction (MercurySolver.cpp:173-184, MercurySolver_Tools.cpp:44)
        ("resist_enabled", ctypes.c_double), ("inver_Rem", ctypes.c_double),
        ("resist_r_inner", ctypes.c_double), ("resist_r_outer", ctypes.c_double), ("resist_width", ctypes.c_double),
        ("resist_rtol", ctypes.c_double), ("resist_atol", ctypes.c_double), ("resist_max_it", ctypes.c_double),
        ("resist_profile", ctypes.c_double),
    ]


def _default_constant() -> Dict[str, float]:
    # Z4_Mercury/9999setup/0_General_Setting.txt: LIST constant
    return {"gamma": 1.6667, "NA": 6.02e23, "R_uni": 8.314, "mu_mag": 1.256637061e-6, "q_e": 1.602e-19}


def _default_ref() -> Dict[str, float]:
    # Z4_Mercury/9999setup/1_Initial_condition.txt: LIST REF
    return {"U": 4.031e5, "T": 1.7407e5, "n": 4e7, "Molecular_mass": 0.001, "B_ref": 1e-9, "L_ref": 2.44e6}


@dataclass
class MercuryUnits:
    """The setup keys the explicit step reads (reference key names)."""

    constant: Dict[str, float] = field(default_factory=_default_constant)
    REF: Dict[str, float] = field(default_factory=_default_ref)
    INITIAL: Dict[str, float] = field(default_factory=dict)
    mole_mass1: float = 0.001
    mole_mass2: float = 0.023
    CFL: float = 0.4
    r_min: float = 1.1
    r_max: float = 1.2
    arti_eta_max: float = 0.0
    J_range_start: float = 800.0
    J_range_on: float = 3000.0
    is_local_artificial_resistivity: bool = False
    local_arti_eta_max: float = 0.1
    local_arti_center: tuple = (1.53, 0.0, 0.25)
    local_arti_r_decay: float = 0.05
    local_arti_r_cutoff: float = 0.05
    is_Ambipolar_Efield: bool = False
    hall_coef_override: float | None = None
    # interior resistivity (Z4_Mercury/9999setup/3_MHD_Control.txt:12-22; eta_max_mercury
    # in Omega m).  The iteration tolerances are those of the build's CG solve.
    is_Mercury_resistance: bool = False
    eta_max_mercury: float = 0.0
    mercury_eta_r_inner: float = 0.84
    mercury_eta_r_outer: float = 1.04
    mercury_eta_width: float = 0.02
    implicit_resistive_rtol: float = 1.0e-10
    implicit_resistive_atol: float = 1.0e-30
    implicit_resistive_max_it: int = 500
    # eta profile: "window" (the reference's, MercurySolver_Tools.cpp:50-69) or "exp",
    # inver_Rem * exp(-(r - r_inner) / width) (an extension for the resistive shell case)
    mercury_eta_profile: str = "window"

    def alfven_units(self) -> "MercuryUnits":
        """Choose B_ref so that inver_MA2 = c_hall = 1 (B in Alfven units)."""
        c, r = self.constant, self.REF
        rho_ref = r["Molecular_mass"] * r["n"] / c["NA"]
        r["B_ref"] = r["U"] * math.sqrt(c["mu_mag"] * rho_ref)
        return self


def derive(u: MercuryUnits) -> MercuryPhysics:
    """calc_physical_constant + BuildMercuryBackgroundState, restated."""
    c, r, ini = u.constant, u.REF, u.INITIAL
    P = MercuryPhysics()
    gamma = c["gamma"]
    NA, R_uni, q_e, mu0 = c["NA"], c["R_uni"], c["q_e"], c["mu_mag"]
    k_B = R_uni / NA
    U_ref, L_ref, B_ref, T_ref, n_ref, M_ref = r["U"], r["L_ref"], r["B_ref"], r["T"], r["n"], r["Molecular_mass"]
    rho_ref = M_ref * n_ref / NA
    M_H, M_Na = u.mole_mass1, u.mole_mass2
    m_Na = M_Na / NA
    P.gamma = gamma
    P.M_H, P.M_Na, P.M_ref = M_H, M_Na, M_ref
    P.state_coeff_H = (R_uni * T_ref) / (M_H * U_ref * U_ref)
    P.state_coeff_Na = (R_uni * T_ref) / (M_Na * U_ref * U_ref)
    P.T_ref = T_ref
    P.inver_MA2 = B_ref * B_ref / (U_ref * U_ref * mu0 * rho_ref)
    hall = B_ref * M_ref / (U_ref * q_e * L_ref * mu0 * rho_ref * NA)
    P.hall_coef = hall if u.hall_coef_override is None else u.hall_coef_override
    P.ne_floor = 0.01
    P.c_ind = (q_e * L_ref * B_ref * n_ref) / (rho_ref * U_ref)
    P.c_hall = (B_ref * B_ref) / (mu0 * rho_ref * U_ref * U_ref)
    P.a1_Na = (L_ref / U_ref) * (m_Na / rho_ref) * 1e6
    P.a4 = L_ref / U_ref
    P.a5 = (3.0 * L_ref * k_B * n_ref) / (rho_ref * U_ref * U_ref * U_ref)
    P.a6 = (1.0e6 * L_ref * k_B) / (rho_ref * U_ref * U_ref * U_ref * (gamma - 1.0))
    P.Tn0 = 1000.0
    P.nu_lit, P.nu_shadow = 5.0e-5, 1.0e-5
    P.cfl = u.CFL
    P.taper_rmin, P.taper_rmax = u.r_min, u.r_max
    P.arti_eta_max, P.J_start, P.J_on = u.arti_eta_max, u.J_range_start, u.J_range_on
    P.local_enabled = 1.0 if u.is_local_artificial_resistivity else 0.0
    P.local_eta_max = u.local_arti_eta_max
    for a in range(3):
        P.local_center[a] = u.local_arti_center[a]
    P.local_r_decay, P.local_r_cutoff = u.local_arti_r_decay, u.local_arti_r_cutoff
    P.ambipolar_enabled = 1.0 if u.is_Ambipolar_Efield else 0.0
    P.pressure_coef = (rho_ref * U_ref) / (q_e * L_ref * B_ref * n_ref)
    # far-field state (0_BackgroundState.h:18-84)
    if ini:
        P.B_imf[0], P.B_imf[1], P.B_imf[2] = ini["Bx"] / B_ref, ini["By"] / B_ref, ini["Bz"] / B_ref
        rho0 = (ini["n"] / n_ref) * (ini["Molecular_mass"] / M_ref)
        cy, cz = ini["c_y"], ini["c_z"]
        cx = -math.sqrt(1.0 - cy * cy - cz * cz)
        u0, v0, w0 = cx * ini["U"] / U_ref, cy * ini["U"] / U_ref, cz * ini["U"] / U_ref
        p0 = ini["n"] * k_B * ini["T"] / (rho_ref * U_ref * U_ref)
        T0 = ini["T"] / T_ref
        for m, v in enumerate((u0, v0, w0, p0, T0)):
            P.q_pv_inf[m] = v
        q = (rho0, rho0 * u0, rho0 * v0, rho0 * w0, 0.5 * rho0 * (u0 * u0 + v0 * v0 + w0 * w0) + p0 / (gamma - 1.0))
        for m, v in enumerate(q):
            P.qinf[m] = v
        pvs = (0.0, 0.0, 0.0, p0 * RHO_SMALL / 23.0, T0)
        qs0 = RHO_SMALL * rho0
        qs = (qs0, 0.0, 0.0, 0.0, pvs[3] / (gamma - 1.0) + 0.5 * qs0 * (pvs[0] ** 2 + pvs[1] ** 2 + pvs[2] ** 2))
        for m, v in enumerate(qs):
            P.qinfs[m] = v
    # wall floors of BC_Solid_Surface_ (Boundary_Condition_Fluid.cpp:58-60)
    P.wall_rho_floor_H, P.wall_rho_floor_Na, P.wall_p_floor = 2.5e-3, 1.0e-12, 1.0e-12
    # interior resistivity: inver_Rem = eta_max_mercury / (mu0 U_ref L_ref) (MercurySolver_Tools.cpp:44)
    P.resist_enabled = 1.0 if u.is_Mercury_resistance else 0.0
    P.inver_Rem = u.eta_max_mercury / (mu0 * U_ref * L_ref)
    P.resist_r_inner, P.resist_r_outer, P.resist_width = u.mercury_eta_r_inner, u.mercury_eta_r_outer, u.mercury_eta_width
    P.resist_rtol, P.resist_atol = u.implicit_resistive_rtol, u.implicit_resistive_atol
    P.resist_max_it = float(u.implicit_resistive_max_it)
    if u.mercury_eta_profile not in ("window", "exp"):
        raise ValueError(f"unknown eta profile {u.mercury_eta_profile!r}")
    P.resist_profile = 1.0 if u.mercury_eta_profile == "exp" else 0.0
    return P
