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

import numpy as np
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
        # ion species: 2 (H+, Na+) or 3 with He++ (no reference counterpart)
        ("n_species", ctypes.c_double), ("M_He", ctypes.c_double), ("state_coeff_He", ctypes.c_double),
        ("Z_He", ctypes.c_double), ("qinf_He", ctypes.c_double * 5), ("wall_rho_floor_He", ctypes.c_double),
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
    # a third ion species, He++ (an extension; the reference has H+ and Na+): its molar
    # mass and its share of the inflow's ions by number
    he_species: bool = False
    mole_mass_he: float = 0.004
    he_number_fraction: float = 0.04
    # eta profile: "window" (the reference's, MercurySolver_Tools.cpp:50-69) or "exp",
    # inver_Rem * exp(-(r - r_inner) / width) (an extension for the resistive shell case)
    mercury_eta_profile: str = "window"
    # intrinsic dipoles (num_of_dipole + LIST Dipole<i>, Z4_Mercury/9999setup/3_MHD_Control.txt:1-9):
    # dimensional x0 y0 z0 [m], mx my mz [A m^2], a [m]
    dipoles: list = field(default_factory=list)

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
    # He++: state coefficient as for the other species, charge 2; far field at the
    # inflow velocity and temperature with he_number_fraction of the H+ number density
    P.n_species = 3.0 if u.he_species else 2.0
    M_He = u.mole_mass_he
    P.M_He, P.Z_He = M_He, 2.0
    P.state_coeff_He = (R_uni * T_ref) / (M_He * U_ref * U_ref)
    P.wall_rho_floor_He = P.wall_rho_floor_H * (M_He / M_H) * u.he_number_fraction
    if ini:
        f = u.he_number_fraction
        rho_he = f * rho0 * (M_He / ini["Molecular_mass"])
        p_he = f * p0
        q = (rho_he, rho_he * u0, rho_he * v0, rho_he * w0, 0.5 * rho_he * (u0 * u0 + v0 * v0 + w0 * w0) + p_he / (gamma - 1.0))
        for m, v in enumerate(q):
            P.qinf_He[m] = v
    return P


# ---------------------------------------------------------------------------
# the reference's setup files (Param, src/0_basic/1_MPCNS_Parameter.cpp:4-434)
# ---------------------------------------------------------------------------
_SEP = set("=\r\n\t#$,;\" ()")
_TYPES = ("string", "int", "double", "bool")


def _words(line: str):
    """Split on the reference's separators: any of = \\r \\n \\t # $ , ; " ( ) and space."""
    out, cur = [], []
    for ch in line:
        if ch in _SEP:
            if cur:
                out.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
    if cur:
        out.append("".join(cur))
    return out


def _value(kind: str, text: str):
    if kind == "int":
        return int(text)
    if kind == "double":
        return float(text)
    if kind == "bool":
        return text != "0"
    return text


def read_setup(setup_dir) -> Dict[str, object]:
    """All parameters of a CASE setup directory: the files listed in ``filenames``
    (up to ###END###), read in order.  Lines are ``type name value`` (string,
    int, double, bool; a bool is true unless its value is "0"), LIST blocks
    (name and count, (int) or (double), the names, the values, the closing LIST
    line) become dicts, IF/ELSEIF/ELSE/ENDIF blocks keep the branch whose
    condition (type, an earlier parameter, a value) holds, END_OF_FILE ends a
    file, and lines starting with // or # are comments.  A later definition
    replaces an earlier one."""
    import os
    params: Dict[str, object] = {}
    with open(os.path.join(setup_dir, "filenames")) as f:
        names = []
        for tok in f.read().split():
            if tok == "###END###":
                break
            names.append(tok)
    for name in names:
        path = os.path.join(setup_dir, name)
        with open(path) as f:
            lines = f.read().split("\n")
        n = 0
        # stack of IF states: [taking this branch, some branch already taken]
        ifs = []

        def active():
            return all(t for t, _ in ifs)

        def cond(ws):
            kind, key, val = ws[1], ws[2], ws[3]
            if key not in params:
                raise KeyError(f"{path}: IF on undefined parameter {key}")
            return params[key] == _value(kind, val)

        while n < len(lines):
            line = lines[n]
            n += 1
            if not any(ch not in _SEP for ch in line) or line.startswith("//") or line.startswith("#"):
                continue
            ws = _words(line)
            kw = ws[0]
            if kw == "END_OF_FILE":
                break
            if kw == "IF":
                take = active() and cond(ws)
                ifs.append([take, take])
                continue
            if kw == "ELSEIF":
                taken = ifs[-1][1]
                ifs[-1][0] = (not taken) and all(t for t, _ in ifs[:-1]) and cond(ws)
                ifs[-1][1] = taken or ifs[-1][0]
                continue
            if kw == "ELSE":
                ifs[-1][0] = not ifs[-1][1] and all(t for t, _ in ifs[:-1])
                ifs[-1][1] = True
                continue
            if kw == "ENDIF":
                ifs.pop()
                continue
            if kw == "LIST":
                head = _words(lines[n])
                kind = _words(lines[n + 1])[0]
                keys = _words(lines[n + 2])
                vals = _words(lines[n + 3])
                n += 5  # the closing LIST line too
                count = int(head[1])
                if kind not in ("int", "double") or len(keys) < count or len(vals) < count:
                    raise ValueError(f"{path}: malformed LIST {head[0]}")
                if active():
                    params[head[0]] = {keys[t]: _value(kind, vals[t]) for t in range(count)}
                continue
            if kw not in _TYPES:
                raise ValueError(f"{path}: unknown keyword {kw}")
            if active():
                params[ws[1]] = _value(kw, ws[2])
    return params


def units_from_setup(params: Dict[str, object]) -> MercuryUnits:
    """MercuryUnits from the keys the Mercury build reads (Param::GetX aborts on a
    missing key: so does this, with the key's name)."""
    def get(key):
        if key not in params:
            raise KeyError(f"setup key {key} is missing")
        return params[key]

    u = MercuryUnits()
    u.constant = dict(get("constant"))
    u.REF = dict(get("REF"))
    u.INITIAL = dict(get("INITIAL"))
    u.mole_mass1, u.mole_mass2 = get("mole_mass1"), get("mole_mass2")
    u.CFL = get("CFL")
    u.r_min, u.r_max = get("r_min"), get("r_max")
    u.arti_eta_max, u.J_range_start, u.J_range_on = get("arti_eta_max"), get("J_range_start"), get("J_range_on")
    u.is_local_artificial_resistivity = bool(get("is_local_artificial_resistivity"))
    u.local_arti_eta_max = get("local_arti_eta_max")
    u.local_arti_center = (get("local_arti_x0"), get("local_arti_y0"), get("local_arti_z0"))
    u.local_arti_r_decay, u.local_arti_r_cutoff = get("local_arti_r_decay"), get("local_arti_r_cutoff")
    u.is_Ambipolar_Efield = bool(get("is_Ambipolar_Efield"))
    u.is_Mercury_resistance = bool(get("is_Mercury_resistance"))
    u.eta_max_mercury = get("eta_max_mercury")
    u.mercury_eta_r_inner, u.mercury_eta_r_outer = get("mercury_eta_r_inner"), get("mercury_eta_r_outer")
    u.mercury_eta_width = get("mercury_eta_width")
    u.implicit_resistive_rtol = get("implicit_resistive_ksp_rtol")
    u.implicit_resistive_atol = get("implicit_resistive_ksp_atol")
    u.implicit_resistive_max_it = get("implicit_resistive_ksp_max_it")
    u.dipoles = [dict(get(f"Dipole{i}")) for i in range(1, int(get("num_of_dipole")) + 1)]
    return u


def dipole_array(u: MercuryUnits) -> np.ndarray:
    """Nondimensional dipoles, 7 values each (x0 y0 z0 mx my mz a), as
    DipoleField::load_from_param derives them (Z4_Mercury/include/operators/Dipole.h:44-87):
    positions and softening length over L_ref, moments over
    4 * 3.1415926535 * B_ref * L_ref^3 / mu_mag (the reference's literal pi)."""
    L_ref, B_ref, mu = u.REF["L_ref"], u.REF["B_ref"], u.constant["mu_mag"]
    inv = 1.0 / (4 * 3.1415926535 * B_ref * L_ref * L_ref * L_ref / mu)
    out = np.zeros((len(u.dipoles), 7))
    for n, d in enumerate(u.dipoles):
        out[n] = (d["x0"] / L_ref, d["y0"] / L_ref, d["z0"] / L_ref, d["mx"] * inv, d["my"] * inv, d["mz"] * inv,
                  d["a"] / L_ref)
    return out
