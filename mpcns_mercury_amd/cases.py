"""Synthetic cases of the Mercury explicit step (grids, topology, seeded initial states).

The reference reads its grids from plot3D files and injects initial states by
restart (src/1_grid/1_MPCNS_Grid.cpp:125-230, src/6_io/IOModule_Bin.cpp:234-420).
No data files exist for the benchmark configurations, so this module builds the
same inputs in memory from a counter-based PRNG (splitmix64), identically for
the HIP path and for the CPU oracle:

* owned node coordinates per block (ghost nodes are extrapolated by the solver,
  as ``Block::Add_ghostmesh`` does);
* the face boundary codes of every block (periodic self-connection, far field,
  the fluid side of the solid planet surface, or a 1:1 interface to a
  neighbour block given by the neighbour and a node-index translation);
* the conserved fluid state of both ion species on the whole allocated cell box
  (ghosts included, as a restart file holds them);
* an edge vector potential ``A`` whose discrete curl is the divergence-free
  face-B perturbation, plus a uniform part ``Bu`` (face flux ``Bu . JD``);
* the static photo-ionisation source field.

Layout of every per-block array: the uniform box [-g, m+g] per axis, k fastest.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .params import MercuryPhysics, MercuryUnits, derive, dipole_array, RHO_SMALL

BC_PERIODIC, BC_FARFIELD, BC_WALL, BC_INTERFACE, BC_EXTERNAL = 0, 1, 2, 3, 4

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser, vectorised over uint64 (wrap-around arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, key: np.ndarray) -> np.ndarray:
    """Deterministic U(-1, 1) per integer key."""
    h = splitmix64(np.uint64(seed) ^ splitmix64(np.uint64(stream) ^ splitmix64(key.astype(np.uint64))))
    return ((h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0


@dataclass
class Block:
    dims: tuple  # cells (mx, my, mz)
    bc: tuple  # 6 face codes: -xi, +xi, -eta, +eta, -zeta, +zeta
    nodes: np.ndarray  # (3, mx+1, my+1, mz+1)
    nbr: tuple = (-1,) * 6  # BC_INTERFACE: neighbour block of each face
    off: tuple = ((0, 0, 0),) * 6  # BC_INTERFACE: node (i,j,k) -> neighbour node (i,j,k) + off


@dataclass
class Case:
    name: str
    g: int
    blocks: List[Block]
    phys: MercuryPhysics
    fields: List[Dict[str, np.ndarray]]  # per block: name -> (ncomp, NI, NJ, NK)
    A: Optional[List[np.ndarray]]  # per block (3, NI, NJ, NK) edge potential, or None
    Bu: tuple = (0.0, 0.0, 0.0)
    steps: int = 100
    integrator: int = 0  # 0 forward Euler (reference), 1 SSP-RK2
    meta: dict = field(default_factory=dict)
    # intrinsic dipoles of Badd: (n, 7) nondimensional x0 y0 z0 mx my mz a, or None
    dipoles: Optional[np.ndarray] = None
    # patch connections (n, 18) int32: block, dir, node box lo[3] hi[3], neighbour
    # block, perm[3], sign[3], offset[3]; None: per-face connections
    patches: Optional[np.ndarray] = None
    # staggered-DOF map of patch connections: 0 the reference's node-index map,
    # 1 positions matched
    patch_rule: int = 0

    def dipole_array(self) -> np.ndarray:
        d = np.zeros((0, 7)) if self.dipoles is None else np.asarray(self.dipoles, dtype=np.float64).reshape(-1, 7)
        return np.ascontiguousarray(d)

    @property
    def nb(self) -> int:
        return len(self.blocks)

    def box(self, ib: int):
        m = self.blocks[ib].dims
        return tuple(mm + 1 + 2 * self.g for mm in m)

    def ncells(self) -> int:
        return int(sum(np.prod(b.dims) for b in self.blocks))

    def dims_array(self) -> np.ndarray:
        return np.array([b.dims for b in self.blocks], dtype=np.int32).reshape(-1)

    def bc_array(self) -> np.ndarray:
        return np.array([b.bc for b in self.blocks], dtype=np.int32).reshape(-1)

    def has_interfaces(self) -> bool:
        return any(c == BC_INTERFACE for b in self.blocks for c in b.bc)

    def nbr_array(self) -> np.ndarray:
        return np.array([b.nbr for b in self.blocks], dtype=np.int32).reshape(-1)

    def trans_array(self) -> np.ndarray:
        """IndexTransform per block face (perm[3], sign[3], offset[3]),
        include/2_topology/TopologyTypes.h:93-98: translations only."""
        t = [[0, 1, 2, 1, 1, 1, *b.off[f]] for b in self.blocks for f in range(6)]
        return np.array(t, dtype=np.int32).reshape(-1)

    def nodes_flat(self) -> np.ndarray:
        return np.concatenate([np.ascontiguousarray(b.nodes).reshape(-1) for b in self.blocks]).astype(np.float64)


def _box_index(dims, g):
    """Integer (i, j, k) index grids of the box [-g, m+g]."""
    rng = [np.arange(-g, m + g + 1) for m in dims]
    return np.meshgrid(*rng, indexing="ij")


def _conserved(rho, u, v, w, p, gamma):
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E])


# ---------------------------------------------------------------------------
# C1: single-fluid ideal MHD on one periodic curvilinear block
# ---------------------------------------------------------------------------

def set_resistivity(units: MercuryUnits, inver_Rem: float, r_inner: float, r_outer: float, width: float,
                    rtol: float = 1e-10) -> MercuryUnits:
    """Switch the interior resistivity on with the non-dimensional amplitude
    inver_Rem (eta_max_mercury = inver_Rem * mu0 * U_ref * L_ref)."""
    c, r = units.constant, units.REF
    units.is_Mercury_resistance = True
    units.eta_max_mercury = inver_Rem * c["mu_mag"] * r["U"] * r["L_ref"]
    units.mercury_eta_r_inner, units.mercury_eta_r_outer, units.mercury_eta_width = r_inner, r_outer, width
    units.implicit_resistive_rtol = rtol
    return units


def make_periodic_cube(n=(48, 48, 48), g: int = 4, seed: int = 1, amp: float = 1e-3,
                       integrator: int = 0, steps: int = 100, hall: bool = False,
                       resist: float | None = None, he: bool = False) -> Case:
    """One block mapped from the unit cube by x += 0.05 sin(2 pi y) sin(2 pi z),
    periodic on every axis; rho = 1, p = 0.1, B = (1,0,0) + curl of a random edge
    potential, seeded velocity perturbations; the Na+ fluid sits at rho_small.
    Hall is off (taper radii beyond the domain) unless ``hall``.  ``resist``
    switches the interior resistivity on with that inver_Rem on the window
    r in [0.5, 0.9] (width 0.05), which crosses the periodic seams.  ``he`` adds
    a He++ fluid at 4 % of the H+ number density with its own seeded velocity
    perturbations."""
    mx, my, mz = n
    nodes = _cube_nodes(n, (0, 0, 0), n)
    units = MercuryUnits().alfven_units()
    units.CFL = 0.4
    if hall:
        units.r_min, units.r_max = -1.0, -1.0
        units.hall_coef_override = 0.02
    else:
        units.r_min, units.r_max = 100.0, 100.0
    if resist is not None:
        set_resistivity(units, resist, 0.5, 0.9, 0.05)
    units.he_species = he
    P = derive(units)
    fields, A = _cube_state(n, (0, 0, 0), n, g, seed, amp, P.gamma, he_frac=units.he_number_fraction if he else None,
                            he_mass_ratio=units.mole_mass_he / units.mole_mass1)
    blk = Block(dims=(mx, my, mz), bc=(BC_PERIODIC,) * 6, nodes=nodes)
    return Case(name=f"periodic_cube_{mx}x{my}x{mz}" + ("_resist" if resist is not None else ""), g=g, blocks=[blk],
                phys=P, fields=[fields], A=[A], Bu=(1.0, 0.0, 0.0), steps=steps, integrator=integrator,
                meta={"seed": seed, "species": 3 if he else 2, "hall": hall, "resist": resist})


def _cube_nodes(n, origin, dims):
    """Nodes of the sub-box [origin, origin + dims] of the C1 mapping of an n-cell cube."""
    mx, my, mz = n
    i, j, k = np.meshgrid(*(np.arange(o, o + d + 1) for o, d in zip(origin, dims)), indexing="ij")
    y = j / my
    z = k / mz
    x = i / mx + 0.05 * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    return np.stack([x, y, z]).astype(np.float64)


def _cube_state(n, origin, dims, g, seed, amp, gamma, he_frac=None, he_mass_ratio=4.0):
    """Seeded state and edge potential on the box of the sub-block at origin,
    keyed by the wrapped global index (identical values at identical global
    positions whatever the block split)."""
    mx, my, mz = n
    I, J, K = _box_index(dims, g)
    I, J, K = I + origin[0], J + origin[1], K + origin[2]
    key = ((np.mod(I, mx) * my + np.mod(J, my)) * mz + np.mod(K, mz)).astype(np.uint64)
    du = [amp * uniform(seed, s, key) for s in range(3)]
    shape = I.shape
    rho = np.ones(shape)
    p = np.full(shape, 0.1)
    UH = _conserved(rho, du[0], du[1], du[2], p, gamma)
    UNa = _conserved(RHO_SMALL * rho, du[0], du[1], du[2], RHO_SMALL * p, gamma)
    h = 1.0 / mx
    A = np.stack([amp * h * uniform(seed, 10 + a, key) for a in range(3)])
    out = {"U_H": UH, "U_Na": UNa, "Photo_rate": np.zeros((1,) + shape)}
    if he_frac is not None:
        dv = [amp * uniform(seed, 30 + s, key) for s in range(3)]
        out["U_He"] = _conserved(he_frac * he_mass_ratio * rho, dv[0], dv[1], dv[2], he_frac * p, gamma)
    return out, A


def split_block_case(case: Case, cuts, name: str | None = None) -> Case:
    """Cut the single block of ``case`` into blocks glued by 1:1 interfaces.

    ``cuts[a]`` lists the cell counts along axis a (summing to the block's size,
    each at least g).  Cut faces become interfaces to the neighbouring piece; a
    periodic axis wraps from the last piece to the first through an interface;
    the other outer faces keep their codes.  Nodes, the state on the box and the
    edge potential are the original arrays sliced, so every piece holds the
    original values at its positions.  Block b = (bx*sy + by)*sz + bz."""
    if case.nb != 1:
        raise ValueError("split_block_case takes a one-block case")
    blk, g = case.blocks[0], case.g
    m = blk.dims
    cuts = [list(int(x) for x in c) for c in cuts]
    for a in range(3):
        if sum(cuts[a]) != m[a] or min(cuts[a]) < g:
            raise ValueError("cuts must sum to the block size and give every piece at least g cells")
    sx, sy, sz = (len(c) for c in cuts)
    org = [np.concatenate([[0], np.cumsum(c)[:-1]]).astype(int) for c in cuts]
    bid = lambda p: (p[0] * sy + p[1]) * sz + p[2]
    blocks, fields, A = [], [], []
    for pos in itertools.product(range(sx), range(sy), range(sz)):
        dims = tuple(int(cuts[a][pos[a]]) for a in range(3))
        o = tuple(int(org[a][pos[a]]) for a in range(3))
        bc, nbr, off = [], [], []
        for f in range(6):
            a, side = f // 2, f & 1
            q = list(pos)
            q[a] += 1 if side else -1
            inner = 0 <= q[a] < len(cuts[a])
            if not inner and blk.bc[f] == BC_PERIODIC:
                q[a] %= len(cuts[a])
                inner = True
            if inner:
                bc.append(BC_INTERFACE)
                nbr.append(bid(q))
                t = [0, 0, 0]
                t[a] = -dims[a] if side else int(cuts[a][q[a]])
                off.append(tuple(t))
            else:
                bc.append(blk.bc[f])
                nbr.append(-1)
                off.append((0, 0, 0))
        nodes = blk.nodes[:, o[0]:o[0] + dims[0] + 1, o[1]:o[1] + dims[1] + 1, o[2]:o[2] + dims[2] + 1]
        blocks.append(Block(dims=dims, bc=tuple(bc), nodes=np.ascontiguousarray(nodes), nbr=tuple(nbr), off=tuple(off)))
        box = tuple(slice(o[a], o[a] + dims[a] + 1 + 2 * g) for a in range(3))
        fields.append({k: np.ascontiguousarray(v[(slice(None),) + box]) for k, v in case.fields[0].items()})
        if case.A is not None:
            A.append(np.ascontiguousarray(case.A[0][(slice(None),) + box]))
    meta = dict(case.meta)
    meta.update({"split": (sx, sy, sz), "cuts": cuts, "origins": [o.tolist() for o in org]})
    return Case(name=name or f"{case.name}_split{sx}x{sy}x{sz}", g=g, blocks=blocks, phys=case.phys, fields=fields,
                A=A if case.A is not None else None, Bu=case.Bu, steps=case.steps, integrator=case.integrator, meta=meta,
                dipoles=case.dipoles)


def case_from_geometry(geo, template: Case, name: str | None = None) -> Case:
    """A case on the blocks of a CASE geometry directory (``hip.CaseGeometry``):
    dimensions, node coordinates and face topology come from the files; the
    physics, the fields on every block box, the edge potential and the uniform
    field come from ``template``, whose blocks must have the same sizes in the
    same order.  Connections must be translations (the solver's restriction)."""
    if geo.nb != template.nb or any(tuple(geo.dims[b]) != tuple(template.blocks[b].dims) for b in range(geo.nb)):
        raise ValueError("the template's blocks differ from the geometry's")
    bc, nbr, tr = geo.topology()
    blocks = []
    for b in range(geo.nb):
        offs = []
        for f in range(6):
            t = tr[9 * (6 * b + f):9 * (6 * b + f + 1)]
            if bc[6 * b + f] == BC_INTERFACE and (tuple(t[:3]) != (0, 1, 2) or tuple(t[3:6]) != (1, 1, 1)):
                raise ValueError(f"block {b} face {f}: rotated or mirrored connection (perm {tuple(t[:3])}, sign {tuple(t[3:6])})")
            offs.append(tuple(int(x) for x in t[6:9]))
        blocks.append(Block(dims=tuple(geo.dims[b]), bc=tuple(int(x) for x in bc[6 * b:6 * b + 6]), nodes=geo.nodes[b],
                            nbr=tuple(int(x) for x in nbr[6 * b:6 * b + 6]), off=tuple(offs)))
    return Case(name=name or f"{template.name}_case", g=template.g, blocks=blocks, phys=template.phys, fields=template.fields,
                A=template.A, Bu=template.Bu, steps=template.steps, integrator=template.integrator, meta=dict(template.meta),
                dipoles=template.dipoles)


def make_split_cube(n=(48, 48, 48), split=(2, 2, 2), g: int = 4, seed: int = 1, amp: float = 1e-3,
                    integrator: int = 0, steps: int = 100, hall: bool = False, resist: float | None = None,
                    cuts=None) -> Case:
    """The C1 cube cut into split[0] x split[1] x split[2] blocks glued by 1:1
    interfaces (the last block of each row wraps around to the first, which
    carries the periodicity).  ``cuts`` optionally gives the cell counts per axis;
    by default each axis is divided evenly."""
    base = make_periodic_cube(n, g=g, seed=seed, amp=amp, integrator=integrator, steps=steps, hall=hall, resist=resist)
    if cuts is None:
        cuts = []
        for a in range(3):
            q, r = divmod(n[a], split[a])
            cuts.append([q + (1 if t < r else 0) for t in range(split[a])])
    return split_block_case(base, cuts, name=base.name.replace("periodic_cube", "split_cube") +
                            "_{}x{}x{}".format(*(len(c) for c in cuts)))


# ---------------------------------------------------------------------------
# C2/C3: two-fluid Hall MHD on six spherical-shell sector blocks
# ---------------------------------------------------------------------------

C2_BLOCKS = ((96, 64, 64), (128, 48, 64), (64, 96, 80), (80, 80, 48), (112, 56, 72), (72, 72, 96))
# panel axes (centre direction, first and second tangent) of the six sectors
_PANELS = (((1, 0, 0), (0, 1, 0), (0, 0, 1)), ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
           ((0, 1, 0), (0, 0, 1), (1, 0, 0)), ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
           ((0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 0, -1), (0, 1, 0), (1, 0, 0)))


# the intrinsic dipole of the reference's example setup (Z4_Mercury/9999setup/3_MHD_Control.txt:4-9)
MERCURY_DIPOLE = {"x0": 0.0, "y0": 0.0, "z0": 4.88e5, "mx": 0.0, "my": 0.0, "mz": -2.83e19, "a": 5.0e2}


def sector_nodes(dims, panel, r0=1.0, r1=4.0, half_angle=np.deg2rad(40.0)):
    """Gnomonic sector of the shell r in [r0, r1]: xi radial, eta/zeta angular."""
    mx, my, mz = dims
    c, e1, e2 = (np.asarray(v, dtype=np.float64) for v in panel)
    r = r0 + (r1 - r0) * np.arange(mx + 1) / mx
    ta = np.tan(-half_angle + 2 * half_angle * np.arange(my + 1) / my)
    tb = np.tan(-half_angle + 2 * half_angle * np.arange(mz + 1) / mz)
    R, TA, TB = np.meshgrid(r, ta, tb, indexing="ij")
    d = c[:, None, None, None] + TA[None] * e1[:, None, None, None] + TB[None] * e2[:, None, None, None]
    d = d / np.sqrt((d * d).sum(axis=0))[None]
    return (R[None] * d).astype(np.float64)


def make_shell(blocks=C2_BLOCKS, g: int = 4, seed: int = 1, steps: int = 200, integrator: int = 0,
               scale: float = 1.0, hall_coef: float | None = None, resist: float | None = None,
               inflow: float = 1.0, imf=(0.0, 0.0, -0.2), na_fraction: float = 0.01, panels=None,
               name: str | None = None, resist_profile: str = "window", dipole: bool = False, he: bool = False,
               half_angle: float = np.deg2rad(40.0)) -> Case:
    """Two-fluid Hall MHD: H+ and Na+ (mass ratio 1:23, density ratio 1:0.01) plus
    electron pressure (ambipolar term), inflow u = (-1,0,0), IMF B = (0,0,-0.2),
    the fluid side of a perfectly conducting sphere at r = 1 on every block's
    xi- face and far field on the other five faces; seeded 1e-3 density noise.
    ``resist`` switches the interior resistivity on with that inver_Rem on the
    reference's window (r_inner 0.84, r_outer 1.04, width 0.02,
    Z4_Mercury/9999setup/3_MHD_Control.txt:12-22).

    ``he`` adds a He++ fluid (solar-wind alphas, 4 % of the H+ number density,
    the far-field state of ``derive``) with the same seeded density noise.
    ``dipole`` adds Mercury's intrinsic dipole of the reference's example setup
    (0, 0, 4.88e5 m; m = (0, 0, -2.83e19) A m^2; a = 500 m;
    Z4_Mercury/9999setup/3_MHD_Control.txt:1-9) to Badd.

    The listed block sizes admit no conformal 1:1 tiling of the shell, so every
    block is a separate gnomonic sector (no inter-block interfaces)."""
    units = MercuryUnits().alfven_units()
    Bref = units.REF["B_ref"]
    units.INITIAL = {"U": inflow * units.REF["U"], "T": units.REF["T"], "n": units.REF["n"],
                     "Molecular_mass": units.REF["Molecular_mass"], "c_y": 0.0, "c_z": 0.0,
                     "Bx": imf[0] * Bref, "By": imf[1] * Bref, "Bz": imf[2] * Bref}
    units.CFL = 0.4
    units.r_min, units.r_max = 1.1, 1.2
    units.is_Ambipolar_Efield = True
    units.is_local_artificial_resistivity = True
    units.hall_coef_override = hall_coef
    if resist is not None:
        if resist_profile == "exp":
            # eta(r) = inver_Rem * exp(-(r - 1) / 0.1): surface-peaked resistivity
            set_resistivity(units, resist, 1.0, 1.0, 0.1)
        else:
            set_resistivity(units, resist, 0.84, 1.04, 0.02)
        units.mercury_eta_profile = resist_profile
    if dipole:
        units.dipoles = [dict(MERCURY_DIPOLE)]
    units.he_species = he
    P = derive(units)
    gamma = P.gamma
    blks, fields = [], []
    for ib, dims0 in enumerate(blocks):
        dims = tuple(max(4, int(round(d * scale))) for d in dims0)
        nodes = sector_nodes(dims, _PANELS[(panels[ib] if panels is not None else ib) % 6], half_angle=half_angle)
        blks.append(Block(dims=dims, bc=(BC_WALL,) + (BC_FARFIELD,) * 5, nodes=nodes))
        I, J, K = _box_index(dims, g)
        key = (((ib * 4096 + (I + g)) * 4096 + (J + g)) * 4096 + (K + g)).astype(np.uint64)
        shape = I.shape
        rhoH = P.qinf[0] * (1.0 + 1e-3 * uniform(seed, 20, key))
        u0, v0, w0, p0 = P.q_pv_inf[0], P.q_pv_inf[1], P.q_pv_inf[2], P.q_pv_inf[3]
        UH = _conserved(rhoH, np.full(shape, u0), np.full(shape, v0), np.full(shape, w0), np.full(shape, p0), gamma)
        rhoN = na_fraction * rhoH
        UNa = _conserved(rhoN, np.full(shape, u0), np.full(shape, v0), np.full(shape, w0),
                         np.full(shape, p0 * na_fraction / 23.0), gamma)
        # cell-centre radius from the owned nodes (photo-ionisation profile of the exosphere)
        cc = np.zeros((3,) + shape)
        nd = nodes
        ccore = 0.125 * (nd[:, :-1, :-1, :-1] + nd[:, 1:, :-1, :-1] + nd[:, :-1, 1:, :-1] + nd[:, 1:, 1:, :-1] +
                         nd[:, :-1, :-1, 1:] + nd[:, 1:, :-1, 1:] + nd[:, :-1, 1:, 1:] + nd[:, 1:, 1:, 1:])
        cc[:, g:g + dims[0], g:g + dims[1], g:g + dims[2]] = ccore
        r = np.sqrt((cc * cc).sum(axis=0))
        photo = np.where(r > 1.0, 1e-2 * np.exp(-(r - 1.0) / 0.2), 0.0)
        fl = {"U_H": UH, "U_Na": UNa, "Photo_rate": photo[None]}
        if he:
            q = P.qinf_He
            rhoHe = q[0] * (1.0 + 1e-3 * uniform(seed, 21, key))
            pHe = (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]) * (P.gamma - 1.0)
            fl["U_He"] = _conserved(rhoHe, np.full(shape, q[1] / q[0]), np.full(shape, q[2] / q[0]), np.full(shape, q[3] / q[0]),
                                    np.full(shape, pHe), gamma)
        fields.append(fl)
    if name is None:
        name = "shell6" if scale == 1.0 else f"shell6_x{scale:g}"
    return Case(name=name + ("_resist" if resist is not None else ""), g=g, blocks=blks, phys=P, fields=fields,
                A=None, Bu=(0.0, 0.0, 0.0), steps=steps, integrator=integrator,
                meta={"seed": seed, "species": 3 if he else 2, "hall": True, "resist": resist, "dipole": dipole},
                dipoles=dipole_array(units) if dipole else None)


def make_uniform_block(nodes: np.ndarray, bc, g: int = 4, rho: float = 1.0, u=(0.0, 0.0, 0.0), p: float = 0.1,
                       rho_na_ratio: float = RHO_SMALL, units: MercuryUnits | None = None, name: str = "block") -> Case:
    """One block with given owned nodes (3, mx+1, my+1, mz+1) and face codes,
    a uniform state on the whole cell box, B = 0 (tests set fields directly)."""
    dims = tuple(int(n) - 1 for n in nodes.shape[1:])
    if units is None:
        units = MercuryUnits().alfven_units()
        units.r_min, units.r_max = 100.0, 100.0
    P = derive(units)
    I, J, K = _box_index(dims, g)
    shape = I.shape
    one = np.ones(shape)
    UH = _conserved(rho * one, u[0] * one, u[1] * one, u[2] * one, p * one, P.gamma)
    UNa = _conserved(rho_na_ratio * rho * one, u[0] * one, u[1] * one, u[2] * one, rho_na_ratio * p * one, P.gamma)
    fields = [{"U_H": UH, "U_Na": UNa, "Photo_rate": np.zeros((1,) + shape)}]
    return Case(name=name, g=g, blocks=[Block(dims=dims, bc=tuple(bc), nodes=nodes.astype(np.float64))], phys=P,
                fields=fields, A=None, Bu=(0.0, 0.0, 0.0), steps=1)


def cartesian_nodes(n=(6, 5, 4), h=(0.1, 0.1, 0.1), origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Owned nodes of a uniform Cartesian block."""
    i, j, k = np.meshgrid(np.arange(n[0] + 1), np.arange(n[1] + 1), np.arange(n[2] + 1), indexing="ij")
    return np.stack([origin[0] + h[0] * i, origin[1] + h[1] * j, origin[2] + h[2] * k]).astype(np.float64)


# ---------------------------------------------------------------------------
# C4: ensemble member r of independent single-fluid + Hall runs
# ---------------------------------------------------------------------------

def make_ensemble_run(r: int, scale: float = 1.0, g: int = 4, steps: int = 100) -> Case:
    """Run r of the batched ensemble (seed 1000 + r): 1-4 shell-sector blocks with
    seeded sizes in [32, 96]^3 (times ``scale``), inflow speed U(0.6, 1.4), IMF of
    strength 0.2 at angle U(0, 2 pi) in the y-z plane, Hall coefficient U(0.02, 0.1);
    single fluid (the Na+ fluid at rho_small)."""
    seed = 1000 + r
    u = lambda k: float(uniform(seed, 90 + k, np.arange(1, dtype=np.uint64))[0] * 0.5 + 0.5)  # U(0, 1)
    nblk = 1 + int(u(0) * 4) % 4
    dims = []
    for b in range(nblk):
        dims.append(tuple(int(round((32 + u(1 + 3 * b + a) * 64) * scale)) for a in range(3)))
    speed = 0.6 + 0.8 * u(20)
    ang = 2.0 * np.pi * u(21)
    hall = 0.02 + 0.08 * u(22)
    panels = [int(u(30 + b) * 6) % 6 for b in range(nblk)]
    return make_shell(blocks=tuple(dims), g=g, seed=seed, steps=steps, scale=1.0, hall_coef=hall, inflow=speed,
                      imf=(0.0, 0.2 * np.sin(ang), -0.2 * np.cos(ang)), na_fraction=RHO_SMALL, panels=panels,
                      name=f"ensemble_run{r}")


def concat_runs(runs, name: str = "ensemble") -> Case:
    """Independent runs as one case of concatenated blocks for an ensemble
    context (``hip.Mercury`` of the result calls mercury_create_ensemble): block
    interfaces are renumbered into the concatenated list; ``meta`` keeps the
    block count and the physics of every run.  The runs must share the ghost
    depth, the uniform field Bu and the dipoles."""
    if not runs:
        raise ValueError("an ensemble needs at least one run")
    g = runs[0].g
    blocks, fields, A, base = [], [], [], 0
    for c in runs:
        if c.g != g or tuple(c.Bu) != tuple(runs[0].Bu):
            raise ValueError("the runs of an ensemble share nghost and Bu")
        if not np.array_equal(c.dipole_array(), runs[0].dipole_array()):
            raise ValueError("the runs of an ensemble share the dipoles")
        for b in c.blocks:
            nbr = tuple(x + base if x >= 0 else x for x in b.nbr)
            blocks.append(Block(dims=b.dims, bc=b.bc, nodes=b.nodes, nbr=nbr, off=b.off))
        fields.extend(c.fields)
        if c.A is not None:
            A.extend(c.A)
        base += c.nb
    if any(c.A is not None for c in runs) and not all(c.A is not None for c in runs):
        raise ValueError("either every run has an edge potential or none")
    meta = {"run_nblocks": [c.nb for c in runs], "run_phys": [c.phys for c in runs], "run_names": [c.name for c in runs]}
    return Case(name=name, g=g, blocks=blocks, phys=runs[0].phys, fields=fields, A=A if runs[0].A is not None else None,
                Bu=runs[0].Bu, steps=runs[0].steps, integrator=runs[0].integrator, meta=meta, dipoles=runs[0].dipoles)


def make_split_sector(dims=(24, 20, 16), cuts=((24,), (10, 10), (16,)), **kw) -> Case:
    """One shell sector of the two-fluid Hall configuration (planet surface on
    xi-, far field on the other faces) cut into blocks glued by interfaces: the
    cut faces meet the wall and far-field faces, so the corner ghosts mix inner
    and physical patches."""
    kw.setdefault("name", "sector_{}x{}x{}".format(*dims))
    base = make_shell(blocks=(tuple(dims),), panels=[0], **kw)
    return split_block_case(base, cuts)


def make_cut_shell(split=(1, 2, 2), scale: float = 1.0, g: int = 4, **kw) -> Case:
    """The six C2 sectors, each cut into split[0] x split[1] x split[2] blocks glued
    by 1:1 interfaces (default: halves across both angular axes, 24 blocks): the
    C2 problem with inter-block ghost copies in every sync.  One run (one dt)."""
    pieces = []
    for ib, dims0 in enumerate(C2_BLOCKS):
        dims = tuple(max(2 * g, int(round(d * scale))) for d in dims0)
        base = make_shell(blocks=(dims,), panels=[ib], g=g, scale=1.0, **kw)
        cuts = []
        for a in range(3):
            q, r = divmod(dims[a], split[a])
            cuts.append([q + (1 if t < r else 0) for t in range(split[a])])
        pieces.append(split_block_case(base, cuts))
    c = concat_runs(pieces, name="shell6_cut_{}x{}x{}".format(*split))
    c.meta = {"split": tuple(split), "sectors": len(pieces)}
    return c


C5_DIMS = (256, 192, 384)


def make_resistive_shell(scale: float = 1.0, g: int = 4, steps: int = 500, integrator: int = 0,
                         inver_Rem: float = 1e-3, species: int = 3) -> Case:
    """The resistive shell configuration: one shell sector of 256 x 192 x 384 cells
    (times ``scale``) on r in [1, 4], cut into 2 x 2 blocks across the two angular
    axes and glued by interfaces, Hall MHD with H+, Na+ and (``species`` = 3, the
    default) He++, with the implicit resistive correction on every edge of the shell
    for eta(r) = inver_Rem * exp(-(r - 1) / 0.1) (``resist_profile`` "exp").  At
    scale 1: 18 874 368 cells in four blocks of 256 x 96 x 192."""
    dims = tuple(max(2 * g, 2 * int(round(d * scale / 2))) for d in C5_DIMS)
    dims = (max(g, int(round(C5_DIMS[0] * scale))),) + dims[1:]
    if species not in (2, 3):
        raise ValueError("species must be 2 or 3")
    base = make_shell(blocks=(dims,), panels=[0], g=g, steps=steps, integrator=integrator, resist=inver_Rem,
                      resist_profile="exp", name="resistive_shell_{}x{}x{}".format(*dims) + ("_he" if species == 3 else ""),
                      he=species == 3)
    cuts = ((dims[0],), (dims[1] // 2, dims[1] - dims[1] // 2), (dims[2] // 2, dims[2] - dims[2] // 2))
    return split_block_case(base, cuts, name=base.name + "_2x2")



# ---------------------------------------------------------------------------
# patch connections: rotated, mirrored and partial-face interfaces
# ---------------------------------------------------------------------------

PATCH_INTS = 18


def _patch_row(block, dirc, lo, hi, nb, perm, sign, off):
    return [int(block), int(dirc), *map(int, lo), *map(int, hi), int(nb), *map(int, perm), *map(int, sign), *map(int, off)]


def patches_of_case(case: Case) -> np.ndarray:
    """The per-face connections of a case (periodic faces and 1:1 translation
    interfaces) as whole-face patches, in block and face order."""
    rows = []
    for b, blk in enumerate(case.blocks):
        for f in range(6):
            a, side = f // 2, f & 1
            if blk.bc[f] == BC_PERIODIC:
                nb, off = b, [0, 0, 0]
                off[a] = -blk.dims[a] if side else blk.dims[a]
            elif blk.bc[f] == BC_INTERFACE:
                nb, off = blk.nbr[f], list(blk.off[f])
            else:
                continue
            lo, hi = [0, 0, 0], [d + 1 for d in blk.dims]
            lo[a] = blk.dims[a] if side else 0
            hi[a] = lo[a] + 1
            rows.append(_patch_row(b, (a + 1) * (1 if side else -1), lo, hi, nb, (0, 1, 2), (1, 1, 1), off))
    return np.array(rows, dtype=np.int32).reshape(-1, PATCH_INTS)


def patched_case(case: Case, rule: int = 0) -> Case:
    """The same case with its connections given as patches."""
    blocks = [Block(dims=b.dims, bc=tuple(BC_INTERFACE if c in (BC_PERIODIC, BC_INTERFACE) else c for c in b.bc), nodes=b.nodes)
              for b in case.blocks]
    return Case(name=case.name + "_patched", g=case.g, blocks=blocks, phys=case.phys, fields=case.fields, A=case.A, Bu=case.Bu,
                steps=case.steps, integrator=case.integrator, meta=dict(case.meta), dipoles=case.dipoles,
                patches=patches_of_case(case), patch_rule=rule)


def patches_of_geometry(geo) -> np.ndarray:
    """The connection patches of a CASE geometry (``hip.CaseGeometry``), every
    rank's, in the reader's order."""
    rows = [_patch_row(p["block"], p["direction"], p["lo"], p["hi"], p["nb_block"], p["perm"], p["sign"], p["offset"])
            for p in geo.interfaces()]
    return np.array(rows, dtype=np.int32).reshape(-1, PATCH_INTS)


def cartesian_periodic_cube(n: int = 64, g: int = 4, seed: int = 1, amp: float = 1e-3, steps: int = 10, hall: bool = False) -> Case:
    """The C1 state (seeded velocities, B = (1,0,0) + curl of a seeded edge
    potential) on the uniform Cartesian unit cube of n^3 cells, periodic."""
    base = make_periodic_cube((n, n, n), g=g, seed=seed, amp=amp, steps=steps, hall=hall)
    blk = base.blocks[0]
    blk.nodes = cartesian_nodes((n, n, n), (1.0 / n,) * 3)
    base.name = f"cartesian_cube_{n}"
    return base


def _lattice_map(nodes: np.ndarray, n: int):
    """Affine map of a Cartesian block on the n^3 lattice: global node = g0 + R @ local node."""
    X = np.rint(nodes * n).astype(np.int64)
    g0 = X[:, 0, 0, 0]
    R = np.stack([X[:, 1, 0, 0] - g0, X[:, 0, 1, 0] - g0, X[:, 0, 0, 1] - g0], axis=1)
    if sorted(np.abs(R).sum(axis=0).tolist()) != [1, 1, 1] or sorted(np.abs(R).sum(axis=1).tolist()) != [1, 1, 1]:
        raise ValueError("block axes are not lattice directions")
    return g0, R


def cube_on_geometry(geo, cube: Case, rule: int = 0, name: str | None = None) -> Case:
    """The state of a one-block periodic Cartesian cube (``cartesian_periodic_cube``)
    cut along a CASE geometry of Cartesian blocks on the same lattice (the
    reference's sample geometries): nodes and patches from the geometry; every
    cell value and every edge-potential value taken from the cube at the same
    physical cell or edge (edge values signed by the block's axis direction),
    ghosts included, so the patched run starts from the cube's state."""
    n, g = cube.blocks[0].dims[0], cube.g
    blocks, fields, A = [], [], []
    for b in range(geo.nb):
        dims = tuple(geo.dims[b])
        nodes = geo.nodes[b]
        g0, R = _lattice_map(nodes, n)
        I, J, K = _box_index(dims, g)
        P = np.stack([I, J, K]).reshape(3, -1)
        # a cell's global index is its lowest corner: local + (R < 0 along that axis)
        cell = g0[:, None] + R @ P + np.minimum(R, 0).sum(axis=1)[:, None]
        ci = np.mod(cell, n) + g
        fl = {}
        for key, arr in cube.fields[0].items():
            fl[key] = np.ascontiguousarray(arr[:, ci[0], ci[1], ci[2]].reshape((arr.shape[0],) + I.shape))
        fields.append(fl)
        Ab = np.zeros((3,) + I.shape)
        for a in range(3):
            col = R[:, a]
            c = int(np.nonzero(col)[0][0])
            s = int(col[c])
            start = g0[:, None] + R @ P + (np.minimum(col, 0))[:, None]  # lower end of the edge along physical axis c
            wi = np.mod(start, n) + g
            Ab[a] = s * cube.A[0][c][wi[0], wi[1], wi[2]].reshape(I.shape)
        A.append(Ab)
        blocks.append(Block(dims=dims, bc=(BC_INTERFACE,) * 6, nodes=np.ascontiguousarray(nodes)))
    return Case(name=name or f"{cube.name}_on_{geo.nb}_blocks", g=g, blocks=blocks, phys=cube.phys, fields=fields, A=A, Bu=cube.Bu,
                steps=cube.steps, integrator=cube.integrator, meta=dict(cube.meta), patches=patches_of_geometry(geo),
                patch_rule=rule)


def _orient_map(perm, sign, m):
    """Node-index map of a block reorientation: new[perm[a]] = sign[a] * old[a] + c[a]."""
    c = [0 if sign[a] > 0 else m[a] for a in range(3)]
    return list(perm), list(sign), c


def _reorient_box_array(arr, perm, sign, loc_delta, lead=1):
    """Values on the box [-g, m+g] of a block, rearranged to the reoriented
    block's box: along a reversed axis a DOF between nodes (delta 1) at index i
    goes to m-1-i, a node-like one to m-i (the box's extra last point of a
    cell-like axis is carried over unchanged)."""
    out = arr
    for a in range(3):
        if sign[a] > 0:
            continue
        ax = lead + a
        n = out.shape[ax]
        idx = np.arange(n)
        idx = idx[::-1].copy() if loc_delta[a] == 0 else np.r_[np.arange(n - 2, -1, -1), n - 1]
        out = np.take(out, idx, axis=ax)
    inv = [0, 0, 0]
    for a in range(3):
        inv[perm[a]] = a
    return np.ascontiguousarray(np.transpose(out, list(range(lead)) + [lead + inv[q] for q in range(3)]))


def reorient_blocks(case: Case, orient: dict, rule: int = 0, name: str | None = None) -> Case:
    """The same case with the index axes of some blocks permuted and reversed:
    orient[b] = (perm, sign) (new axis perm[a] = old axis a, direction sign[a]).
    Node coordinates, the cell fields, the edge potential (signed along reversed
    axes) and the face codes move with the axes; every connection (periodic
    faces and 1:1 interfaces) becomes a patch whose transform composes the two
    blocks' reorientations, so the connections between reoriented blocks rotate
    and mirror.  The physical problem is unchanged."""
    R = {}
    blocks, fields, A = [], [], []
    for b, blk in enumerate(case.blocks):
        perm, sign = orient.get(b, ((0, 1, 2), (1, 1, 1)))
        m = blk.dims
        R[b] = _orient_map(perm, sign, m)
        nd = [0, 0, 0]
        for a in range(3):
            nd[perm[a]] = m[a]
        nodes = _reorient_box_array(blk.nodes, perm, sign, (0, 0, 0))
        fl = {k: _reorient_box_array(v, perm, sign, (1, 1, 1)) for k, v in case.fields[b].items()}
        fields.append(fl)
        if case.A is not None:
            Ab = np.zeros((3,) + tuple(mm + 1 + 2 * case.g for mm in nd))
            for a in range(3):
                delta = [0, 0, 0]
                delta[a] = 1
                Ab[perm[a]] = sign[a] * _reorient_box_array(case.A[b][a], perm, sign, delta, lead=0)
            A.append(Ab)
        bc = [0] * 6
        for f in range(6):
            a, side = f // 2, f & 1
            nside = side if sign[a] > 0 else 1 - side
            c = blk.bc[f]
            bc[2 * perm[a] + nside] = BC_INTERFACE if c in (BC_PERIODIC, BC_INTERFACE) else c
        blocks.append(Block(dims=tuple(nd), bc=tuple(bc), nodes=nodes))
    rows = []
    for row in patches_of_case(case):
        b, nb = int(row[0]), int(row[8])
        off = row[15:18]
        pb, sb, cb = R[b]
        pn, sn, cn = R[nb]
        # new-b index q -> old-b p = Rb^-1(q) -> old-nb p + off -> new-nb Rn(p + off)
        lo, hi = row[2:5], row[5:8]
        corners = []
        for q in ([lo[0], lo[1], lo[2]], [hi[0] - 1, hi[1] - 1, hi[2] - 1]):
            r = [0, 0, 0]
            for a in range(3):
                r[pb[a]] = sb[a] * q[a] + cb[a]
            corners.append(r)
        nlo = [min(corners[0][a], corners[1][a]) for a in range(3)]
        nhi = [max(corners[0][a], corners[1][a]) + 1 for a in range(3)]
        perm, sign, offs = [0, 0, 0], [0, 0, 0], [0, 0, 0]
        for ax in range(3):  # new-b axis ax comes from old-b axis a with perm_b[a] = ax
            a = pb.index(ax)
            # old-b p[a] = sb[a] * (q[ax] - cb[a]); old-nb axis a -> new-nb axis pn[a]
            perm[ax] = pn[a]
            sign[ax] = sn[a] * sb[a]
            offs[ax] = sn[a] * (-sb[a] * cb[a] + off[a]) + cn[a]
        d = int(row[1])
        a0 = abs(d) - 1
        nd_ = pb[a0] + 1
        ndir = nd_ if (d > 0) == (sb[a0] > 0) else -nd_
        rows.append(_patch_row(b, ndir, nlo, nhi, nb, perm, sign, offs))
    return Case(name=name or f"{case.name}_reoriented", g=case.g, blocks=blocks, phys=case.phys, fields=fields,
                A=A if case.A is not None else None, Bu=case.Bu, steps=case.steps, integrator=case.integrator, meta=dict(case.meta),
                dipoles=case.dipoles, patches=np.array(rows, dtype=np.int32).reshape(-1, PATCH_INTS), patch_rule=rule)


def glue_patches(blocks, tol: float = 1e-9) -> np.ndarray:
    """Patches joining every interface face of ``blocks`` to the face of
    another block whose nodes coincide with its own (whole faces, conforming
    nodes): the IndexTransform maps each node onto the node at the same point
    and continues the block across the face into the neighbour."""
    faces = []
    for b, blk in enumerate(blocks):
        for f in range(6):
            if blk.bc[f] != BC_INTERFACE:
                continue
            a, side = f // 2, f & 1
            sl = [slice(None)] * 3
            sl[a] = blk.dims[a] if side else 0
            pts = blk.nodes[(slice(None),) + tuple(sl)].reshape(3, -1).T
            faces.append((b, f, pts))
    scale = max(np.abs(blk.nodes).max() for blk in blocks)
    key = lambda x: tuple(np.rint(np.asarray(x) / (tol * scale)).astype(np.int64))
    rows = []
    for b, f, pts in faces:
        blk = blocks[b]
        a, side = f // 2, f & 1
        corners = sorted(key(p) for p in pts[[0, -1]])
        match = None
        for nb, fn, npts in faces:
            if (nb, fn) == (b, f) or len(npts) != len(pts):
                continue
            if {key(p) for p in npts[[0, -1]]} <= {key(p) for p in npts} and set(corners) <= {key(p) for p in npts}:
                if {key(p) for p in pts} == {key(p) for p in npts}:
                    match = (nb, fn)
                    break
        if match is None:
            raise ValueError(f"block {b} face {f}: no coinciding face")
        nb, fn = match
        nblk = blocks[nb]
        where = {}
        an, sn = fn // 2, fn & 1
        for idx in np.ndindex(*(nblk.dims[t] + 1 for t in range(3))):
            if idx[an] == (nblk.dims[an] if sn else 0):
                where[key(nblk.nodes[(slice(None),) + idx])] = idx
        p0 = [0, 0, 0]
        p0[a] = blk.dims[a] if side else 0
        q0 = where[key(blk.nodes[(slice(None),) + tuple(p0)])]
        perm, sign = [0, 0, 0], [0, 0, 0]
        for t in range(3):
            if t == a:
                continue
            p1 = list(p0)
            p1[t] += 1
            d = np.subtract(where[key(blk.nodes[(slice(None),) + tuple(p1)])], q0)
            perm[t] = int(np.nonzero(d)[0][0])
            sign[t] = int(d[perm[t]])
        perm[a] = an
        inward_nb = -1 if sn else 1
        sign[a] = inward_nb if side else -inward_nb
        off = [int(q0[perm[t]] - sign[t] * p0[t]) for t in range(3)]
        lo, hi = [0, 0, 0], [d + 1 for d in blk.dims]
        lo[a] = blk.dims[a] if side else 0
        hi[a] = lo[a] + 1
        rows.append(_patch_row(b, (a + 1) * (1 if side else -1), lo, hi, nb, perm, sign, off))
    return np.array(rows, dtype=np.int32).reshape(-1, PATCH_INTS)


def edge_potential(case: Case, A) -> List[np.ndarray]:
    """Edge line integrals of a vector potential A(x, y, z) -> (Ax, Ay, Az)
    (quadratic or lower: Simpson's rule is exact) along the straight edges
    between owned nodes, per block (3, box) in the layout of Case.A; edges
    with a node outside the owned range stay 0 (the faces they bound are
    ghosts)."""
    out = []
    g = case.g
    for blk in case.blocks:
        nd = np.asarray(blk.nodes, dtype=np.float64)
        shape = tuple(m + 1 + 2 * g for m in blk.dims)
        arr = np.zeros((3,) + shape)
        for a in range(3):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[a] = slice(0, -1)
            hi[a] = slice(1, None)
            p, q = nd[(slice(None),) + tuple(lo)], nd[(slice(None),) + tuple(hi)]
            mid = 0.5 * (p + q)
            fp, fm, fq = (np.stack(A(*x)) for x in (p, mid, q))
            val = ((q - p) * (fp + 4.0 * fm + fq) / 6.0).sum(axis=0)
            sl = tuple(slice(g, g + n) for n in val.shape)
            arr[(a,) + sl] = val
        out.append(arr)
    return out


def make_cubed_sphere(n: int = 8, nr: int = 8, g: int = 4, rule: int = 0, **kw) -> Case:
    """The two-fluid Hall problem of the C2 shell on a cubed sphere: six
    gnomonic panels of nr x n x n cells (radial, two angular axes) over the
    whole shell, the perfectly conducting sphere at r = 1 and the far field
    outside, the panels joined by rotated patches along their sides.  The
    eight corner lines where three panels meet are singular edges (valence 3)."""
    kw.setdefault("name", f"cubed_sphere_{n}x{nr}")
    c = make_shell(blocks=((nr, n, n),) * 6, panels=list(range(6)), g=g, half_angle=np.pi / 4, **kw)
    for blk in c.blocks:
        blk.bc = (BC_WALL, BC_FARFIELD) + (BC_INTERFACE,) * 4
    c.patches = glue_patches(c.blocks)
    c.patch_rule = rule
    return c
