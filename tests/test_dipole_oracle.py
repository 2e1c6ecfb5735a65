"""The oracle's intrinsic-dipole part of Badd (CPU only).

DipoleField::Build_Badd_FaceFlux (Z4_Mercury/include/operators/Dipole.h:89-374)
takes the edge 1-form A.dr of the softened dipole potential
A = (m x r) / (|r|^2 + a^2)^(3/2) and its curl d1 as the face flux.  Pinned by
* the closed-form curl of that potential,
  B = (3 r (r.m) - m (|r|^2 - 2 a^2)) / (|r|^2 + a^2)^(5/2),
  dotted with the face area vectors (second order in the cell size; the
  mutants of tests/test_oracle_mutants.py -- dropped softening, flipped cross
  product, endpoint instead of midpoint value, wrong power -- each fail it);
* d2 Badd = 0 to roundoff (a curl has no divergence);
* the face boxes it covers (ghost depth 1, Dipole.h:213-324);
* Mercury's equatorial surface field, mu0 |m| / (4 pi R_M^3) ~ 195 nT for the
  reference's example dipole (Z4_Mercury/9999setup/3_MHD_Control.txt:1-9),
  which fixes the nondimensionalisation of DipoleField::load_from_param (:44-87).
"""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.params import MercuryUnits, dipole_array

AX = ("_xi", "_eta", "_zeta")
FACE_DELTA = [(0, 1, 1), (1, 0, 1), (1, 1, 0)]


def b_closed(x, D):
    """Field of one softened dipole (row of 7) at points x (3, ...)."""
    r = x - D[:3].reshape((3,) + (1,) * (x.ndim - 1))
    m = D[3:6].reshape((3,) + (1,) * (x.ndim - 1))
    r2 = (r * r).sum(axis=0)
    s = r2 + D[6] ** 2
    return (3.0 * r * (r * m).sum(axis=0) - m * (r2 - 2.0 * D[6] ** 2)) / s ** 2.5


def _block(n=24, h=0.0125, origin=(0.6, 0.1, 0.2), dip=((0.02, -0.03, 0.05, 0.3, -0.5, 1.0, 0.3),)):
    nodes = cases.cartesian_nodes(n=(n, n, n), h=(h, h, h), origin=origin)
    case = cases.make_uniform_block(nodes, (cases.BC_FARFIELD,) * 6, g=2)
    case.phys.B_imf[0] = case.phys.B_imf[1] = case.phys.B_imf[2] = 0.0
    case.dipoles = np.array(dip, dtype=np.float64)
    return case


def _face_centres(nodes, a):
    """Centres of the owned faces of axis a: the mean of each face's 4 corner
    nodes, from the owned node array (3, m0+1, m1+1, m2+1)."""
    b, c = [t for t in range(3) if t != a]
    n = nodes.shape[1:]
    acc = 0.0
    for db in (0, 1):
        for dc in (0, 1):
            sl = [slice(None)] * 3
            sl[b] = slice(db, db + n[b] - 1)
            sl[c] = slice(dc, dc + n[c] - 1)
            acc = acc + nodes[(slice(None),) + tuple(sl)]
    return 0.25 * acc


def test_face_flux_matches_closed_form_field(oracle_mod):
    case = _block()
    o = oracle_mod.Oracle(case)
    D = case.dipoles[0]
    g = case.g
    m = case.blocks[0].dims
    worst = 0.0
    scale = 0.0
    for a, s in enumerate(AX):
        flux = o.get("Badd" + s, 0)[0]
        JD = o.get("JD" + ("xi", "et", "ze")[a], 0)
        own = tuple(slice(g, g + mm + 1 - d) for mm, d in zip(m, FACE_DELTA[a]))
        xc = _face_centres(case.blocks[0].nodes, a)
        expect = (b_closed(xc, D) * JD[(slice(None),) + own]).sum(axis=0)
        worst = max(worst, np.abs(flux[own] - expect).max())
        scale = max(scale, np.abs(expect).max())
    assert scale > 1e-4
    assert worst <= 2e-3 * scale, (worst, scale)


def test_badd_is_divergence_free_and_covers_depth_one(oracle_mod):
    case = _block(n=10, h=0.05)
    o = oracle_mod.Oracle(case)
    g = case.g
    m = case.blocks[0].dims
    f = [o.get("Badd" + s, 0)[0] for s in AX]
    # d2 over cells whose six faces lie in the depth-1 face boxes
    c = (f[0][1:, :-1, :-1] - f[0][:-1, :-1, :-1]) + (f[1][:-1, 1:, :-1] - f[1][:-1, :-1, :-1]) + \
        (f[2][:-1, :-1, 1:] - f[2][:-1, :-1, :-1])
    inner = tuple(slice(g - 1, g + mm + 1) for mm in m)
    assert np.abs(c[inner]).max() <= 1e-14 * max(np.abs(x).max() for x in f)
    for a in range(3):
        box = tuple(slice(g - 1, g + mm + 2 - d) for mm, d in zip(m, FACE_DELTA[a]))
        mask = np.zeros(f[a].shape, bool)
        mask[box] = True
        assert np.all(f[a][~mask] == 0.0)
        assert np.count_nonzero(f[a][mask]) > 0.9 * mask.sum()


def test_mercury_surface_field_from_the_example_dipole():
    u = MercuryUnits().alfven_units()
    u.dipoles = [dict(cases.MERCURY_DIPOLE)]
    D = dipole_array(u)[0]
    assert D[0] == 0.0 and abs(D[2] - 4.88e5 / u.REF["L_ref"]) <= 1e-15
    # field at the equator one radius out, in the dipole's own frame, dimensional
    x = np.array(D[:3]) + np.array([1.0, 0.0, 0.0])
    B = b_closed(x.reshape(3, 1), D)[:, 0] * u.REF["B_ref"]
    expect = u.constant["mu_mag"] * 2.83e19 / (4 * np.pi * u.REF["L_ref"] ** 3)
    assert abs(np.linalg.norm(B) - expect) <= 1e-6 * expect
    assert B[2] > 0.0  # a southward moment gives a northward equatorial field
    assert 1.9e-7 < expect < 2.0e-7  # ~195 nT
