"""DEC operators of the CPU oracle on one block (CPU only).

The oracle restates src/8_dec/DecExteriorDerivative.cpp and DecDiagnostics.cpp.
Pinned here by
* the reference's own DEC chain test: d1 d0 phi = 0 and d2 d1 E = 0 on its
  polynomial fills (Z0_NULL/test/dec_chain/Z0_DecChainTest.cpp:13-51,85-100),
  exactly, since every value is a dyadic rational;
* the incidence pattern of each operator: one unit value on one DOF must give
  the signed boundary / coboundary of that DOF (Stokes' orientation of a unit
  cube, edges along +xi/+eta/+zeta, faces oriented by their normal axis);
* the owned ranges: points outside [0, m + 1 - delta) keep their values.
"""
import numpy as np
import pytest

import oracle as orc


def _phi(dims, g):
    """fill_phi of Z0_DecChainTest.cpp:13-27 over the whole box."""
    I, J, K = np.meshgrid(*[np.arange(-g, m + 1 + g) for m in dims], indexing="ij")
    return 0.125 * I * I - 0.25 * J * K + 0.5 * K + 3.0


def _edge_fill(dims, g):
    """fill_edge_form of Z0_DecChainTest.cpp:29-51 (biases 1, -2, 0.75)."""
    I, J, K = np.meshgrid(*[np.arange(-g, m + 1 + g) for m in dims], indexing="ij")
    base = 0.375 * I - 0.5 * J + 0.625 * K + 0.0625 * I * J - 0.03125 * J * K
    return [1.0 + base, -2.0 + base, 0.75 + base]


def _owned(dims, g, delta):
    return tuple(slice(g, g + m + 1 - d) for m, d in zip(dims, delta))


EDGE_DELTA = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
FACE_DELTA = [(0, 1, 1), (1, 0, 1), (1, 1, 0)]


@pytest.mark.parametrize("dims,g", [((6, 5, 4), 4), ((1, 3, 7), 2), ((4, 4, 4), 0)])
def test_reference_chain_fills_give_zero(oracle_mod, dims, g):
    phi = _phi(dims, g)
    e = orc.dec_d0(dims, g, phi)
    f = orc.dec_d1(dims, g, e)
    for a in range(3):
        own = f[a][_owned(dims, g, FACE_DELTA[a])]
        assert np.all(own == 0.0), a
    assert orc.dec_max_abs_curl_grad(dims, g, phi) == 0.0
    E = _edge_fill(dims, g)
    c = orc.dec_d2(dims, g, orc.dec_d1(dims, g, E))
    assert np.all(c[_owned(dims, g, (1, 1, 1))] == 0.0)
    assert orc.dec_max_abs_div_curl(dims, g, E) == 0.0


def test_d0_incidence(oracle_mod):
    dims, g = (4, 3, 5), 2
    n = np.zeros(orc._box_shape(dims, g))
    P = (2, 1, 3)
    n[tuple(g + p for p in P)] = 1.0
    e = orc.dec_d0(dims, g, n)
    for a in range(3):
        own = e[a][_owned(dims, g, EDGE_DELTA[a])]
        expect = np.zeros_like(own)
        expect[P] = -1.0  # the edge leaving the node
        Q = list(P)
        Q[a] -= 1
        expect[tuple(Q)] = 1.0  # the edge arriving at it
        np.testing.assert_array_equal(own, expect)


def test_d1_incidence(oracle_mod):
    """A unit xi-edge at (i,j,k) circulates around the faces that contain it:
    eta-faces (i,j,k-1) and (i,j,k) with signs +1, -1, zeta-faces (i,j,k) and
    (i,j-1,k) with +1, -1 (right-hand rule about each face normal)."""
    dims, g = (4, 4, 4), 1
    sh = orc._box_shape(dims, g)
    P = (1, 2, 2)
    for a in range(3):
        e = [np.zeros(sh) for _ in range(3)]
        e[a][tuple(g + p for p in P)] = 1.0
        f = orc.dec_d1(dims, g, e)
        b, c = (a + 1) % 3, (a + 2) % 3  # cyclic partner axes
        expect = [np.zeros(sh) for _ in range(3)]
        # face normal b contains the edge at P and P - e_c: +1 at P - e_c, -1 at P
        Pc = list(P)
        Pc[c] -= 1
        expect[b][tuple(g + p for p in Pc)] = 1.0
        expect[b][tuple(g + p for p in P)] = -1.0
        # face normal c: +1 at P, -1 at P - e_b
        Pb = list(P)
        Pb[b] -= 1
        expect[c][tuple(g + p for p in P)] = 1.0
        expect[c][tuple(g + p for p in Pb)] = -1.0
        for d in range(3):
            sl = _owned(dims, g, FACE_DELTA[d])
            np.testing.assert_array_equal(f[d][sl], expect[d][sl], err_msg=f"edge axis {a}, face axis {d}")


def test_d2_incidence_and_faraday(oracle_mod):
    dims, g = (3, 4, 5), 1
    sh = orc._box_shape(dims, g)
    P = (1, 2, 3)
    for a in range(3):
        f = [np.zeros(sh) for _ in range(3)]
        f[a][tuple(g + p for p in P)] = 1.0
        c = orc.dec_d2(dims, g, f)[_owned(dims, g, (1, 1, 1))]
        expect = np.zeros_like(c)
        expect[P] = -1.0  # the face is the -side face of cell P ...
        Q = list(P)
        Q[a] -= 1
        expect[tuple(Q)] = 1.0  # ... and the +side face of cell P - e_a
        np.testing.assert_array_equal(c, expect)
    # Faraday: face - dt * d1(edge), owned faces only
    rng = np.random.default_rng(5)
    e = [rng.standard_normal(sh) for _ in range(3)]
    f0 = [rng.standard_normal(sh) for _ in range(3)]
    d1 = orc.dec_d1(dims, g, e)
    f1 = orc.dec_d1(dims, g, e, out=f0, update=True, dt=0.25)
    for a in range(3):
        sl = _owned(dims, g, FACE_DELTA[a])
        np.testing.assert_array_equal(f1[a][sl], f0[a][sl] - 0.25 * d1[a][sl])
        mask = np.ones(sh, bool)
        mask[sl] = False
        np.testing.assert_array_equal(f1[a][mask], f0[a][mask])


def test_owned_ranges_only(oracle_mod):
    """Outputs outside the owned range keep their values (NaN sentinels stay)."""
    dims, g = (5, 4, 3), 3
    sh = orc._box_shape(dims, g)
    rng = np.random.default_rng(1)
    n = rng.standard_normal(sh)
    e = orc.dec_d0(dims, g, n)
    for a in range(3):
        mask = np.ones(sh, bool)
        mask[_owned(dims, g, EDGE_DELTA[a])] = False
        assert np.all(np.isnan(e[a][mask])) and not np.any(np.isnan(e[a][~mask]))
    f = orc.dec_d1(dims, g, e)
    for a in range(3):
        mask = np.ones(sh, bool)
        mask[_owned(dims, g, FACE_DELTA[a])] = False
        assert np.all(np.isnan(f[a][mask])) and not np.any(np.isnan(f[a][~mask]))


def test_diagnostic_maxima_see_a_broken_chain(oracle_mod):
    """max|d2 d1 E| and max|d1 d0 phi| are 0 for any integer-valued input
    (exact arithmetic), roundoff-sized and positive for random real input, and
    never read the ghost layers."""
    dims, g = (6, 5, 7), 2
    sh = orc._box_shape(dims, g)
    rng = np.random.default_rng(2)
    E = [rng.integers(-1000, 1000, sh).astype(float) for _ in range(3)]
    assert orc.dec_max_abs_div_curl(dims, g, E) == 0.0
    n = rng.integers(-1000, 1000, sh).astype(float)
    assert orc.dec_max_abs_curl_grad(dims, g, n) == 0.0
    # random real values: roundoff only
    R = [rng.standard_normal(sh) for _ in range(3)]
    v = orc.dec_max_abs_div_curl(dims, g, R)
    assert 0.0 < v < 1e-14
    # NaN in the ghosts is never read
    for x in R:
        x[:g] = np.nan
    assert orc.dec_max_abs_div_curl(dims, g, R) == v
