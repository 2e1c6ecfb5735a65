"""The library's CASE geometry reader (mercury_case_*) on the CPU.

* the reference's own multi-rank sample geometries (tests/golden, from
  Z0_NULL/CASE/geometry.tar.gz): block sizes and coordinates as an independent
  parse of the grid files gives them; every connection patch's transform maps
  its node box onto the partner's box, carries each patch node onto the same
  point of the neighbour (or, for periodic patches, onto a point shifted by one
  constant period vector), and is the inverse of the partner patch's transform;
  their faces are split into partial patches, which the solver topology rejects
  by name;
* round trip: cases written in the reference's formats over several ranks read
  back with the same nodes and the same per-face topology the solver is given.
"""
import struct
import tarfile

import numpy as np
import pytest

from mpcns_mercury_amd import cases, hip

import case_format as cf

GOLDEN = __import__("pathlib").Path(__file__).parent / "golden" / "z0_null_case_geometry.tar.gz"


@pytest.fixture(scope="module")
def samples(tmp_path_factory):
    d = tmp_path_factory.mktemp("geometry")
    with tarfile.open(GOLDEN) as t:
        t.extractall(d, members=[m for m in t.getmembers() if "Zone.Identifier" not in m.name])
    return d


def _parse_grid(path):
    """Independent parse of a binary grid file: [(name, (imax, jmax, kmax), xyz (3, imax, jmax, kmax))]."""
    data = open(path, "rb").read()
    nb = struct.unpack_from("<i", data, 0)[0]
    p, names = 4, []
    for _ in range(nb):
        n = struct.unpack_from("<i", data, p)[0]
        names.append(data[p + 4:p + 4 + n].split(b"\0")[0].decode())
        p += 4 + n
    dims = [struct.unpack_from("<3i", data, p + 12 * b) for b in range(nb)]
    p += 12 * nb
    out = []
    for b in range(nb):
        n = dims[b][0] * dims[b][1] * dims[b][2]
        xyz = np.frombuffer(data, "<f8", count=3 * n, offset=p).reshape(3, dims[b][2], dims[b][1], dims[b][0])
        p += 24 * n
        out.append((names[b], dims[b], xyz.transpose(0, 3, 2, 1)))
    assert p == len(data)
    return out


def _map(pt, p):
    out = [0, 0, 0]
    for a in range(3):
        out[p["perm"][a]] = p["sign"][a] * pt[a] + p["offset"][a]
    return tuple(out)


@pytest.mark.parametrize("which,nrank,nblock", [("geometry_3", 3, 8), ("geometry_6", 6, 9)])
def test_reference_sample_geometry(samples, which, nrank, nblock):
    d = samples / which
    geo = hip.CaseGeometry(d)
    try:
        assert geo.nb == nblock
        grids = [g for r in range(nrank) for g in _parse_grid(d / f"grid{r:4d}.grd")]
        assert len(grids) == nblock
        for b, (name, nn, xyz) in enumerate(grids):
            assert geo.names[b] == name and geo.dims[b] == tuple(x - 1 for x in nn)
            np.testing.assert_array_equal(geo.nodes[b], xyz)
        pats = geo.interfaces()
        assert not geo.physical()  # every face of these samples is a connection
        kinds = {p["kind"] for p in pats}
        assert kinds == {"inner", "parallel"}
        scale = max(np.abs(x).max() for x in geo.nodes)
        nperiodic = 0
        for p in pats:
            lo, hi = p["lo"], p["hi"]
            c0, c1 = _map(lo, p), _map(tuple(h - 1 for h in hi), p)
            assert tuple(min(a, b) for a, b in zip(c0, c1)) == p["nb_lo"], p
            assert tuple(max(a, b) + 1 for a, b in zip(c0, c1)) == p["nb_hi"], p
            X, Y = geo.nodes[p["block"]], geo.nodes[p["nb_block"]]
            shifts = []
            for i in range(lo[0], hi[0]):
                for j in range(lo[1], hi[1]):
                    for k in range(lo[2], hi[2]):
                        q = _map((i, j, k), p)
                        shifts.append(Y[:, q[0], q[1], q[2]] - X[:, i, j, k])
            shifts = np.array(shifts)
            # one rigid shift per patch: zero for a plain connection, the period for a periodic one
            np.testing.assert_allclose(shifts, np.broadcast_to(shifts[0], shifts.shape), atol=1e-12 * scale, err_msg=str(p))
            if np.abs(shifts[0]).max() > 1e-12 * scale:
                nperiodic += 1
            # the partner patch (declared by the neighbour) carries the transform's inverse
            partners = [q for q in pats if q["block"] == p["nb_block"] and q["nb_block"] == p["block"]
                        and q["lo"] == p["nb_lo"] and q["hi"] == p["nb_hi"]]
            assert len(partners) == 1, p
            for pt in [lo, tuple(h - 1 for h in hi)]:
                assert _map(_map(pt, p), partners[0]) == tuple(pt)
        assert nperiodic > 0
        with pytest.raises(hip.MercuryError, match="split into"):
            geo.topology()
        # as patches: every face a connection, the reader's records in order
        bc, pt = geo.patch_topology()
        assert (bc == cases.BC_INTERFACE).all()
        np.testing.assert_array_equal(pt, cases.patches_of_geometry(geo))
    finally:
        geo.close()


def _topology_of(case):
    """The per-face arrays the solver receives for an in-memory case."""
    bc, nbr, tr = case.bc_array().copy(), case.nbr_array().copy(), case.trans_array().copy()
    for b in range(case.nb):
        for f in range(6):
            if bc[6 * b + f] != cases.BC_INTERFACE:
                nbr[6 * b + f] = -1
                tr[9 * (6 * b + f):9 * (6 * b + f + 1)] = [0, 1, 2, 1, 1, 1, 0, 0, 0]
    return bc, nbr, tr


@pytest.mark.parametrize("which,ranks", [("cube", [0, 0, 1, 1]), ("cube", [0, 1, 2, 3]), ("sector", [0, 0, 1, 1]),
                                         ("sector", [0, 0, 0, 0])])
def test_round_trip_through_the_case_files(tmp_path, which, ranks):
    if which == "cube":
        c = cases.make_split_cube((12, 10, 8), (2, 2, 1))
    else:
        c = cases.make_split_sector(dims=(16, 16, 12), cuts=((8, 8), (8, 8), (12,)), scale=1.0)
    cf.write_case(tmp_path, c, ranks)
    geo = hip.CaseGeometry(tmp_path)
    try:
        assert geo.nb == c.nb and geo.dims == [b.dims for b in c.blocks]
        for b in range(c.nb):
            np.testing.assert_array_equal(geo.nodes[b], c.blocks[b].nodes)
        for got, want in zip(geo.topology(), _topology_of(c)):
            np.testing.assert_array_equal(got, want)
        pats = geo.interfaces()
        assert {p["kind"] for p in pats} == ({"inner", "parallel"} if len(set(ranks)) > 1 else {"inner"})
        if which == "cube":
            assert any(p["periodic"] for p in pats if p["kind"] == "inner")
        r = cases.case_from_geometry(geo, c)
        for b in range(c.nb):
            assert r.blocks[b].bc == c.blocks[b].bc and r.blocks[b].nbr == c.blocks[b].nbr
            assert all(r.blocks[b].off[f] == c.blocks[b].off[f] for f in range(6) if c.blocks[b].bc[f] == cases.BC_INTERFACE)
    finally:
        geo.close()


def test_text_grids(tmp_path):
    """Read_Method 0: text grids hold the same nodes (written with repr, so they
    read back exactly)."""
    c = cases.make_split_sector(dims=(16, 16, 12), cuts=((8, 8), (16,), (12,)), scale=1.0)
    cf.write_case(tmp_path, c, [0, 1], binary=False)
    geo = hip.CaseGeometry(tmp_path, binary=False)
    try:
        assert geo.nb == 2 and geo.names == ["Fluid", "Fluid"]
        for b in range(c.nb):
            np.testing.assert_array_equal(geo.nodes[b], c.blocks[b].nodes)
        for got, want in zip(geo.topology(), _topology_of(c)):
            np.testing.assert_array_equal(got, want)
    finally:
        geo.close()
    with pytest.raises(hip.MercuryError):  # a text grid is not a binary one
        hip.CaseGeometry(tmp_path, binary=True)


def test_unsupported_boundaries_are_named(tmp_path):
    c = cases.make_split_sector(dims=(16, 16, 12), cuts=((16,), (8, 8), (12,)), scale=1.0)
    cf.write_case(tmp_path, c, [0, 0])
    p = tmp_path / "boundary_condition" / f"boundary{0:4d}.txt"
    txt = p.read_text().replace(f"\t{cf.BC_ID['Solid_Surface']}\n", f"\t{cf.BC_ID['Pole']}\n")
    p.write_text(txt)
    geo = hip.CaseGeometry(tmp_path)
    try:
        assert {q["bc_id"] for q in geo.physical()} >= {cf.BC_ID["Pole"], cf.BC_ID["Farfield"]}
        with pytest.raises(hip.MercuryError, match="Pole"):
            geo.topology()
    finally:
        geo.close()
    with pytest.raises(hip.MercuryError, match="no grid file|cannot open"):
        hip.CaseGeometry(tmp_path / "missing")


@pytest.mark.parametrize("binary", [True, False])
def test_cubed_sphere_through_the_case_files(tmp_path, binary):
    """A cubed sphere written as the reference's CASE files (rotated connections
    as inner records) reads back with the same nodes, face codes and, by the
    reference's IndexTransform rule, the same patch list in the same order; the
    singular edges found from the read-back topology are the in-memory case's."""
    c = cases.make_cubed_sphere(6, 4)
    cf.write_case(tmp_path, c, [0] * c.nb, binary=binary)
    geo = hip.CaseGeometry(tmp_path, binary=binary)
    try:
        assert geo.nb == c.nb and geo.dims == [b.dims for b in c.blocks]
        for b in range(c.nb):
            np.testing.assert_array_equal(geo.nodes[b], c.blocks[b].nodes)
        bc, pt = geo.patch_topology()
        np.testing.assert_array_equal(bc, c.bc_array())
        np.testing.assert_array_equal(pt, np.asarray(c.patches).reshape(-1, 18))
        np.testing.assert_array_equal(cases.patches_of_geometry(geo), pt)
    finally:
        geo.close()
    r = cases.make_cubed_sphere(6, 4)
    r.patches = pt
    want = {e["owner"]: e for e in hip.singular_edges(c)}
    got = {e["owner"]: e for e in hip.singular_edges(r)}
    assert len(got) == 32 and set(got) == set(want)
    for k, e in got.items():
        assert e["inverse_hodge"] == want[k]["inverse_hodge"]
