"""libimx mesh builder + gridprop against the reference's own meshes.

The builder carries the icosahedral subdivision as explicit triangles instead
of re-running a Delaunay insertion, so numbering differs from the reference;
the mesh itself must be the same: bit-identical node coordinates, identical
edge/triangle sets and CCW neighbour rings, geometry equal to rounding.
"""
import os

import numpy as np
import pytest

from imodel_amd.swm import Mesh
from tests.meshmatch import match


@pytest.mark.parametrize("name,g,pos", [("ref_mesh_g2.bin", 2, "eqs"), ("ref_mesh_g3.bin", 3, "eqs"),
                                        ("ref_mesh_g2_pol.bin", 2, "pol")])
def test_same_mesh_as_reference(golden, name, g, pos):
    ref = golden(name)
    m = Mesh.icos(g, pos)
    mine = m.arrays
    assert list(mine["dims"][:3]) == list(ref["dims"][:3])
    nmap, emap, tmap, flip = match(mine, ref)
    if pos == "eqs":
        assert (nmap >= 0).all(), "node coordinates not bit-identical"
    else:
        # pol uses libm cos/sin of the single-rounded pi; allow a last-bit slip
        from scipy.spatial import cKDTree
        d, nmap = cKDTree(ref["v_p"].reshape(-1, 3)).query(mine["v_p"].reshape(-1, 3))
        assert d.max() < 1e-15
        nmap, emap, tmap, flip = match({**mine, "v_p": ref["v_p"].reshape(-1, 3)[nmap].ravel()}, ref)
    assert (emap >= 0).all() and len(set(emap)) == len(emap)
    assert (tmap >= 0).all() and len(set(tmap)) == len(tmap)
    # CCW rings equal up to rotation
    for i in range(m.nv):
        n = mine["v_nnb"][i]
        assert ref["v_nnb"][nmap[i]] == n
        mr = [nmap[x - 1] for x in mine["v_nb"][6 * i:6 * i + n]]
        rr = list(ref["v_nb"][6 * nmap[i]:6 * nmap[i] + n] - 1)
        k = rr.index(mr[0])
        assert rr[k:] + rr[:k] == mr
    rel = lambda a, b: np.max(np.abs(a - b) / np.abs(b))  # noqa: E731
    assert rel(mine["ed_leng"], ref["ed_leng"][emap]) < 1e-14
    assert rel(mine["edhx_leng"], ref["edhx_leng"][emap]) < 1e-12
    assert rel(mine["tr_areag"], ref["tr_areag"][tmap]) < 1e-13
    assert rel(mine["hx_areag"], ref["hx_areag"][nmap]) < 1e-12
    # sign conventions: tr%nr = -tr%tg, edge flips only reverse orientation
    assert np.array_equal(mine["tr_nr"] if "tr_nr" in mine else -mine["tr_tg"], -mine["tr_tg"])


@pytest.mark.parametrize("name", ["ref_mesh_g2.bin", "ref_mesh_g3.bin"])
def test_gridprop_bitexact_on_reference_mesh(golden, name):
    """initialize_gridprop restated in libimx reproduces the reference's
    kite areas and TRiSK weights bit for bit on the reference's mesh."""
    ref = dict(golden(name))
    expect = {k: ref.pop(k) for k in ("tr_trhx_area", "hx_hxtr_area", "hx_trskw")}
    trhx, hxtr, trskw = Mesh(ref).gridprop()
    assert np.array_equal(trhx, expect["tr_trhx_area"])
    assert np.array_equal(hxtr, expect["hx_hxtr_area"])
    assert np.array_equal(trskw, expect["hx_trskw"])


@pytest.mark.parametrize("g", [0, 1, 4, 5])
def test_counts_and_closure(g):
    m = Mesh.icos(g)
    a = m.arrays
    assert m.nv == 10 * 4 ** g + 2 and m.ne == 3 * m.nv - 6 and m.nt == 2 * m.nv - 4
    assert np.sum(a["v_nnb"] == 5) == 12
    assert abs(np.sum(a["tr_areag"]) - 4 * np.pi) < 1e-10
    assert abs(np.sum(a["hx_areag"]) - 4 * np.pi) < 1e-10
    # kites of a cell tile it; kites of a triangle tile it
    R = a["hx_hxtr_area"].reshape(-1, 6).sum(axis=1)
    # L'Huilier areas lose digits on small triangles: tolerance grows with level
    tol = 1e-13 * 4.0 ** g
    assert np.max(np.abs(R - 1)) < tol
    K = a["tr_trhx_area"].reshape(-1, 3).sum(axis=1)
    assert np.max(np.abs(K / a["tr_areag"] - 1)) < tol
    # TRiSK weights are antisymmetric: w(i,j) = -w(j,i) (Thuburn et al. 2009)
    W = a["hx_trskw"].reshape(-1, 6, 6)
    assert np.max(np.abs(W + W.transpose(0, 2, 1))) < tol
    # every edge has two distinct cells and two distinct triangles
    ev = a["ed_v"].reshape(-1, 2)
    es = a["ed_sh"].reshape(-1, 2)
    assert (ev[:, 0] < ev[:, 1]).all() and (es[:, 0] < es[:, 1]).all()


def test_tc2_matches_reference(golden):
    """TC2 initial state from libimx equals the reference's on the same mesh,
    up to the orientation of renumbered edges."""
    ref_mesh = golden("ref_mesh_g3.bin")
    r = golden("ref_run_g3_tc2.bin")
    m = Mesh.icos(3)
    h, u = m.tc2()
    nmap, emap, tmap, flip = match(m.arrays, ref_mesh)
    assert np.array_equal(h, r["h0"][nmap])
    assert np.array_equal(np.where(flip, -u, u), r["u0"][emap])


def test_tc2_hc_matches_reference(golden):
    """TC2 with HC staggering (u along edhx%nr at the Voronoi edge
    midpoints) equals the reference's up to the orientation of renumbered
    dual edges and the rounding of the circumcentres it is built from."""
    ref_mesh = golden("ref_mesh_g3.bin")
    r = golden("ref_run_g3_tc2hc.bin")
    m = Mesh.icos(3)
    h, u = m.tc2(stag="HC")
    nmap, emap, tmap, flip = match(m.arrays, ref_mesh)
    assert np.array_equal(h, r["h0"][nmap])
    sh = m.arrays["ed_sh"].reshape(-1, 2) - 1
    rsh = ref_mesh["ed_sh"].reshape(-1, 2) - 1
    fsh = tmap[sh[:, 0]] != rsh[emap, 0]   # dual edge walked the other way
    us = np.where(fsh, -u, u)
    assert np.max(np.abs(us - r["u0"][emap])) < 1e-14 * np.max(np.abs(u))


def test_mesh_read_round_trip(golden, tmp_path):
    """imx_mesh_read takes the reference harness's mesh dump and what
    imx_mesh_write writes, record for record."""
    from imodel_amd.swm import Mesh
    from tests.conftest import GOLDEN
    r = golden("ref_mesh_g3.bin")
    m = Mesh.read(os.path.join(GOLDEN, "ref_mesh_g3.bin"), gridprop=False)
    for k, v in m.arrays.items():
        if k in r:
            assert np.array_equal(np.asarray(v).ravel(), np.asarray(r[k]).ravel()), k
    a = Mesh.icos(2)
    p = str(tmp_path / "m.bin")
    a.write(p)
    b = Mesh.read(p, gridprop=False)
    for k, v in a.arrays.items():
        assert np.array_equal(np.asarray(v), np.asarray(b.arrays[k])), k
    with pytest.raises(Exception, match="IMXBIN"):
        bad = tmp_path / "bad.bin"
        bad.write_bytes(b"not a mesh")
        Mesh.read(str(bad))
