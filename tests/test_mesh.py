"""libimx mesh builder + gridprop against the reference's own meshes.

The builder triangulates with a restatement of STRIPACK's incremental
Delaunay insertion (trmesh) at every refinement level, so the meshes carry
the reference's numbering: every array of grid_structure the shallow-water
path reads, and the gridprop products, equal the reference's bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from imodel_amd import imxbin
from imodel_amd.swm import Mesh
from tests.conftest import ROOT

REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def assert_same_arrays(mine, ref):
    common = [k for k in ref if k in mine]
    assert len(common) >= 30, common
    for k in common:
        a, b = np.asarray(mine[k]).ravel(), np.asarray(ref[k]).ravel()
        assert a.dtype == b.dtype and a.shape == b.shape, k
        assert np.array_equal(a, b), (k, int(np.sum(a != b)))


@pytest.mark.parametrize("name,g,pos", [("ref_mesh_g2.bin", 2, "eqs"), ("ref_mesh_g3.bin", 3, "eqs"),
                                        ("ref_mesh_g2_pol.bin", 2, "pol")])
def test_same_mesh_as_reference(golden, name, g, pos):
    """Every record of the reference's mesh dump (index tables, signs,
    coordinates, lengths, areas, kite areas, TRiSK weights) is identical."""
    assert_same_arrays(Mesh.icos(g, pos).full_arrays(), golden(name))


@pytest.mark.skipif(not os.path.exists(REF_HARNESS), reason="reference harness not built")
@pytest.mark.parametrize("g,pos", [(g, "eqs") for g in range(0, 6)] + [(g, "pol") for g in range(0, 5)])
def test_same_mesh_as_reference_meshbuild(tmp_path, g, pos):
    """Against the reference's own meshbuild run now (oracle/_ref), glevel
    0-5 eqs and 0-4 pol."""
    out = str(tmp_path / "m.bin")
    subprocess.run([REF_HARNESS, "build", str(g), pos, out], cwd=tmp_path, check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    assert_same_arrays(Mesh.icos(g, pos).full_arrays(), imxbin.read(out))


def test_planted_ring_rotation_is_caught(golden):
    """The comparison sees a ring rotated by one slot (what a different
    insertion history gives)."""
    ref = dict(golden("ref_mesh_g2.bin"))
    mine = dict(Mesh.icos(2).full_arrays())
    nb = np.array(mine["v_nb"]).reshape(-1, 6)
    n = int(mine["v_nnb"][7])
    nb[7, :n] = np.roll(nb[7, :n], 1)
    mine["v_nb"] = nb.ravel()
    with pytest.raises(AssertionError):
        assert_same_arrays(mine, ref)


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
    """TC2 initial state from libimx equals the reference's on the same mesh."""
    r = golden("ref_run_g3_tc2.bin")
    h, u = Mesh.icos(3).tc2()
    assert np.array_equal(h, r["h0"]) and np.array_equal(u, r["u0"])


def test_tc2_hc_matches_reference(golden):
    """TC2 with HC staggering (u along edhx%nr at the Voronoi edge
    midpoints) equals the reference's."""
    r = golden("ref_run_g3_tc2hc.bin")
    h, u = Mesh.icos(3).tc2(stag="HC")
    assert np.array_equal(h, r["h0"]) and np.array_equal(u, r["u0"])


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
