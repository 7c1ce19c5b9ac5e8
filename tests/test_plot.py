"""plot_scalarfield's lat-lon sampling (imx_remap_ll): bit for bit the
stream the reference's own plot_scalarfield writes for fields on cells,
triangles and both edge kinds, and the containing-triangle walk against a
brute-force search."""
import os
import subprocess

import numpy as np
import pytest

from imodel_amd import imxbin
from imodel_amd.swm import Mesh
from oracle.oracle import REF_HARNESS, ref_harness_available


@pytest.mark.skipif(not ref_harness_available(), reason="reference harness not built (oracle/Makefile.ref)")
def test_matches_reference_plot_scalarfield(tmp_path):
    subprocess.run([REF_HARNESS, "build", "3", "eqs", str(tmp_path / "ref.bin")], cwd=tmp_path, check=True,
                   stdout=subprocess.DEVNULL)
    m = Mesh(imxbin.read(str(tmp_path / "ref.bin")))   # the reference's mesh and numbering
    for pos, n in ((0, m.nv), (1, m.nt), (3, m.ne), (6, m.ne)):
        f = np.random.default_rng(pos).standard_normal(n)
        imxbin.write(str(tmp_path / "f.bin"), {"f": f})
        subprocess.run([REF_HARNESS, "plot", "3", str(tmp_path / "f.bin"), str(pos)], cwd=tmp_path, check=True,
                       stdout=subprocess.DEVNULL)
        ref = np.fromfile(tmp_path / "data" / "plottest_icos_eqs_nopt_3.dat", dtype=np.float32).reshape(-1, 3)
        ours = m.remap_ll(f, pos)
        assert ours.shape == ref.shape == (720 * 1440, 3)
        assert np.array_equal(ours, ref), (pos, int(np.sum(ours != ref)))


def test_walk_finds_containing_triangle():
    """pos 1 returns the value of the triangle containing each point: with
    f = triangle index, every sample must lie inside that triangle, and
    cell values must come from one of its vertices (checked brute force)."""
    m = Mesh.icos(2)
    a = m.arrays
    tv = a["tr_v"].reshape(-1, 3) - 1
    P = a["v_p"].reshape(-1, 3)
    nlon, nlat = 96, 48
    t = m.remap_ll(np.arange(m.nt, dtype=np.float64), 1, nlon, nlat)
    cells = m.remap_ll(np.arange(m.nv, dtype=np.float64), 0, nlon, nlat)
    lon, lat = np.deg2rad(t[:, 0].astype(np.float64)), np.deg2rad(t[:, 1].astype(np.float64))
    p = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], 1)
    k = t[:, 2].astype(int)
    A, B, C = P[tv[k, 0]], P[tv[k, 1]], P[tv[k, 2]]
    for X, Y in ((A, B), (B, C), (C, A)):
        assert (np.einsum("ij,ij->i", np.cross(X, Y), p) > -1e-6).all()
    assert all(int(c) in tv[kk] for c, kk in zip(cells[:, 2], k))
    # coordinates: cell-centred, longitude fastest
    assert np.allclose(t[:3, 0], [-180 + 1.875, -180 + 3 * 1.875, -180 + 5 * 1.875])
    assert np.allclose(t[::nlon, 1][:2], [-90 + 1.875, -90 + 3 * 1.875])


def test_bad_position_raises():
    from imodel_amd.capi import ImxError
    m = Mesh.icos(1)
    with pytest.raises(ImxError):
        m.remap_ll(np.zeros(m.nv), 4, 8, 4)
