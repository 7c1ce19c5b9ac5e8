"""meshbuild for mesh kind "read" (imx_mesh_build_read): nodes from .xyz or
.gmt files and the neighbour rings from the .ngb file, as meshread and
meshstruct take them (src/smeshpack.f90:7549-7700, :370-1195).

The reference's own meshbuild (oracle/_ref/ref_harness buildread) reads the
same files; every array of the two meshes must be identical, with the .ngb
rings and, without them, with both sides triangulating the nodes. The files are
made from a glevel-2 icosahedral grid with its nodes shuffled and each ring
started at a random neighbour, so neither the node order nor the ring starts
follow the icosahedral builder's rules.
"""
import os
import subprocess

import numpy as np
import pytest

from imodel_amd import capi, imxbin
from imodel_amd.swm import Mesh
from oracle.oracle import REF_HARNESS, ref_harness_available


def _write_grid(td, name, seed, fmt, glevel=2):
    m = Mesh.icos(glevel, gridprop=False).arrays
    nv = int(m["dims"][0])
    rng = np.random.default_rng(seed)
    perm = rng.permutation(nv)            # new index of old node i
    inv = np.argsort(perm)
    p = m["v_p"].reshape(nv, 3)[inv]
    nb = m["v_nb"].reshape(nv, 6)[inv]
    nnb = m["v_nnb"][inv]
    os.makedirs(os.path.join(td, "grid"), exist_ok=True)
    if fmt == "xyz":
        with open(os.path.join(td, "grid", name + ".xyz"), "w") as f:
            f.write(f"{nv}\n")
            for x in p:
                f.write("%.17g %.17g %.17g\n" % tuple(x))
    else:
        lat = np.degrees(np.arcsin(p[:, 2]))
        lon = np.degrees(np.arctan2(p[:, 1], p[:, 0]))
        with open(os.path.join(td, "grid", name + ".gmt"), "w") as f:
            for a, b in zip(lon, lat):
                f.write("%.17g %.17g\n" % (a, b))
    with open(os.path.join(td, "grid", name + ".ngb"), "w") as f:
        for i in range(nv):
            ring = [int(perm[k - 1]) + 1 for k in nb[i, :nnb[i]]]
            r = int(rng.integers(nnb[i]))
            ring = ring[r:] + ring[:r]
            f.write(" ".join(str(k) for k in ring + [0] * (6 - len(ring))) + "\n")
    return nv


@pytest.mark.skipif(not ref_harness_available(), reason="reference harness not built (oracle/Makefile.ref)")
@pytest.mark.parametrize("fmt", ["xyz", "gmt"])
def test_build_read_matches_reference(tmp_path, fmt):
    td = str(tmp_path)
    nv = _write_grid(td, "shuffled", 5, fmt)
    r = subprocess.run([REF_HARNESS, "buildread", f"shuffled.{fmt}", str(nv), "ref.bin"], cwd=td,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0 and os.path.exists(os.path.join(td, "ref.bin")), r.stdout[-2000:]
    ref = imxbin.read(os.path.join(td, "ref.bin"))
    m = Mesh.build_read(os.path.join(td, "grid", f"shuffled.{fmt}"))
    m.write(os.path.join(td, "ours.bin"))
    ours = imxbin.read(os.path.join(td, "ours.bin"))
    checked = 0
    for k, v in ref.items():
        if k in ours and k != "dims":
            assert np.array_equal(np.asarray(ours[k]).ravel(), np.asarray(v).ravel()), k
            checked += 1
    assert checked >= 33, checked


@pytest.mark.skipif(not ref_harness_available(), reason="reference harness not built (oracle/Makefile.ref)")
@pytest.mark.parametrize("seed,glevel", [(4, 2), (11, 3)])
def test_build_read_triangulated_matches_reference(tmp_path, seed, glevel):
    """Nodes without an .ngb file: both sides triangulate the shuffled nodes
    (STRIPACK trmesh in the reference, its restatement here) and every array
    of the two meshes is identical."""
    td = str(tmp_path)
    nv = _write_grid(td, "bare", seed, "xyz", glevel=glevel)
    os.remove(os.path.join(td, "grid", "bare.ngb"))
    r = subprocess.run([REF_HARNESS, "buildread", "bare.xyz", str(nv), "ref.bin"], cwd=td,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    ref = imxbin.read(os.path.join(td, "ref.bin"))
    m = Mesh.build_read(os.path.join(td, "grid", "bare.xyz"))
    m.write(os.path.join(td, "ours.bin"))
    ours = imxbin.read(os.path.join(td, "ours.bin"))
    checked = 0
    for k, v in ref.items():
        if k in ours and k != "dims":
            assert np.array_equal(np.asarray(ours[k]).ravel(), np.asarray(v).ravel()), k
            checked += 1
    assert checked >= 33, checked


def test_build_read_errors(tmp_path):
    td = str(tmp_path)
    # shuffled glevel-2 nodes whose first three lie on one great circle
    # (nodes 1 and 3 antipodal): the reference's trmesh stops there too
    _write_grid(td, "coll", 3, "xyz")
    os.remove(os.path.join(td, "grid", "coll.ngb"))
    with pytest.raises(capi.ImxError, match="collinear"):
        Mesh.build_read(os.path.join(td, "grid", "coll.xyz"))
    with pytest.raises(capi.ImxError, match="Extension"):
        Mesh.build_read(os.path.join(td, "grid", "g.txt"))
    os.makedirs(os.path.join(td, "grid"), exist_ok=True)
    with open(os.path.join(td, "grid", "dup.xyz"), "w") as f:   # a repeated node
        f.write("5\n1 0 0\n0 1 0\n0 0 1\n0 1 0\n-1 0 0\n")
    with pytest.raises(capi.ImxError, match="equals node"):
        Mesh.build_read(os.path.join(td, "grid", "dup.xyz"))
