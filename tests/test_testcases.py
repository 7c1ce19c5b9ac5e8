"""initialize_fields' test cases (src/swm.f90:1043-2414) and the bottom
topography they carry, against fixtures the reference wrote
(tests/golden/ref_run_g3_tc*.bin, make_golden.py tc).

* imx_ic_testcase reproduces the reference's h, u and bt bit for bit
  (test cases 1, 5 HTC and HC, 51, 6, 21, 53, 11, 32, 33, 34 HC, 40, 41, 42
  HTC and HC at glevel 3; the others of the build's list are checked
  against the reference harness when it is built).
* the oracle, given bt and the case's Coriolis setting (f-sphere), reproduces
  the reference's tendency (with g (h + b) in the Bernoulli potential), its
  intermediates, the diagnostics (potential and available energy with b)
  and 3 RK4 steps bit for bit; dropping bt, or the f-sphere, fails the
  check.
"""
import os
import subprocess

import numpy as np
import pytest

from imodel_amd import capi, imxbin
from imodel_amd.swm import Mesh
from oracle.oracle import REF_HARNESS, Oracle, ref_harness_available

CASES = {"tc1": (1, "HTC"), "tc5": (5, "HTC"), "tc5hc": (5, "HC"), "tc51": (51, "HTC"), "tc6": (6, "HTC"),
         "tc21": (21, "HTC"), "tc53": (53, "HTC"), "tc11": (11, "HTC"), "tc32": (32, "HTC"),
         "tc33": (33, "HTC"), "tc34hc": (34, "HC"), "tc40": (40, "HTC"), "tc41": (41, "HTC"),
         "tc42": (42, "HTC"), "tc42hc": (42, "HC")}
FIELDS = ["h_ed", "uh", "eta", "h_tr", "q_tr", "q_ed", "divuh", "ke_hx", "ghbK", "grad_ghbK", "uhq_perp"]


@pytest.mark.parametrize("name", sorted(CASES))
def test_testcase_state_bitexact(golden, name):
    tc, stag = CASES[name]
    r = golden(f"ref_run_g3_{name}.bin")
    h, u, bt = Mesh.icos(3).testcase(tc, stag)
    assert np.array_equal(h, r["h0"]) and np.array_equal(u, r["u0"]) and np.array_equal(bt, r["bt"])


@pytest.mark.skipif(not ref_harness_available(), reason="reference harness not built (oracle/Makefile.ref)")
@pytest.mark.parametrize("tc", [0, 2, 12, 22, 23, 34, 35, 52])
@pytest.mark.parametrize("stag", ["HTC", "HC"])
def test_testcase_state_vs_harness(tmp_path, tc, stag):
    spec = f"tc={tc}" + (",stag=HC" if stag == "HC" else "")
    subprocess.run([REF_HARNESS, "run", "build:4:eqs", "tc", "0", "900", "0", "out.bin", spec], cwd=tmp_path,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True)
    r = imxbin.read(os.path.join(tmp_path, "out.bin"))
    h, u, bt = Mesh.icos(4).testcase(tc, stag)
    assert np.array_equal(h, r["h0"]) and np.array_equal(u, r["u0"]) and np.array_equal(bt, r["bt"])


def test_testcase_errors():
    m = Mesh.icos(2)
    with pytest.raises(capi.ImxError, match="not part of this build"):
        m.testcase(24)
    with pytest.raises(capi.ImxError, match="stag"):
        capi.check(capi.lib().imx_ic_testcase(m._handle, 2, 7, 6.37122e6, 2.0, capi.f64p(np.zeros(m.nv)),
                                              capi.f64p(np.zeros(m.ne)), capi.f64p(np.zeros(m.nv))), "tc")


@pytest.mark.skipif(not ref_harness_available(), reason="reference harness not built (oracle/Makefile.ref)")
@pytest.mark.parametrize("tc", [5, 6, 21])
def test_testcase_state_pol_vs_harness(tmp_path, tc):
    """Nodes on the poles (the jet table's last interval, cos(lat) = 0)."""
    subprocess.run([REF_HARNESS, "run", "build:3:pol", "tc", "0", "900", "0", "out.bin", f"tc={tc}"],
                   cwd=tmp_path, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True)
    r = imxbin.read(os.path.join(tmp_path, "out.bin"))
    h, u, bt = Mesh.icos(3, "pol").testcase(tc)
    assert np.array_equal(h, r["h0"]) and np.array_equal(u, r["u0"]) and np.array_equal(bt, r["bt"])


def _oracle(golden, name):
    tc, stag = CASES[name]
    fsphere, fcte = Mesh.testcase_coriolis(tc)
    return Oracle(golden("ref_mesh_g3.bin"), testcase=tc, scheme={"stag": stag} if stag == "HC" else None,
                  fsphere=fsphere, fcte=fcte)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_with_topography_bitexact(golden, name):
    r = golden(f"ref_run_g3_{name}.bin")
    o = _oracle(golden, name)
    o.bt = np.array(r["bt"])
    m, f = o.tendency(r["h0"], r["u0"])
    assert np.array_equal(m, r["masseq"]) and np.array_equal(f, r["momeq"])
    for k in FIELDS:
        assert np.array_equal(o.field(k.lower()), r[k]), k
    d0 = o.diagnostics(r["h0"], r["u0"], refresh=True)
    assert np.array_equal(d0[[0, 1, 2, 3, 4, 6]], r["diag0"])
    ser = r["series"].reshape(-1, 7)
    h, u = r["h0"], r["u0"]
    for k in range(ser.shape[0]):
        h, u = o.step(h, u, 900.0, 1)
        d = o.diagnostics(h, u, refresh=False)
        assert np.array_equal(d[[0, 1, 2, 3, 4, 6]], ser[k, 1:]), k
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


def test_coriolis_of_test_cases():
    """The f-sphere settings initialize_fields makes: f = 0 for the
    non-rotating Hollingsworth and localised cases, 1.4584e-4 on the
    f-sphere of 42 and the linearised 11."""
    assert Mesh.testcase_coriolis(2) == (0, 0.0)
    assert Mesh.testcase_coriolis(32) == (0, 0.0) and Mesh.testcase_coriolis(34) == (0, 0.0)
    for tc in (33, 35, 40, 41):
        assert Mesh.testcase_coriolis(tc) == (2, 0.0)
    assert Mesh.testcase_coriolis(42) == (2, 1.4584e-4)
    assert Mesh.testcase_coriolis(11) == (1, 1.4584e-4)


def test_wrong_coriolis_is_caught(golden):
    """Test case 42 on the f-sphere: the oracle with the variable f
    disagrees with the reference's tendency."""
    r = golden("ref_run_g3_tc42.bin")
    o = Oracle(golden("ref_mesh_g3.bin"), testcase=42)
    m, f = o.tendency(r["h0"], r["u0"])
    assert not np.array_equal(f, r["momeq"])


def test_dropped_topography_is_caught(golden):
    """The mountain of test case 5 enters the momentum tendency and the
    energies: the oracle without bt disagrees with the fixture."""
    r = golden("ref_run_g3_tc5.bin")
    o = _oracle(golden, "tc5")
    m, f = o.tendency(r["h0"], r["u0"])
    assert np.array_equal(m, r["masseq"])          # the mass equation does not see b
    assert not np.array_equal(f, r["momeq"])
    d0 = o.diagnostics(r["h0"], r["u0"], refresh=True)
    assert d0[1] != r["diag0"][1] and d0[4] != r["diag0"][4]
