"""initialize_fields' test cases with bottom topography on the GPU.

* FP64_REF: the tendency, its intermediates and 3 RK4 steps equal the
  reference's fixtures bit for bit for test cases 1, 5 (HTC and HC), 51, 6,
  21, 53, 11, 32, 33, 34 (HC), 40, 41 and 42 (HTC and HC), with the state,
  topography and Coriolis setting (f-sphere) of imx_ic_testcase.
* the fused FP64 and FP32 paths against the oracle: the tendency, the
  diagnostics with b, and one simulated day of test case 5 at glevel 4.
"""
import numpy as np
import pytest

from imodel_amd.swm import ERAD_REF, Mesh, Swm
from oracle.oracle import Oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = {"tc1": (1, "HTC"), "tc5": (5, "HTC"), "tc5hc": (5, "HC"), "tc51": (51, "HTC"), "tc6": (6, "HTC"),
         "tc21": (21, "HTC"), "tc53": (53, "HTC"), "tc11": (11, "HTC"), "tc32": (32, "HTC"),
         "tc33": (33, "HTC"), "tc34hc": (34, "HC"), "tc40": (40, "HTC"), "tc41": (41, "HTC"),
         "tc42": (42, "HTC"), "tc42hc": (42, "HC")}
FIELDS = ["h_ed", "uh", "eta", "h_tr", "q_tr", "q_ed", "divuh", "ke_hx", "ghbK", "grad_ghbK", "uhq_perp"]


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def g3():
    return Mesh.icos(3)


@pytest.mark.parametrize("name", sorted(CASES))
def test_ref_testcase_bitexact(g3, golden, name):
    say("fp64_ref test case", name)
    tc, stag = CASES[name]
    r = golden(f"ref_run_g3_{name}.bin")
    h, u, bt = g3.testcase(tc, stag)
    with Swm(g3, precision="fp64_ref", dt=900.0, testcase=tc) as s:
        if stag == "HC":
            s.set_scheme(stag="HC")
        s.set_topography(bt)
        s.set_coriolis(*Mesh.testcase_coriolis(tc))
        m, f = s.tendency(h, u)
        assert np.array_equal(m, r["masseq"]) and np.array_equal(f, r["momeq"])
        for k in FIELDS:
            assert np.array_equal(s.field(k.lower()), r[k]), k
        s.set_state(h, u)
        s.step(3)
        hg, ug = s.get_state()
    assert np.array_equal(hg, r["h"]) and np.array_equal(ug, r["u"])


def flux_terms(m, uh):
    """Norm over cells of sum_e |l_e (hu)_e| / (A_c a): the size of the terms
    the mass tendency (a flux divergence) sums."""
    a = m.arrays
    nv = a["hx_areag"].size
    ve = a["v_ed"].reshape(nv, 6)
    e = np.maximum(ve - 1, 0)
    w = (ve > 0) * np.abs(a["edhx_leng"][e] * uh[e])
    return np.linalg.norm(w.sum(1) / a["hx_areag"]) / ERAD_REF


@pytest.mark.parametrize("name", ["tc5", "tc5hc", "tc6", "tc21", "tc11", "tc33", "tc42hc"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 5e-5)])
def test_fast_testcase_tendency_vs_oracle(name, prec, tol):
    say("fast tendency", name, prec)
    tc, stag = CASES[name]
    m = Mesh.icos(4)
    h, u, bt = m.testcase(tc, stag)
    fsphere, fcte = Mesh.testcase_coriolis(tc)
    o = Oracle(m.arrays, testcase=tc, scheme={"stag": stag} if stag == "HC" else None, fsphere=fsphere,
               fcte=fcte)
    o.bt = bt
    mo, fo = o.tendency(h, u)
    # the balanced states (5, 6, 21: geostrophic to a few per cent) have a
    # momentum tendency far below its two terms, and the zonal flows (2, 33)
    # a mass tendency far below its fluxes' divergence terms; errors are
    # measured against the larger of the tendency and its terms
    scale = max(np.linalg.norm(fo), np.linalg.norm(o.field("grad_ghbk")))
    mscale = max(np.linalg.norm(mo), flux_terms(m, o.field("uh")))
    with Swm(m, precision=prec, testcase=tc) as s:
        if stag == "HC":
            s.set_scheme(stag="HC")
        s.set_topography(bt)
        s.set_coriolis(fsphere, fcte)
        mg, fg = s.tendency(h, u)
    # the fluid at rest of test case 11 has a zero tendency: absolute then
    em = np.linalg.norm(mg - mo) / max(mscale, 1e-300)
    ef = np.linalg.norm(fg - fo) / max(scale, 1e-300)
    assert em < tol and ef < tol, (em, ef)


@pytest.mark.parametrize("prec,tol", [("fp64_ref", 1e-12), ("fp64", 1e-12), ("fp32", 1e-6)])
def test_diagnostics_with_topography(prec, tol):
    say("diagnostics with topography", prec)
    m = Mesh.icos(4)
    h, u, bt = m.testcase(5)
    o = Oracle(m.arrays, testcase=5)
    o.bt = bt
    d_o = o.diagnostics(h, u, refresh=True)
    with Swm(m, precision=prec, testcase=5) as s:
        s.set_topography(bt)
        s.set_state(h, u)
        d = s.diagnostics()
        s.set_topography(None)
        d_flat = s.diagnostics()
    for k, name in ((0, "mass"), (1, "penergy"), (2, "kenergy"), (3, "tenergy")):
        assert abs(d[name] - d_o[k]) <= tol * abs(d_o[k]), (name, d[name], d_o[k])
    assert abs(d["aenergy"] - d_o[4]) <= tol * abs(d_o[3]), (d["aenergy"], d_o[4])
    assert d_flat["penergy"] < d["penergy"]  # b > 0 raises the potential energy


@pytest.mark.parametrize("order", ["topography_first", "state_first"])
@pytest.mark.parametrize("prec,tol_h,tol_u", [("fp64", 1e-11, 1e-10), ("fp32", 1e-6, 1e-4)])
def test_fast_tc5_one_day_vs_oracle(prec, tol_h, tol_u, order):
    """Flow over the mountain, glevel 4, 96 steps of 900 s, with the
    topography set before or after the state (the first set_topography
    allocates the g b buffer: its zero fill must be complete before the
    handle's stream writes it)."""
    say("fast tc5 day", prec, order)
    m = Mesh.icos(4)
    h, u, bt = m.testcase(5)
    o = Oracle(m.arrays, testcase=5)
    o.bt = bt
    ho, uo = o.step(h, u, 900.0, 96)
    with Swm(m, precision=prec, dt=900.0, testcase=5) as s:
        if order == "topography_first":
            s.set_topography(bt)
            s.set_state(h, u)
        else:
            s.set_state(h, u)
            s.set_topography(bt)
        s.step(96)
        hg, ug = s.get_state()
        dg = s.diagnostics()
    A = m.arrays["hx_areag"]
    say(prec, "h", rel_l2(hg, ho, A), "u", rel_l2(ug, uo))
    assert rel_l2(hg, ho, A) < tol_h and rel_l2(ug, uo) < tol_u
    assert dg["badflag"] == 0


def test_topography_off_restores_plain_step():
    say("topography off")
    m = Mesh.icos(3)
    h, u = m.spectral(3)
    out = []
    for topo in (False, True):
        with Swm(m, precision="fp32", dt=900.0) as s:
            s.set_state(h, u)
            s.step(1)  # graph captured without topography
            if topo:
                s.set_topography(np.full(m.nv, 100.0))
                s.set_topography(None)
            s.set_state(h, u)
            s.step(3)
            out.append(s.get_state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
