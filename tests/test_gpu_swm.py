"""GPU parity of libimx against the reference golden vectors and the oracle.

FP64_REF must be bit-identical to the reference (built -ffp-contract=off);
the fused FP64 path agrees with the oracle to rounding; the FP32 path is held
to the tolerances written in each test.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from imodel_amd import capi, imxbin
from imodel_amd.swm import Mesh, Swm
from oracle.oracle import REF_HARNESS, Oracle, ref_harness_available
from tests.conftest import rel_l2, relmax

pytestmark = pytest.mark.gpu

FIELDS = ["h_ed", "uh", "eta", "h_tr", "q_tr", "q_ed", "divuh", "ke_hx", "ghbK", "grad_ghbK",
          "uhq_perp", "divu", "zeta"]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("sfc", [False, True])
@pytest.mark.parametrize("run", ["ref_run_g3_tc2.bin", "ref_run_g3_spec.bin"])
def test_ref_tendency_bitexact(golden, run, sfc):
    say("ref tendency", run, "sfc", sfc)
    mesh = Mesh(golden("ref_mesh_g3.bin"))
    r = golden(run)
    with Swm(mesh, precision="fp64_ref", dt=900.0, use_sfc=sfc) as s:
        m, f = s.tendency(r["h0"], r["u0"])
        assert np.array_equal(m, r["masseq"])
        assert np.array_equal(f, r["momeq"])
        for name in FIELDS:
            assert np.array_equal(s.field(name.lower()), r[name]), name


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("run", ["ref_run_g3_tc2.bin", "ref_run_g3_spec.bin"])
def test_ref_steps_bitexact(golden, run, graph):
    say("ref steps", run, "graph", graph)
    mesh = Mesh(golden("ref_mesh_g3.bin"))
    r = golden(run)
    with Swm(mesh, precision="fp64_ref", dt=900.0, use_graph=graph) as s:
        s.set_state(r["h0"], r["u0"])
        s.step(5)
        h, u = s.get_state()
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


def test_ref_96_steps_bitexact(golden):
    say("ref 96 steps g2")
    mesh = Mesh(golden("ref_mesh_g2.bin"))
    r = golden("ref_run_g2_spec96.bin")
    with Swm(mesh, precision="fp64_ref", dt=900.0) as s:
        s.set_state(r["h0"], r["u0"])
        s.step(96)
        h, u = s.get_state()
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


@pytest.mark.skipif(not ref_harness_available(), reason="oracle/_ref not built on this host")
def test_g5_day_bitexact_against_reference_binary():
    """glevel 5, dt=900 s, 96 steps, FP64 both sides: the GPU REF path against
    the reference Fortran itself run on the same (libimx-built) mesh."""
    say("g5 day vs reference binary")
    m = Mesh.icos(5)
    h0, u0 = m.spectral(1)
    with tempfile.TemporaryDirectory() as td:
        m.write(os.path.join(td, "mesh.bin"))
        imxbin.write(os.path.join(td, "state.bin"), {"h": h0, "u": u0})
        subprocess.run([REF_HARNESS, "run", "mesh.bin", "state.bin", "96", "900", "0", "out.bin"],
                       cwd=td, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       timeout=300)
        r = imxbin.read(os.path.join(td, "out.bin"))
    with Swm(m, precision="fp64_ref", dt=900.0) as s:
        m0, f0 = s.tendency(h0, u0)
        assert np.array_equal(m0, r["masseq"]) and np.array_equal(f0, r["momeq"])
        s.set_state(h0, u0)
        s.step(96)
        h, u = s.get_state()
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


def _case(g, seed):
    m = Mesh.icos(g)
    h, u = m.spectral(seed)
    return m, h, u


# FP32: u and h - h_ref are stored in FP32 (relative rounding 6e-8); the
# divergence and the Bernoulli gradient are differences of O(100x) larger
# terms, so their relative error is ~1e-5.
@pytest.mark.parametrize("prec,tol_m,tol_f", [("fp64", 1e-12, 1e-12), ("fp32", 5e-5, 5e-5)])
def test_fast_tendency_vs_oracle(prec, tol_m, tol_f):
    say("fast tendency", prec)
    m, h, u = _case(4, 21)
    o = Oracle(m.arrays)
    mo, fo = o.tendency(h, u)
    with Swm(m, precision=prec) as s:
        mg, fg = s.tendency(h, u)
    # relative L2 against the oracle's field (tolerances: FP64 rounding; FP32
    # storage of u and h - h_ref with FP64 mass-flux accumulation)
    assert rel_l2(mg, mo) < tol_m, rel_l2(mg, mo)
    assert rel_l2(fg, fo) < tol_f, rel_l2(fg, fo)


@pytest.mark.parametrize("prec,tol_h,tol_u", [("fp64", 1e-12, 1e-11), ("fp32", 1e-7, 1e-5)])
def test_fast_steps_vs_oracle(prec, tol_h, tol_u):
    say("fast steps", prec)
    m, h, u = _case(4, 22)
    o = Oracle(m.arrays)
    ho, uo = o.step(h, u, 900.0, 10)
    with Swm(m, precision=prec, dt=900.0) as s:
        s.set_state(h, u)
        s.step(10)
        hg, ug = s.get_state()
    A = m.arrays["hx_areag"]
    # h: area-weighted relative L2 of the full depth, as error_calc (swm.f90:3524)
    assert rel_l2(hg, ho, A) < tol_h, rel_l2(hg, ho, A)
    assert rel_l2(ug, uo) < tol_u, rel_l2(ug, uo)


@pytest.mark.parametrize("prec,tol,tol_div", [("fp64_ref", 1e-12, 1e-10), ("fp64", 1e-12, 1e-10),
                                              ("fp32", 1e-6, 1e-3)])
def test_diagnostics_vs_oracle(prec, tol, tol_div):
    say("diagnostics", prec)
    m, h, u = _case(4, 23)
    o = Oracle(m.arrays)
    h, u = o.step(h, u, 900.0, 4)  # develop some divergence
    d_o = o.diagnostics(h, u, refresh=True)
    with Swm(m, precision=prec) as s:
        s.set_state(h, u)
        d = s.diagnostics()
    names = capi.DIAG_NAMES
    for k in (0, 1, 2, 3, 5):
        assert abs(d[names[k]] - d_o[k]) <= tol * abs(d_o[k]), (names[k], d[names[k]], d_o[k])
    # RMS divergence: cancellation in div of FP32-rounded u
    assert abs(d["rmsdiv"] - d_o[6]) <= tol_div * abs(d_o[6]), (d["rmsdiv"], d_o[6])
    # available energy is a difference of large terms: compare against the total
    assert abs(d["aenergy"] - d_o[4]) <= tol * abs(d_o[3])
    assert d["badflag"] == 0.0


def test_graph_equals_eager_fp32():
    say("graph vs eager fp32")
    m, h, u = _case(4, 24)
    out = []
    for graph in (False, True):
        with Swm(m, precision="fp32", dt=900.0, use_graph=graph) as s:
            s.set_state(h, u)
            for _ in range(3):
                s.step(2)
            out.append(s.get_state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_sfc_invariant_fp64():
    say("sfc on/off fp64")
    m, h, u = _case(4, 25)
    res = []
    for sfc in (False, True):
        with Swm(m, precision="fp64", dt=900.0, use_sfc=sfc) as s:
            s.set_state(h, u)
            s.step(4)
            res.append(s.get_state())
    assert relmax(res[0][0], res[1][0]) < 1e-13 and relmax(res[0][1], res[1][1]) < 1e-12


def test_zero_depth_guard_reports_error():
    """h = 0 trips the reference's division guard (swm_operators.f90:1603):
    every kernel still exits; the call returns IMX_ESTATE."""
    say("zero depth guard")
    m, h, u = _case(3, 1)
    for prec in ("fp64_ref", "fp32"):
        with Swm(m, precision=prec) as s:
            with pytest.raises(capi.ImxError, match="division by zero"):
                s.tendency(np.zeros_like(h), u)


@pytest.mark.parametrize("glevel", [0, 1])
def test_smallest_meshes_vs_oracle(glevel):
    """The icosahedron itself (12 pentagons, every cell with a padded slot)
    and glevel 1 (42 cells), both far smaller than one block: FP64_REF equals
    the oracle bit for bit over 5 steps, the fused paths agree to their
    tolerances."""
    say("smallest meshes", glevel)
    m, h, u = _case(glevel, 31)
    o = Oracle(m.arrays)
    mo, fo = o.tendency(h, u)
    ho, uo = o.step(h, u, 900.0, 5)
    A = m.arrays["hx_areag"]
    for prec, tol_f, tol_h, tol_u in (("fp64_ref", 0.0, 0.0, 0.0), ("fp64", 1e-12, 1e-12, 1e-11),
                                      ("fp32", 5e-5, 1e-7, 1e-5)):
        with Swm(m, precision=prec, dt=900.0) as s:
            mg, fg = s.tendency(h, u)
            s.set_state(h, u)
            s.step(5)
            hg, ug = s.get_state()
        if prec == "fp64_ref":
            assert np.array_equal(mg, mo) and np.array_equal(fg, fo)
            assert np.array_equal(hg, ho) and np.array_equal(ug, uo)
        else:
            assert rel_l2(mg, mo) < tol_f and rel_l2(fg, fo) < tol_f, (prec, rel_l2(mg, mo), rel_l2(fg, fo))
            assert rel_l2(hg, ho, A) < tol_h and rel_l2(ug, uo) < tol_u, (prec, rel_l2(hg, ho, A), rel_l2(ug, uo))


def test_nonfinite_state_sets_badflag():
    """A NaN in the state runs through every kernel without a fault and
    raises the diagnostics' NaN/Inf flag, on the reference-order and the
    fused paths (the driver's blow-up stop reads it)."""
    say("nonfinite state")
    m, h, u = _case(3, 5)
    u = u.copy()
    u[17] = np.nan
    for prec in ("fp64_ref", "fp64", "fp32"):
        with Swm(m, precision=prec, dt=900.0) as s:
            s.set_state(h, u)
            assert s.diagnostics()["badflag"] == 1
            s.step(2)
            assert s.diagnostics()["badflag"] == 1
            hg, ug = s.get_state()
        assert not np.all(np.isfinite(hg))


def test_mass_conservation_fp32_day_g6():
    """One simulated day at glevel 6 in FP32: relative mass drift stays at
    FP32-storage level, energy drift small, no NaN."""
    say("fp32 day g6")
    m, h, u = _case(6, 3)
    with Swm(m, precision="fp32", dt=450.0) as s:
        s.set_state(h, u)
        d0 = s.diagnostics()
        s.step(192)
        d1 = s.diagnostics()
    assert d1["badflag"] == 0
    assert abs(d1["mass"] - d0["mass"]) / d0["mass"] < 1e-7
    assert abs(d1["tenergy"] - d0["tenergy"]) / d0["tenergy"] < 1e-4


@pytest.fixture(scope="module")
def g8_case():
    m = Mesh.icos(8)
    h, u = m.spectral(5)
    return m, h, u


def test_ref_g8_bitexact_vs_oracle(g8_case):
    """The configuration the throughput is quoted on (glevel 8, 655,362
    cells, dt = 112.5 s): FP64_REF with SFC ordering and the graph equals the
    oracle bit for bit over 2 RK4 steps."""
    say("fp64_ref g8 2 steps vs oracle")
    m, h, u = g8_case
    ho, uo = Oracle(m.arrays).step(h, u, 112.5, 2)
    with Swm(m, precision="fp64_ref", dt=112.5) as s:
        s.set_state(h, u)
        s.step(2)
        hg, ug = s.get_state()
    assert np.array_equal(hg, ho) and np.array_equal(ug, uo)


def test_fast_g8_vs_oracle(g8_case):
    """Glevel 8, 16 steps: the fused FP64 path to rounding and the FP32 path
    to its storage tolerances (h - h_ref and u in FP32)."""
    say("fast g8 16 steps vs oracle")
    m, h, u = g8_case
    ho, uo = Oracle(m.arrays).step(h, u, 112.5, 16)
    A = m.arrays["hx_areag"]
    for prec, tol_h, tol_u in (("fp64", 1e-12, 1e-11), ("fp32", 1e-7, 1e-5)):
        with Swm(m, precision=prec, dt=112.5) as s:
            s.set_state(h, u)
            s.step(16)
            hg, ug = s.get_state()
        assert rel_l2(hg, ho, A) < tol_h, (prec, rel_l2(hg, ho, A))
        assert rel_l2(ug, uo) < tol_u, (prec, rel_l2(ug, uo))
