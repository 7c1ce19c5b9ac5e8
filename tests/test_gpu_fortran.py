"""The Fortran host (imodel driver + swm module over ISO_C_BINDING) on the GPU.

Runs imodel_amd/fortran/imodel with the reference's positional par files on
glevel 3, TC2, dt = 900 s, one day, FP64_REF: its final state must equal the
oracle's bit for bit, and its write_evol file must carry the reference's
layout (src/swm.f90:3088-3152).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from imodel_amd import imxbin
from imodel_amd.build import IMODEL
from imodel_amd.swm import Mesh
from oracle.oracle import REF_HARNESS, Oracle, ref_harness_available

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _par(td, nv, testcase, dt, days, nprints, precision, dump=1, loadable=0, nplots=0):
    os.makedirs(os.path.join(td, "par"), exist_ok=True)
    src = os.path.join(ROOT, "par")
    mesh = open(os.path.join(src, "mesh.par")).read().splitlines()
    mesh[2] = str(nv)
    mesh[10] = str(loadable)
    open(os.path.join(td, "par", "mesh.par"), "w").write("\n".join(mesh) + "\n")
    swm = open(os.path.join(src, "swm.par")).read().splitlines()
    swm[2] = f"{testcase} 0"
    swm[4] = f"{days} {days}"
    swm[6] = f"{dt} 0 0"
    swm[22] = f"{nplots} {nprints} 0"
    open(os.path.join(td, "par", "swm.par"), "w").write("\n".join(swm) + "\n")
    gpu = open(os.path.join(src, "swm_gpu.par")).read().splitlines()
    gpu[1] = precision
    gpu[7] = str(dump)
    open(os.path.join(td, "par", "swm_gpu.par"), "w").write("\n".join(gpu) + "\n")


def _error_calc(f, fref, w):
    """error_calc (src/swm.f90:3492-3543) with sumfsq_areas weights."""
    e2 = np.sqrt(np.sum(w * (f - fref) ** 2) / np.sum(w)) / np.sqrt(np.sum(w * fref ** 2) / np.sum(w))
    return np.abs(f - fref).max() / np.abs(fref).max(), e2


def test_fortran_driver_tc2_day_bitexact():
    """TC2 has an analytic solution, so the driver writes the reference's
    error outputs: the _evolution_ file every nprints steps (errors of h, u,
    q_tr, max div, max grad K) and data/swm_errors.txt at the end; the
    final state equals the oracle's bit for bit."""
    print("[gpu-case] fortran imodel g3 tc2 fp64_ref", flush=True)
    assert os.access(IMODEL, os.X_OK), "Fortran host not built"
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 24, "fp64_ref")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", STATE))
        base = "swm_tc2_dt900_HTC_trsk10_areageo"
        evo = open(os.path.join(td, "data", f"{base}_evolution_icos_eqs_nopt_3.txt")).read().splitlines()
        me = open(os.path.join(td, "data", f"{base}_mass_energy_icos_eqs_nopt_3.txt")).read().splitlines()
        errs = open(os.path.join(td, "data", "swm_errors.txt")).read().splitlines()
    m = Mesh.icos(3)
    a = m.arrays
    h0, u0 = m.tc2()
    o = Oracle(a)
    h, u = o.step(h0, u0, 900.0, 96)
    assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)
    # the mass/energy file holds only the initial absolute values (swm.f90:231)
    assert len(me) == 2 and int(me[1].split()[3]) == 0
    assert evo[0].split()[:4] == ["n", "mvdist", "tcase", "k"]
    rows = [line.split() for line in evo[1:]]
    assert [int(x[3]) for x in rows] == [0, 24, 48, 72, 96]
    assert all(float(v) == 0.0 for v in rows[0][8:14])
    q_exact = np.zeros(m.nt)
    from imodel_amd import capi
    capi.check(capi.lib().imx_ic_tc2_pv(m._handle, 6.37122e6, capi.f64p(q_exact)), "imx_ic_tc2_pv")
    A, At = a["hx_areag"], a["tr_areag"]
    We = a["ed_leng"] * a["edhx_leng"] * 0.5
    ev = a["ed_v"].reshape(-1, 2) - 1
    hk, uk = h0, u0
    for row in rows[1:]:
        hk, uk = o.step(hk, uk, 900.0, 24)
        q, divu, ke = o.field("q_tr"), o.field("divu"), o.field("ke_hx")
        want = [*_error_calc(hk, h0, A), *_error_calc(uk, u0, We), *_error_calc(q, q_exact, At),
                np.sqrt(np.sum(A * divu ** 2) / np.sum(A)), np.abs(divu).max(),
                np.abs((ke[ev[:, 1]] - ke[ev[:, 0]]) / a["ed_leng"] / 6.37122e6).max()]
        got = [float(row[i]) for i in (8, 9, 10, 11, 12, 13, 19, 20, 21)]
        assert np.allclose(got, want, rtol=1e-9, atol=0), (row[3], got, want)
        assert float(row[22]) == 1.0
    e = errs[1].split()
    assert errs[0].split()[:3] == ["Mesh", "TCase", "nTime"] and len(errs) == 2
    assert (int(e[0]), int(e[1]), int(e[2])) == (642, 2, 96)
    assert np.allclose([float(x) for x in e[5:9]], [float(rows[-1][i]) for i in (8, 9, 10, 11)], rtol=1e-9)


def test_fortran_driver_seeded_fp32_runs():
    print("[gpu-case] fortran imodel g5 seeded fp32", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 10242, 100, 900, 1, 96, "fp32", dump=0)
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        evol = open(os.path.join(td, "data", "swm_tc100_dt900_HTC_trsk10_areageo_mass_energy_icos_eqs_nopt_5.txt")).read().splitlines()
    last = evol[-1].split()
    assert int(last[3]) == 96
    assert abs(float(last[8])) < 1e-8       # mass drift, FP32 state
    assert abs(float(last[11])) < 1e-4      # total energy drift


STATE = "swm_tc2_dt900_HTC_trsk10_areageo_state_icos_eqs_nopt_3.bin"


@pytest.mark.skipif(not ref_harness_available(), reason="reference harness not built (oracle/Makefile.ref)")
def test_fortran_driver_runs_reference_grid_files():
    """loadable = 1: the driver loads the grid files the reference's own
    meshbuild stored (its numbering, not ours) and its one-day FP64_REF run
    equals the oracle on the reference's in-memory mesh bit for bit."""
    print("[gpu-case] fortran imodel g3 tc2 on reference grid files", flush=True)
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([REF_HARNESS, "build", "3", "eqs", os.path.join(td, "ref.bin")], cwd=td, check=True,
                       stdout=subprocess.DEVNULL)
        ref = imxbin.read(os.path.join(td, "ref.bin"))
        shutil.rmtree(os.path.join(td, "par"))
        _par(td, 642, 2, 900, 1, 24, "fp64_ref", loadable=1)
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        assert "loading grid/icos_eqs_nopt_3" in r.stdout
        st = imxbin.read(os.path.join(td, "data", STATE))
        m = Mesh.load_ref(os.path.join(td, "grid"), "icos_eqs_nopt_3")
    h0, u0 = m.tc2()
    h, u = Oracle(ref).step(h0, u0, 900.0, 96)
    assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)


def test_fortran_driver_stores_then_loads_grid():
    """loadable = 1 without grid files: the driver builds and stores the grid
    (meshstore); a second run loads it and reproduces the first bit for bit."""
    print("[gpu-case] fortran imodel g3 store/load grid", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 24, "fp64_ref", loadable=1)
        runs = []
        for _ in range(2):
            r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True, timeout=120)
            assert r.returncode == 0, r.stdout[-2000:]
            runs.append((r.stdout, imxbin.read(os.path.join(td, "data", STATE))))
        assert len([f for f in os.listdir(os.path.join(td, "grid")) if f.endswith(".dat")]) == 11
    assert "loading grid/" not in runs[0][0] and "loading grid/icos_eqs_nopt_3" in runs[1][0]
    assert np.array_equal(runs[0][1]["h"], runs[1][1]["h"]) and np.array_equal(runs[0][1]["u"], runs[1][1]["u"])


def test_fortran_driver_plot_files():
    """nplots = 2: plotfields writes the reference's lat-lon streams at t = 0,
    12 h and 24 h; h and pv at the end equal the lat-lon sampling of the
    oracle's final state and last-stage q_tr."""
    print("[gpu-case] fortran imodel g3 tc2 plots", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 24, "fp64_ref", nplots=2)
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        base = "swm_tc2_dt900_HTC_trsk10_areageo"
        names = set(os.listdir(os.path.join(td, "data")))
        for t in (0, 43200, 86400):
            for f in ("h", "eta", "Kenergy", "pv", "pv_ed", "divuh", "h_error", "u_error", "q_tr_error"):
                n = f"{base}_{f}_t{t}_icos_eqs_nopt_3.dat"
                assert n in names, n
                assert os.path.getsize(os.path.join(td, "data", n)) == 720 * 1440 * 3 * 4
        hp = np.fromfile(os.path.join(td, "data", f"{base}_h_t86400_icos_eqs_nopt_3.dat"), np.float32).reshape(-1, 3)
        qp = np.fromfile(os.path.join(td, "data", f"{base}_pv_t86400_icos_eqs_nopt_3.dat"), np.float32).reshape(-1, 3)
    m = Mesh.icos(3)
    h0, u0 = m.tc2()
    o = Oracle(m.arrays)
    h, u = o.step(h0, u0, 900.0, 96)
    assert np.array_equal(hp, m.remap_ll(h, 0))
    assert np.array_equal(qp, m.remap_ll(o.field("q_tr"), 1))


def test_fortran_driver_scheme_options_bitexact():
    """APVM, diffusion (align) and hyperdiffusion (diam) read from par/swm.par
    as swmpars does: the file names carry the reference's suffixes and the
    final state equals the oracle with the same options bit for bit."""
    print("[gpu-case] fortran imodel g3 tc2 fp64_ref apvm+diffusion", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 48, "fp64_ref")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        swm[24] = "1e5 align"
        swm[26] = "1e15 diam"
        swm[32] = "apvm 0.5"
        open(p, "w").write("\n".join(swm) + "\n")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        base = ("swm_tc2_dt900_HTC_trsk10_areageo_pvsapvm0.50_align_diffusion_10to5.000"
                "_diam_hyperdiffusion_10to15.000")
        st = imxbin.read(os.path.join(td, "data", f"{base}_state_icos_eqs_nopt_3.bin"))
        assert os.path.exists(os.path.join(td, "data", f"{base}_evolution_icos_eqs_nopt_3.txt"))
    m = Mesh.icos(3)
    h0, u0 = m.tc2()
    o = Oracle(m.arrays, scheme={"pv": "apvm", "pvspar": 0.5, "dt": 900.0, "k2": 1e5, "diffus": "align",
                                 "k4": 1e15, "hyperdiffus": "diam"})
    h, u = o.step(h0, u0, 900.0, 96)
    assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)


def test_fortran_driver_scheme_needs_fp64_ref():
    print("[gpu-case] fortran imodel scheme on fp32 refused", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 48, "fp32")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        swm[10] = "none"
        swm[14] = "pered"
        swm[16] = "gass"
        open(p, "w").write("\n".join(swm) + "\n")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode != 0 and "fp64_ref" in r.stdout


FUSED_OPTIONS = {
    "k2": ({24: "1e5 align"}, {"k2": 1e5, "diffus": "align"}, "_align_diffusion_10to5.000"),
    "mix": ({24: "1e5 align", 26: "1e14 const", 32: "apvm 0.5"},
            {"k2": 1e5, "diffus": "align", "k4": 1e14, "hyperdiffus": "const", "pv": "apvm", "pvspar": 0.5},
            "_pvsapvm0.50_align_diffusion_10to5.000_const_hyperdiffusion_10to14.000"),
    "gass18": ({10: "gass18"}, {"wrap": "gass18"}, ""),
    "gassapvm": ({10: "gass18", 32: "apvm 0.5"}, {"wrap": "gass18", "pv": "apvm", "pvspar": 0.5}, "_pvsapvm0.50"),
    "hcpx": ({8: "HC", 10: "pxt16"}, {"stag": "HC", "wrap": "pxt16"}, ""),
}


@pytest.mark.parametrize("name", sorted(FUSED_OPTIONS))
def test_fortran_driver_fp32_options(name):
    """Diffusion (and, in "mix", hyperdiffusion and APVM; the gass18 wrapper
    with and without APVM; pxt16 with HC) run on the fused FP32 path from
    par/swm.par: a day of TC2 on glevel 3 against the oracle with the same
    options (FP32 storage tolerances), and the options' effect over the day
    is far above the FP32 error."""
    print(f"[gpu-case] fortran imodel g3 tc2 fp32 {name}", flush=True)
    lines, sc, suffix = FUSED_OPTIONS[name]
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 48, "fp32")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        for k, v in lines.items():
            swm[k] = v
        open(p, "w").write("\n".join(swm) + "\n")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        stag = lines.get(8, "HTC")
        base = f"swm_tc2_dt900_{stag}_{lines.get(10, 'trsk10')}_areageo" + suffix
        st = imxbin.read(os.path.join(td, "data", f"{base}_state_icos_eqs_nopt_3.bin"))
    m = Mesh.icos(3)
    h0, u0 = m.tc2(stag=stag)
    h, u = Oracle(m.arrays, scheme=dict(sc, dt=900.0)).step(h0, u0, 900.0, 96)
    # the same staggering without the options
    hp, up = Oracle(m.arrays, scheme={"stag": stag, "dt": 900.0} if stag == "HC" else None).step(h0, u0, 900.0, 96)
    A = m.arrays["hx_areag"]
    eh = np.sqrt(np.sum(A * (st["h"] - h) ** 2) / np.sum(A * h ** 2))
    eu = np.linalg.norm(st["u"] - u) / np.linalg.norm(u)
    du = np.linalg.norm(up - u) / np.linalg.norm(u)
    assert eh < 1e-7 and eu < 1e-5 and du > 20 * eu, (eh, eu, du)


@pytest.mark.parametrize("wrap", ["pxt16", "gass18"])
def test_fortran_driver_wrappers_bitexact(wrap):
    """The pxt16 / gass18 wrapper line of par/swm.par: a day of TC2 on
    glevel 3 equals the oracle with that wrapper bit for bit."""
    print(f"[gpu-case] fortran imodel g3 tc2 fp64_ref {wrap}", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 48, "fp64_ref")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        swm[10] = wrap
        open(p, "w").write("\n".join(swm) + "\n")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", f"swm_tc2_dt900_HTC_{wrap}_areageo_state_icos_eqs_nopt_3.bin"))
    m = Mesh.icos(3)
    h0, u0 = m.tc2()
    o = Oracle(m.arrays, scheme={"wrap": wrap, "dt": 900.0})
    h, u = o.step(h0, u0, 900.0, 96)
    assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)


@pytest.mark.skipif(not ref_harness_available(), reason="reference harness not built (oracle/Makefile.ref)")
def test_fortran_driver_g5_tc5_day_vs_reference_end_to_end():
    """Both sides end to end from the same parameters: the reference builds
    its own glevel-5 mesh (meshbuild + trmesh), initialises test case 5 with
    its mountain and runs 96 steps of 900 s; the Fortran driver does the same
    from par/ files on the GPU (FP64_REF). The final states are identical."""
    print("[gpu-case] fortran imodel g5 tc5 day vs reference meshbuild end to end", flush=True)
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([REF_HARNESS, "run", "build:5:eqs", "tc", "96", "900", "0", "ref.bin", "tc=5"], cwd=td,
                       check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
        r = imxbin.read(os.path.join(td, "ref.bin"))
        _par(td, 10242, 5, 900, 1, 96, "fp64_ref")
        res = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             timeout=300)
        assert res.returncode == 0, res.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", "swm_tc5_dt900_HTC_trsk10_areageo_state_icos_eqs_nopt_5.bin"))
    assert np.array_equal(st["h"], r["h"]) and np.array_equal(st["u"], r["u"])


@pytest.mark.parametrize("tc,tag", [(33, "_hol2.00"), (42, ""), (11, "")])
def test_fortran_driver_fsphere_cases_bitexact(golden, tc, tag):
    """Test cases that set the Coriolis term (33: f = 0 under the
    Hollingsworth layer, 42: the f-sphere, 11: the linearised equations):
    3 steps from par/ files equal the reference's run of the same case."""
    print(f"[gpu-case] fortran imodel g3 tc{tc} fp64_ref f-sphere", flush=True)
    r = golden(f"ref_run_g3_tc{tc}.bin")
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, tc, 900, 0.03125, 3, "fp64_ref")
        res = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", f"swm_tc{tc}_dt900_HTC_trsk10_areageo{tag}_state_icos_eqs_nopt_3.bin"))
    assert np.array_equal(st["h"], r["h"]) and np.array_equal(st["u"], r["u"])


# wrapper "none" in par/swm.par (stag, reconmtd + gasscoef, Coriolis,
# sinterpol, PV stabilisation, level model lines) against the reference's
# own 3-step runs from the seeded state of seed 7 (ref_run_g3_m<name>.bin)
DRIVER_METHODS = {
    "mixnone": ({10: "none", 12: "gass 0.5", 14: "dtred", 16: "gass", 32: "apvm 0.4"},
                "HTC_vrecgass0.50_crecdtred_sintgass_grdtrsk_areageo_pvsapvm0.40"),
    "rmelvhc": ({8: "HC", 10: "none", 12: "melv 1.0"}, "HC_vrecmelv_crectrsk_sinttrsk_grdtrsk_areageo"),
    "gassnopv": ({10: "none", 14: "gass", 16: "gass", 34: "1"},
                 "HTC_vrectrsk_crecgass_sintgass_grdtrsk_areageo_nopv"),
}


# par/swm.par lines of options the fused kernels take beyond the wrappers,
# with the reference's own 3-step run of seed 7 (None: the driver's FP64_REF
# run) and a token of the output name
FUSED_METHODS = {
    "nopv": ({34: "1"}, "ref_run_g3_mnopv.bin", "_nopv"),
    "hcnopv": ({8: "HC", 34: "1"}, None, "_nopv"),
    "rgass": ({10: "none", 12: "gass 0.75"}, "ref_run_g3_mrgass.bin", "vrecgass0.75"),
    "rdubos": ({10: "none", 12: "dubos 1.0"}, "ref_run_g3_mrdubos.bin", "vrecdubos"),
    "rmelv": ({10: "none", 12: "melv 1.0"}, "ref_run_g3_mrmelv.bin", "vrecmelv"),
    "rmelvhc": ({8: "HC", 10: "none", 12: "melv 1.0"}, "ref_run_g3_mrmelvhc.bin", "vrecmelv"),
    "cdtred": ({10: "none", 14: "dtred"}, "ref_run_g3_mcdtred.bin", "crecdtred"),
    "cdtredhc": ({8: "HC", 10: "none", 14: "dtred"}, "ref_run_g3_mcdtredhc.bin", "crecdtred"),
    "baryperhx": ({10: "none", 12: "perhx 1.0", 16: "bary"}, "ref_run_g3_mbaryperhx.bin", "vrecperhx"),
}


@pytest.mark.parametrize("name", sorted(FUSED_METHODS))
def test_fortran_driver_fused_level_model(golden, name):
    """The level-model line (noPV) with trsk10 and Gassmann's kinetic energy
    (reconmtd gass / dubos) on the fused FP32 path: 3 steps of the seeded
    state against the reference's own run (or the driver's FP64_REF run), to
    the FP32 storage tolerances."""
    print(f"[gpu-case] fortran imodel g3 seeded fp32 fused methods {name}", flush=True)
    lines, fixture, token = FUSED_METHODS[name]

    def run(prec):
        with tempfile.TemporaryDirectory() as td:
            _par(td, 642, 100, 900, 0.03125, 3, prec)
            p = os.path.join(td, "par", "swm.par")
            swm = open(p).read().splitlines()
            for k, v in lines.items():
                swm[k] = v
            open(p, "w").write("\n".join(swm) + "\n")
            p = os.path.join(td, "par", "swm_gpu.par")
            gpu = open(p).read().splitlines()
            gpu[3] = "7"
            open(p, "w").write("\n".join(gpu) + "\n")
            res = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                 text=True, timeout=120)
            assert res.returncode == 0, res.stdout[-2000:]
            f = [x for x in os.listdir(os.path.join(td, "data")) if x.endswith("_state_icos_eqs_nopt_3.bin")]
            assert len(f) == 1 and token in f[0], f
            return imxbin.read(os.path.join(td, "data", f[0]))

    st = run("fp32")
    r = golden(fixture) if fixture else run("fp64_ref")
    A = Mesh.icos(3).arrays["hx_areag"]
    eh = np.sqrt(np.sum(A * (st["h"] - r["h"]) ** 2) / np.sum(A * r["h"] ** 2))
    eu = np.linalg.norm(st["u"] - r["u"]) / np.linalg.norm(r["u"])
    assert eh < 1e-7 and eu < 1e-5, (eh, eu)


@pytest.mark.parametrize("name", sorted(DRIVER_METHODS))
def test_fortran_driver_methods_bitexact(golden, name):
    """The individual method lines read as swmpars reads them: the output
    names carry the reference's method suffixes and 3 steps equal the
    reference's bit for bit."""
    print(f"[gpu-case] fortran imodel g3 seeded fp64_ref methods {name}", flush=True)
    lines, tag = DRIVER_METHODS[name]
    r = golden(f"ref_run_g3_m{name}.bin")
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 100, 900, 0.03125, 3, "fp64_ref")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        for k, v in lines.items():
            swm[k] = v
        open(p, "w").write("\n".join(swm) + "\n")
        p = os.path.join(td, "par", "swm_gpu.par")
        gpu = open(p).read().splitlines()
        gpu[3] = "7"
        open(p, "w").write("\n".join(gpu) + "\n")
        res = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True, timeout=120)
        assert res.returncode == 0, res.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", f"swm_tc100_dt900_{tag}_state_icos_eqs_nopt_3.bin"))
    assert np.array_equal(st["h"], r["h"]) and np.array_equal(st["u"], r["u"])


def test_fortran_driver_nopv_pered_refused():
    """noPV with the pered Coriolis term stops as swmpars does."""
    print("[gpu-case] fortran imodel nopv + pered refused", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 48, "fp64_ref")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        swm[10], swm[14], swm[34] = "none", "pered", "1"
        open(p, "w").write("\n".join(swm) + "\n")
        res = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True, timeout=120)
        assert res.returncode != 0 and "level model (noPV)" in res.stdout


def test_fortran_driver_hc_tc2_bitexact():
    """HC staggering from par/swm.par: TC2 initialised on the Voronoi edge
    midpoints, a day on glevel 3 equal to the oracle with HC bit for bit."""
    print("[gpu-case] fortran imodel g3 tc2 fp64_ref HC", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 48, "fp64_ref")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        swm[8] = "HC"
        open(p, "w").write("\n".join(swm) + "\n")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", "swm_tc2_dt900_HC_trsk10_areageo_state_icos_eqs_nopt_3.bin"))
    m = Mesh.icos(3)
    h0, u0 = m.tc2(stag="HC")
    o = Oracle(m.arrays, scheme={"stag": "HC", "dt": 900.0})
    h, u = o.step(h0, u0, 900.0, 96)
    assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)


@pytest.mark.skipif(not ref_harness_available(), reason="oracle/_ref/ref_harness not built")
@pytest.mark.parametrize("stag", ["HTC", "HC"])
def test_fortran_driver_lte(golden, stag):
    """simulcase 14 (swm_horiz_loc_trunc_er) from par/mesh.par, on the grid
    files the reference's meshbuild stored: the printed table equals the
    reference's error triples to its 8 printed digits, the mimetic sums
    (printed with 17) bit for bit, and data/swm_loc_trunc_errors.txt has the
    reference's header and row layout."""
    print(f"[gpu-case] fortran imodel g3 lte {stag}", flush=True)
    from imodel_amd.capi import LTE_ROWS
    r = golden("ref_lte_g3_trsk10.bin" if stag == "HTC" else "ref_lte_g3_hc.bin")
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([REF_HARNESS, "build", "3", "eqs", os.path.join(td, "ref.bin")], cwd=td, check=True,
                       stdout=subprocess.DEVNULL)
        shutil.rmtree(os.path.join(td, "par"))
        _par(td, 642, 2, 900, 1, 48, "fp64_ref", loadable=1)
        p = os.path.join(td, "par", "mesh.par")
        mesh = open(p).read().splitlines()
        mesh[14] = "14"
        open(p, "w").write("\n".join(mesh) + "\n")
        p = os.path.join(td, "par", "swm.par")
        swm = open(p).read().splitlines()
        swm[8] = stag
        open(p, "w").write("\n".join(swm) + "\n")
        res = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True, timeout=120)
        assert res.returncode == 0, res.stdout[-2000:]
        assert "loading grid/icos_eqs_nopt_3" in res.stdout
        lines = open(os.path.join(td, "data", "swm_loc_trunc_errors.txt")).read().splitlines()
    out = res.stdout.splitlines()
    table = {}
    for ln in out:
        f = ln.split()
        if len(f) == 5 and f[1] == "642":
            table[f[0]] = np.array([float(x) for x in f[2:]])
    for row in LTE_ROWS:
        if f"lte_{row}" in r:
            assert np.allclose(table[row], r[f"lte_{row}"], rtol=6e-8, atol=0), (row, table[row])
    mim = {}
    for ln in out:
        t = ln.strip()
        for i, k in enumerate(("Sum of h*Div(u):", "Sum of h*Div(u)+u*Grad(h):", "Sum of Div(hu):",
                               "Difference:", "Sum of k.Rot(Grad(h)):",
                               "Sum of Coriolis energy (Ae*q*h^2*u*u_perp):",
                               "Sum of Coriolis energy / Tenergy:")):
            if t.startswith(k):
                mim[i] = float(t[len(k):])
    assert [mim[i] for i in range(7)] == list(r["mim"])
    # the list-directed header wraps after 80 columns, as the reference's
    # own file does when built with the same compiler
    assert " ".join(lines[:2]).split() == ["Field", "Mesh", "MaxErrorRel", "RMSError", "MaxError", "Methods",
                                           "Grid"]
    assert len(lines) == 2 + len(LTE_ROWS)
    assert lines[2].split()[-2:] == [f"swm_tc2_{stag}_trsk10_areageo", "icos_eqs_nopt_3"]


def test_fortran_driver_read_grid():
    """mesh kind "read" in par/mesh.par: the driver reads grid/<name>.xyz and
    its .ngb (a glevel-3 grid with shuffled nodes and ring starts), stores
    the grid files with loadable = 1, and its one-day FP64_REF TC2 run equals
    the oracle on the same mesh bit for bit; the output files carry the grid
    name."""
    print("[gpu-case] fortran imodel read-kind grid g3", flush=True)
    from tests.test_meshread import _write_grid
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 2, 900, 1, 48, "fp64_ref", loadable=1)
        _write_grid(td, "shuf3", 11, "xyz", glevel=3)
        p = os.path.join(td, "par", "mesh.par")
        mesh = open(p).read().splitlines()
        mesh[4] = "read"
        mesh[16] = "shuf3.xyz"
        open(p, "w").write("\n".join(mesh) + "\n")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        assert os.path.exists(os.path.join(td, "grid", "shuf3_header.dat"))
        st = imxbin.read(os.path.join(td, "data", "swm_tc2_dt900_HTC_trsk10_areageo_state_shuf3.bin"))
        m = Mesh.build_read(os.path.join(td, "grid", "shuf3.xyz"))
    h0, u0 = m.tc2()
    h, u = Oracle(m.arrays).step(h0, u0, 900.0, 96)
    assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)


def test_fortran_caller_reference_surface_bitexact(golden):
    """A Fortran program written against the reference's call surface
    (tests/fortran/swm_caller.f90, shaped as moist_swm's import of
    initialize_gridprop, sumf_areas, calc_energies, tendency, error_calc):
    test case 5 with its mountain on glevel 3, FP64_REF. The tendency of the
    initial state, sumf_areas, calc_energies and three ode_rk4 steps equal
    the reference's own run (ref_run_g3_tc5.bin) bit for bit."""
    print("[gpu-case] fortran reference-surface caller g3 tc5 fp64_ref", flush=True)
    from imodel_amd.build import CALLER
    assert os.access(CALLER, os.X_OK), "test caller not built"
    r = golden("ref_run_g3_tc5.bin")
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 5, 900, 1, 24, "fp64_ref")
        res = subprocess.run([CALLER], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stdout[-2000:]
        out = imxbin.read(os.path.join(td, "out.bin"))
    assert np.array_equal(out["masseq"], r["masseq"]) and np.array_equal(out["momeq"], r["momeq"])
    assert np.array_equal(out["diag0"], r["diag0"][:5])   # mass, P, K, T, available
    assert np.array_equal(out["h"], r["h"]) and np.array_equal(out["u"], r["u"])


def test_fortran_caller_reference_surface_fp32():
    """The same caller on the fused FP32 path: the tendency and the three
    steps against the oracle with the mountain, at the FP32 tolerances."""
    print("[gpu-case] fortran reference-surface caller g3 tc5 fp32", flush=True)
    from imodel_amd.build import CALLER
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 5, 900, 1, 24, "fp32")
        res = subprocess.run([CALLER], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stdout[-2000:]
        out = imxbin.read(os.path.join(td, "out.bin"))
    m = Mesh.icos(3)
    h0, u0, bt = m.testcase(5)
    o = Oracle(m.arrays, testcase=5)
    o.bt = bt
    mo, fo = o.tendency(h0, u0)
    scale = max(np.linalg.norm(fo), np.linalg.norm(o.field("grad_ghbk")))
    assert np.linalg.norm(out["masseq"] - mo) / np.linalg.norm(mo) < 5e-5
    assert np.linalg.norm(out["momeq"] - fo) / scale < 5e-5
    h, u = o.step(h0, u0, 900.0, 3)
    A = m.arrays["hx_areag"]
    assert np.sqrt(np.sum(A * (out["h"] - h) ** 2) / np.sum(A * h ** 2)) < 1e-7
    assert np.linalg.norm(out["u"] - u) / np.linalg.norm(u) < 1e-5


def test_fortran_driver_tc21_shipped_parameters():
    """The reference's shipped par/swm.par runs the Galewsky jet (test case
    21, dt = 120 s, HTC, trsk10, prints every 60 steps); one day of it on
    glevel 3 with FP64_REF equals the oracle bit for bit."""
    print("[gpu-case] fortran imodel g3 tc21 dt120 fp64_ref", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 642, 21, 120, 1, 60, "fp64_ref")
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", "swm_tc21_dt120_HTC_trsk10_areageo_state_icos_eqs_nopt_3.bin"))
        me = open(os.path.join(td, "data", "swm_tc21_dt120_HTC_trsk10_areageo_mass_energy_icos_eqs_nopt_3.txt")
                  ).read().splitlines()
    m = Mesh.icos(3)
    h0, u0, _ = m.testcase(21)
    h, u = Oracle(m.arrays, testcase=21).step(h0, u0, 120.0, 720)
    assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)
    assert [int(x.split()[3]) for x in me[1:]][-1] == 720


@pytest.mark.parametrize("prec", ["fp64_ref", "fp32"])
def test_fortran_driver_tc5_mountain(prec):
    """Test case 5 (flow over the mountain) from par/swm.par: the driver sets
    the topography; a day on glevel 4 equals the oracle with the mountain
    (bit for bit on FP64_REF, to FP32 storage tolerances on FP32) and the
    energies of the mass/energy file carry b."""
    print(f"[gpu-case] fortran imodel g4 tc5 {prec}", flush=True)
    with tempfile.TemporaryDirectory() as td:
        _par(td, 2562, 5, 900, 1, 48, prec)
        r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:]
        st = imxbin.read(os.path.join(td, "data", "swm_tc5_dt900_HTC_trsk10_areageo_state_icos_eqs_nopt_4.bin"))
        me = open(os.path.join(td, "data", "swm_tc5_dt900_HTC_trsk10_areageo_mass_energy_icos_eqs_nopt_4.txt")
                  ).read().splitlines()
    m = Mesh.icos(4)
    h0, u0, bt = m.testcase(5)
    o = Oracle(m.arrays, testcase=5)
    o.bt = bt
    h, u = o.step(h0, u0, 900.0, 96)
    if prec == "fp64_ref":
        assert np.array_equal(st["h"], h) and np.array_equal(st["u"], u)
    else:
        A = m.arrays["hx_areag"]
        assert np.sqrt(np.sum(A * (st["h"] - h) ** 2) / np.sum(A * h ** 2)) < 1e-6
    p0 = float(me[1].split()[9])      # absolute initial potential energy
    want = o.diagnostics(h0, u0, refresh=True)[1]
    # FP32 holds h - h_ref rounded to FP32 (relative 6e-8 of the 10% mountain
    # perturbation) from the first step on, the initial energies included
    assert abs(p0 - want) <= (1e-12 if prec == "fp64_ref" else 1e-8) * abs(want)


def test_fortran_driver_fused_hc_divergence_matches_ref():
    """HC staggering on the fused FP32 path: the evolution file's RMS and max
    div u use the reference's HC signs (they come from a reference-order
    tendency of the state), so they agree with an FP64_REF HC run to FP32
    storage accuracy."""
    print("[gpu-case] fortran imodel g3 tc2 HC fp32 vs fp64_ref divergence", flush=True)
    cols = {}
    for prec in ("fp64_ref", "fp32"):
        with tempfile.TemporaryDirectory() as td:
            _par(td, 642, 2, 900, 1, 24, prec)
            p = os.path.join(td, "par", "swm.par")
            swm = open(p).read().splitlines()
            swm[8] = "HC"
            open(p, "w").write("\n".join(swm) + "\n")
            r = subprocess.run([IMODEL], cwd=td, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True, timeout=120)
            assert r.returncode == 0, r.stdout[-2000:]
            evo = open(os.path.join(td, "data", "swm_tc2_dt900_HC_trsk10_areageo_evolution_icos_eqs_nopt_3.txt")
                       ).read().splitlines()
        cols[prec] = np.array([[float(x) for x in ln.split()[19:21]] for ln in evo[2:]])
    # fp64_ref reads the last RK stage's div u, fp32 the state's, so they
    # differ; HTC signs on an HC run (half the terms flipped) would be off
    # by orders of magnitude
    ratio = cols["fp32"] / cols["fp64_ref"]
    assert np.all((ratio > 0.5) & (ratio < 2.0)), cols
