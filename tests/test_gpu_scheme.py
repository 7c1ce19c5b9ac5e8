"""Scheme options of the FP64_REF path on the GPU (imx_swm_set_scheme):
HC staggering, the pxt16 and gass18 wrappers, APVM, diffusion and hyperdiffusion with their
coefficient functions, bit for bit against the reference's own golden vectors (tests/golden/make_golden.py)
and against the oracle on a larger grid; the normal-mode Jacobian with the
widened stencil."""
import numpy as np
import pytest

from imodel_amd import capi
from imodel_amd.swm import Mesh, Swm
from oracle.oracle import ERAD_REF, Oracle
from tests.test_gpu_jacobian import per_column
from tests.test_oracle_golden import SCHEMES

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def set_scheme(s, sc):
    s.set_scheme(**{k: v for k, v in sc.items() if k != "dt"})


@pytest.mark.parametrize("sfc", [False, True])
@pytest.mark.parametrize("name", sorted(SCHEMES))
def test_scheme_tendency_and_steps_match_reference(golden, name, sfc):
    say("scheme vs reference", name, "sfc", sfc)
    mesh = Mesh(golden("ref_mesh_g3.bin"))
    r = golden(f"ref_run_g3_{name}.bin")
    with Swm(mesh, precision="fp64_ref", dt=900.0, use_sfc=sfc) as s:
        set_scheme(s, SCHEMES[name])
        m, f = s.tendency(r["h0"], r["u0"])
        assert np.array_equal(m, r["masseq"]) and np.array_equal(f, r["momeq"])
        for k in ("q_grad_ed", "lapu", "lap_lapu", "q_ed", "uhq_perp", "divu", "zeta"):
            if k in r:
                assert np.array_equal(s.field(k), r[k]), k
        s.set_state(r["h0"], r["u0"])
        s.step(3)
        h, u = s.get_state()
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


# wrapper "none": the individual method choices of the reference's
# ref_run_g3_m<name>.bin fixtures (tests/golden/make_golden.py)
METHODS = {
    "rgass": {"wrap": "none", "rec": "gass", "gasscoef": 0.75},
    "rdubos": {"wrap": "none", "rec": "dubos"},
    "rmelv": {"wrap": "none", "rec": "melv"},
    "rmelvhc": {"stag": "HC", "wrap": "none", "rec": "melv"},
    "cdtred": {"wrap": "none", "cor": "dtred"},
    "cdtredhc": {"stag": "HC", "wrap": "none", "cor": "dtred"},
    "nopv": {"wrap": "none", "nopv": True},
    "gassnopv": {"wrap": "none", "sint": "gass", "cor": "gass", "nopv": True},
    "baryperhx": {"wrap": "none", "sint": "bary", "rec": "perhx"},
    "mixnone": {"wrap": "none", "sint": "gass", "cor": "dtred", "rec": "gass", "gasscoef": 0.5, "pv": "apvm",
                "pvspar": 0.4},
}
METHOD_FIELDS = {"h_ed": "h_ed", "uh": "uh", "eta": "eta", "h_tr": "h_tr", "q_tr": "q_tr", "q_ed": "q_ed",
                 "divuh": "divuh", "ke_hx": "ke_hx", "ghbk": "ghbK", "grad_ghbk": "grad_ghbK",
                 "uhq_perp": "uhq_perp", "divu": "divu", "zeta": "zeta", "q_grad_ed": "q_grad_ed"}


@pytest.mark.parametrize("sfc", [False, True])
@pytest.mark.parametrize("name", sorted(METHODS))
def test_methods_match_reference(golden, name, sfc):
    """Each method choice: the tendency, its intermediates and 3 RK4
    steps, bit for bit."""
    say("methods vs reference", name, "sfc", sfc)
    mesh = Mesh(golden("ref_mesh_g3.bin"))
    r = golden(f"ref_run_g3_m{name}.bin")
    with Swm(mesh, precision="fp64_ref", dt=900.0, use_sfc=sfc) as s:
        s.set_scheme(**METHODS[name])
        m, f = s.tendency(r["h0"], r["u0"])
        bad = [k for k, g in METHOD_FIELDS.items() if g in r and not np.array_equal(s.field(k), r[g])]
        assert not bad, bad
        assert np.array_equal(m, r["masseq"]) and np.array_equal(f, r["momeq"])
        s.set_state(r["h0"], r["u0"])
        s.step(3)
        h, u = s.get_state()
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


def test_method_refusals():
    """nopv with pered or dtred, gass interpolation with diffusion, bad
    method values; fused handles take the triples of trsk or bary
    interpolation, trsk, pered or dtred Coriolis and any kinetic energy,
    gass18's interpolation, Coriolis term or both (trsk's or bary's for
    the other) with any kinetic energy and no diffusion, and the level model with
    trsk10, and refuse the other triples."""
    m = Mesh.icos(3)
    with Swm(m, precision="fp64_ref", dt=900.0) as s:
        for kw in ({"wrap": "none", "cor": "pered", "nopv": True}, {"wrap": "none", "cor": "dtred", "nopv": True},
                   {"wrap": "pxt16", "nopv": True}, {"wrap": "none", "sint": "gass", "k2": 1e5}):
            with pytest.raises(capi.ImxError):
                s.set_scheme(**kw)
        s.set_scheme(wrap="none", cor="gass", nopv=True)
    with Swm(m, precision="fp32", dt=900.0) as s:
        for kw in ({"wrap": "none", "sint": "bary", "cor": "gass", "k2": 1e5},
                   {"wrap": "none", "cor": "gass", "k2": 1e5}, {"wrap": "none", "sint": "gass", "cor": "pered"},
                   {"wrap": "none", "nopv": True, "pv": "apvm"},
                   {"wrap": "gass18", "nopv": True}, {"wrap": "none", "sint": "gass", "cor": "dtred"}):
            with pytest.raises(capi.ImxError):
                s.set_scheme(**kw)
        # the trsk10, pxt16 and gass18 triples spelled out run on the fused
        # kernels, and so does the level model with trsk10
        s.set_scheme(wrap="none", sint="bary", cor="pered", rec="perhx")
        s.set_scheme(wrap="none", sint="gass", cor="gass")
        s.set_scheme(wrap="none", nopv=True)
        s.set_scheme(wrap="none", rec="gass", gasscoef=0.5)
        s.set_scheme(wrap="none", rec="melv")
        s.set_scheme(wrap="none", cor="dtred")
        s.set_scheme(stag="HC", wrap="none", sint="bary", cor="dtred")
        s.set_scheme(wrap="none", sint="bary", cor="pered", rec="melv")
        s.set_scheme(wrap="none", sint="gass", cor="gass", rec="dubos")
        s.set_scheme(stag="HC", wrap="none", sint="gass", cor="gass", rec="melv")
        s.set_scheme(wrap="none", sint="gass", cor="trsk", rec="melv")
        s.set_scheme(stag="HC", wrap="none", sint="trsk", cor="gass", pv="apvm")
        s.set_scheme(wrap="none")


# glevel-5 cases: with K4 > 0 the Laplacian term also runs with sqrt(K4)
# (the reference's set-up), so K4 is kept small enough for dt = 450 s
G5 = {"mix": {"k2": 1e5, "diffus": "diam", "k4": 1e14, "hyperdiffus": "align", "pv": "apvm", "pvspar": 0.5},
      "k2align": {"k2": 1e5, "diffus": "align"},
      "pxt16": {"wrap": "pxt16", "pv": "apvm", "pvspar": 0.5, "k2": 1e5, "diffus": "const"},
      "gass18": {"wrap": "gass18", "pv": "apvm", "pvspar": 0.5},
      "hcpx": {"stag": "HC", "wrap": "pxt16", "pv": "apvm", "pvspar": 0.5, "k4": 1e14, "hyperdiffus": "align"}}


@pytest.mark.parametrize("name", sorted(G5))
def test_scheme_g5_matches_oracle(name):
    """glevel 5 (10242 cells), spectral state, 12 steps with the graph."""
    say("scheme g5 vs oracle", name)
    m = Mesh.icos(5)
    h, u = m.spectral(11)
    sc = dict(G5[name], dt=450.0)
    o = Oracle(m.arrays, scheme=sc)
    ho, uo = o.step(h, u, 450.0, 12)
    with Swm(m, precision="fp64_ref", dt=450.0) as s:
        set_scheme(s, sc)
        s.set_state(h, u)
        s.step(12)
        hg, ug = s.get_state()
    assert np.all(np.isfinite(ho)) and np.all(np.isfinite(uo))
    assert np.array_equal(hg, ho) and np.array_equal(ug, uo)


@pytest.mark.parametrize("name", ["apvm", "mix", "pxt16", "gass18", "gassapvm", "hc", "hcpx", "hcmix"])
def test_scheme_jacobian_equals_per_column(name):
    say("scheme jacobian g2", name)
    m = Mesh.icos(2)
    h, u = m.spectral(5)
    sc = dict(SCHEMES[name], dt=900.0)
    s = Swm(m, precision="fp64_ref", dt=900.0, erad=ERAD_REF)
    set_scheme(s, sc)
    pert = float(np.mean(m.arrays["hx_areag"]))
    J, nev = s.jacobian(h, u)
    M = per_column(Oracle(m.arrays, scheme=sc), h, u, pert)
    assert np.array_equal(J.toarray(), M)
    assert nev < m.nv + m.ne


def test_scheme_errors():
    say("scheme errors")
    m = Mesh.icos(2)
    with Swm(m, precision="fp32", dt=900.0) as s:
        with pytest.raises(capi.ImxError, match="FP64_REF"):
            s.set_scheme(wrap="none", sint="gass", cor="pered")
    with Swm(m, precision="fp64_ref", dt=900.0) as s:
        with pytest.raises(capi.ImxError, match="bad option"):
            s.set_scheme(k2=-1.0)
        with pytest.raises(capi.ImxError, match="gass interpolation with diffusion"):
            s.set_scheme(wrap="gass18", k4=1e15)
        s.set_scheme(k2=1e5)
        h, u = m.spectral(3)
        s.tendency(h, u)
        with pytest.raises(capi.ImxError, match="option is off"):
            s.field("q_grad_ed")
        # back to the defaults: the plain trsk10 tendency
        s.set_scheme()
        f0 = s.tendency(h, u)[1]
    with Swm(m, precision="fp64_ref", dt=900.0) as p:
        assert np.array_equal(p.tendency(h, u)[1], f0)


@pytest.mark.parametrize("sfc", [False, True])
def test_hc_tc2_matches_reference(golden, sfc):
    """TC2 with HC staggering as the reference initialises it: tendency and
    one step bit for bit."""
    say("hc tc2 vs reference", sfc)
    mesh = Mesh(golden("ref_mesh_g3.bin"))
    r = golden("ref_run_g3_tc2hc.bin")
    with Swm(mesh, precision="fp64_ref", dt=900.0, use_sfc=sfc) as s:
        s.set_scheme(stag="HC")
        m, f = s.tendency(r["h0"], r["u0"])
        assert np.array_equal(m, r["masseq"]) and np.array_equal(f, r["momeq"])
        s.set_state(r["h0"], r["u0"])
        s.step(1)
        h, u = s.get_state()
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


# APVM and K2 diffusion on the fused kernels (HTC, trsk10): the tendency
# against the oracle, the option's part on its own (tendency with minus
# without), and steps with the graph
FAST_K2 = {"const": {"k2": 1e5}, "align": {"k2": 1e5, "diffus": "align"},
           "diam": {"k2": 2e5, "diffus": "diam"}, "apvm": {"pv": "apvm", "pvspar": 0.5},
           "apvm_k2": {"pv": "apvm", "pvspar": 0.3, "k2": 1e5, "diffus": "align"},
           "hc_apvm_k2": {"stag": "HC", "pv": "apvm", "pvspar": 0.4, "k2": 1e5, "diffus": "diam"},
           "k4diam": {"k4": 1e14, "hyperdiffus": "diam"},
           "mix": {"k2": 1e5, "diffus": "diam", "k4": 1e14, "hyperdiffus": "align", "pv": "apvm", "pvspar": 0.5},
           "hck4": {"stag": "HC", "pv": "apvm", "pvspar": 0.3, "k4": 1e14, "hyperdiffus": "const"},
           "pxt16": {"wrap": "pxt16"},
           "pxk2": {"wrap": "pxt16", "k2": 1e5, "diffus": "align"},
           "pxmix": {"wrap": "pxt16", "pv": "apvm", "pvspar": 0.5, "k4": 1e14, "hyperdiffus": "const"},
           "gass18": {"wrap": "gass18"}, "hcgass": {"stag": "HC", "wrap": "gass18"},
           "gassapvm": {"wrap": "gass18", "pv": "apvm", "pvspar": 0.5},
           "hcgassapvm": {"stag": "HC", "wrap": "gass18", "pv": "apvm", "pvspar": 0.4},
           "hcpx": {"stag": "HC", "wrap": "pxt16"},
           "hcpxmix": {"stag": "HC", "wrap": "pxt16", "pv": "apvm", "pvspar": 0.5, "k2": 1e5, "diffus": "align"}}


@pytest.mark.parametrize("name", sorted(FAST_K2))
def test_fast_k2_tendency_vs_oracle(name):
    from tests.conftest import rel_l2
    say("fast k2 tendency", name)
    m = Mesh.icos(4)
    h, u = m.spectral(21)
    sc = dict(FAST_K2[name], dt=900.0)
    mo, fo = Oracle(m.arrays, scheme=sc).tendency(h, u)
    _, fo0 = Oracle(m.arrays).tendency(h, u)
    for prec, tol, tol_d in (("fp64", 1e-12, 1e-10), ("fp32", 5e-5, None)):
        with Swm(m, precision=prec) as s:
            _, f0 = s.tendency(h, u)
            set_scheme(s, sc)
            mg, fg = s.tendency(h, u)
        assert rel_l2(mg, mo) < tol, (prec, rel_l2(mg, mo))
        assert rel_l2(fg, fo) < tol, (prec, rel_l2(fg, fo))
        if tol_d is not None:
            # the diffusion term alone: both parts, with their signs
            dg, do = fg - f0, fo - fo0
            assert np.linalg.norm(do) > 1e-6 * np.linalg.norm(fo)
            assert rel_l2(dg, do) < tol_d, (prec, rel_l2(dg, do))


@pytest.mark.parametrize("name", ["align", "apvm_k2", "hc_apvm_k2", "mix", "hck4", "pxt16", "pxmix", "gass18",
                                  "hcgass", "gassapvm", "hcgassapvm", "hcpx", "hcpxmix"])
def test_fast_k2_steps_vs_oracle(name):
    from tests.conftest import rel_l2
    say("fast scheme steps fp64 / fp32", name)
    m = Mesh.icos(4)
    h, u = m.spectral(22)
    sc = dict(FAST_K2[name], dt=900.0)
    ho, uo = Oracle(m.arrays, scheme=sc).step(h, u, 900.0, 10)
    A = m.arrays["hx_areag"]
    for prec, tol_h, tol_u in (("fp64", 1e-12, 1e-11), ("fp32", 1e-7, 1e-5)):
        with Swm(m, precision=prec, dt=900.0) as s:
            set_scheme(s, sc)
            s.set_state(h, u)
            s.step(10)
            hg, ug = s.get_state()
            # back to no diffusion: the plain fused step
            s.set_scheme()
            s.set_state(h, u)
            s.step(10)
            hp, up = s.get_state()
        assert rel_l2(hg, ho, A) < tol_h, (prec, rel_l2(hg, ho, A))
        assert rel_l2(ug, uo) < tol_u, (prec, rel_l2(ug, uo))
        with Swm(m, precision=prec, dt=900.0) as p:
            p.set_state(h, u)
            p.step(10)
            hq, uq = p.get_state()
        assert np.array_equal(hp, hq) and np.array_equal(up, uq)


# options the oracle does not restate (the level model, Gassmann's kinetic
# energy): (scheme, the same scheme without the option)
FAST_VS_REF = {
    "nopv": ({"nopv": True}, {}),
    "hcnopv": ({"stag": "HC", "nopv": True}, {"stag": "HC"}),
    "nopvk2": ({"nopv": True, "k2": 1e5, "diffus": "align"}, {"k2": 1e5, "diffus": "align"}),
    "nopvk4": ({"nopv": True, "k2": 1e5, "diffus": "const", "k4": 1e14, "hyperdiffus": "const"},
               {"k2": 1e5, "diffus": "const", "k4": 1e14, "hyperdiffus": "const"}),
    "hcnopvk4": ({"stag": "HC", "nopv": True, "k4": 1e14, "hyperdiffus": "align"},
                 {"stag": "HC", "k4": 1e14, "hyperdiffus": "align"}),
    "rgass": ({"wrap": "none", "rec": "gass", "gasscoef": 0.75}, {}),
    "rdubos": ({"wrap": "none", "rec": "dubos"}, {}),
    "hcrgass": ({"stag": "HC", "wrap": "none", "rec": "gass", "gasscoef": 0.75}, {"stag": "HC"}),
    "gassgk": ({"wrap": "none", "sint": "gass", "cor": "gass", "rec": "gass", "gasscoef": 0.5}, {"wrap": "gass18"}),
    "rgassk2": ({"wrap": "none", "rec": "gass", "gasscoef": 0.25, "k2": 1e5, "diffus": "const", "pv": "apvm",
                 "pvspar": 0.5}, {"k2": 1e5, "diffus": "const", "pv": "apvm", "pvspar": 0.5}),
    "rmelv": ({"wrap": "none", "rec": "melv"}, {}),
    "rmelvhc": ({"stag": "HC", "wrap": "none", "rec": "melv"}, {"stag": "HC"}),
    "rmelvmix": ({"wrap": "none", "rec": "melv", "k2": 1e5, "diffus": "align", "pv": "apvm", "pvspar": 0.4},
                 {"k2": 1e5, "diffus": "align", "pv": "apvm", "pvspar": 0.4}),
    "cdtred": ({"wrap": "none", "cor": "dtred"}, {}),
    "cdtredhc": ({"stag": "HC", "wrap": "none", "cor": "dtred"}, {"stag": "HC"}),
    "cdtredmix": ({"wrap": "none", "cor": "dtred", "rec": "gass", "gasscoef": 0.5, "k2": 1e5, "diffus": "const"},
                  {"k2": 1e5, "diffus": "const"}),
    "baryperhx": ({"wrap": "none", "sint": "bary", "rec": "perhx"}, {}),
    "barytrsk": ({"wrap": "none", "sint": "bary"}, {}),
    "peredtrsk": ({"wrap": "none", "cor": "pered"}, {}),
    "hcpered": ({"stag": "HC", "wrap": "none", "cor": "pered", "rec": "perhx"}, {"stag": "HC"}),
    "pxgk": ({"wrap": "none", "sint": "bary", "cor": "pered", "rec": "gass", "gasscoef": 0.5}, {}),
    "hcbarymelv": ({"stag": "HC", "wrap": "none", "sint": "bary", "rec": "melv"}, {"stag": "HC"}),
    "dtredperhx": ({"wrap": "none", "cor": "dtred", "rec": "perhx", "k2": 1e5, "diffus": "diam"},
                   {"k2": 1e5, "diffus": "diam"}),
    "barydtred": ({"wrap": "none", "sint": "bary", "cor": "dtred"}, {}),
    "gasstrsk": ({"wrap": "none", "sint": "gass", "cor": "trsk"}, {"wrap": "gass18"}),
    "hcgasstrsk": ({"stag": "HC", "wrap": "none", "sint": "gass", "cor": "trsk", "rec": "melv", "pv": "apvm",
                    "pvspar": 0.4}, {"stag": "HC", "wrap": "gass18", "pv": "apvm", "pvspar": 0.4}),
    "trskgass": ({"wrap": "none", "sint": "trsk", "cor": "gass", "rec": "gass", "gasscoef": 0.5}, {}),
    "trskgassapvm": ({"wrap": "none", "sint": "trsk", "cor": "gass", "pv": "apvm", "pvspar": 0.4},
                     {"pv": "apvm", "pvspar": 0.4}),
    "barygass": ({"wrap": "none", "sint": "bary", "cor": "gass", "pv": "apvm", "pvspar": 0.4},
                 {"pv": "apvm", "pvspar": 0.4}),
    "barygassmelv": ({"wrap": "none", "sint": "bary", "cor": "gass", "rec": "melv"}, {}),
    "hcbarygass": ({"stag": "HC", "wrap": "none", "sint": "bary", "cor": "gass", "rec": "melv", "pv": "apvm",
                    "pvspar": 0.4}, {"stag": "HC", "pv": "apvm", "pvspar": 0.4}),
    "hctrskgass": ({"stag": "HC", "wrap": "none", "sint": "trsk", "cor": "gass", "rec": "perhx"}, {"stag": "HC"}),
    "gassperhx": ({"wrap": "none", "sint": "gass", "cor": "gass", "rec": "perhx"}, {"wrap": "gass18"}),
    "gassmelv": ({"wrap": "none", "sint": "gass", "cor": "gass", "rec": "melv", "pv": "apvm", "pvspar": 0.4},
                 {"wrap": "gass18", "pv": "apvm", "pvspar": 0.4}),
    "hcgassmelv": ({"stag": "HC", "wrap": "none", "sint": "gass", "cor": "gass", "rec": "melv"},
                   {"stag": "HC", "wrap": "gass18"}),
    "hcgassperhx": ({"stag": "HC", "wrap": "none", "sint": "gass", "cor": "gass", "rec": "perhx", "pv": "apvm",
                     "pvspar": 0.4}, {"stag": "HC", "wrap": "gass18", "pv": "apvm", "pvspar": 0.4}),
    "hcbarydtred": ({"stag": "HC", "wrap": "none", "sint": "bary", "cor": "dtred", "rec": "perhx"},
                    {"stag": "HC"}),
    "hcbarydtredk2": ({"stag": "HC", "wrap": "none", "sint": "bary", "cor": "dtred", "k2": 1e5,
                       "diffus": "const"}, {"stag": "HC", "k2": 1e5, "diffus": "const"}),
}


@pytest.mark.parametrize("name", sorted(FAST_VS_REF))
def test_fast_vs_ref(name):
    """The level model (noPV) with trsk10, and method triples other than the
    wrappers' (Gassmann's and Melvin's kinetic energies, the dtred Coriolis
    term, pxt16's methods mixed with trsk10's), HTC and HC, on the
    fused kernels against the FP64_REF handle, which equals the reference's
    own runs with these options bit for bit: the tendency (FP64 < 1e-12,
    FP32 < 5e-5) and 10 steps; the option changes the tendency far above the
    FP64 tolerance; switching back restores the plain fused step bit for
    bit."""
    from tests.conftest import rel_l2
    say("fast vs ref", name)
    m = Mesh.icos(4)
    h, u = m.spectral(24)
    sc, base = FAST_VS_REF[name]
    with Swm(m, precision="fp64_ref", dt=900.0) as r:
        r.set_scheme(**sc)
        mo, fo = r.tendency(h, u)
        r.set_state(h, u)
        r.step(10)
        ho, uo = r.get_state()
        r.set_scheme(**base)
        _, fp = r.tendency(h, u)
    assert rel_l2(fp, fo) > 1e-6, rel_l2(fp, fo)
    A = m.arrays["hx_areag"]
    for prec, tol, tol_h, tol_u in (("fp64", 1e-12, 1e-12, 1e-11), ("fp32", 5e-5, 1e-7, 1e-5)):
        with Swm(m, precision=prec, dt=900.0) as s:
            s.set_scheme(**sc)
            mg, fg = s.tendency(h, u)
            s.set_state(h, u)
            s.step(10)
            hg, ug = s.get_state()
            s.set_scheme()
            s.set_state(h, u)
            s.step(3)
            hp, up = s.get_state()
        assert rel_l2(mg, mo) < tol and rel_l2(fg, fo) < tol, (prec, rel_l2(mg, mo), rel_l2(fg, fo))
        assert rel_l2(hg, ho, A) < tol_h and rel_l2(ug, uo) < tol_u, (prec, rel_l2(hg, ho, A), rel_l2(ug, uo))
        with Swm(m, precision=prec, dt=900.0) as p:
            p.set_state(h, u)
            p.step(3)
            hq, uq = p.get_state()
        assert np.array_equal(hp, hq) and np.array_equal(up, uq)


def test_fast_scheme_refusals():
    say("fast scheme refusals")
    m = Mesh.icos(2)
    with Swm(m, precision="fp64", dt=900.0) as s:
        for kw in ({"wrap": "none", "sint": "gass", "cor": "pered", "rec": "perhx"},
                   {"wrap": "none", "sint": "gass", "cor": "dtred"},
                   {"stag": "HC", "wrap": "none", "cor": "gass", "k4": 1e14},
                   {"wrap": "none", "sint": "gass", "cor": "dtred"}):
            with pytest.raises(capi.ImxError, match="fused handles take"):
                s.set_scheme(**kw)


def test_fast_hc_vs_oracle():
    """HC staggering on the fused kernels (trsk10): the tendency against the
    oracle's HC tendency, 10 steps of TC2 initialised along edhx%nr, the
    diagnostics against an FP64_REF HC handle on the same state, and the
    switch back to HTC restoring the plain step bit for bit."""
    from tests.conftest import rel_l2
    say("fast hc")
    m = Mesh.icos(4)
    h, u = m.spectral(23)
    sc = {"stag": "HC", "dt": 900.0}
    mo, fo = Oracle(m.arrays, scheme=sc).tendency(h, u)
    for prec, tol in (("fp64", 1e-12), ("fp32", 5e-5)):
        with Swm(m, precision=prec) as s:
            s.set_scheme(stag="HC")
            mg, fg = s.tendency(h, u)
        assert rel_l2(mg, mo) < tol, (prec, rel_l2(mg, mo))
        assert rel_l2(fg, fo) < tol, (prec, rel_l2(fg, fo))
    h2, u2 = m.tc2(stag="HC")
    ho, uo = Oracle(m.arrays, scheme=sc).step(h2, u2, 900.0, 10)
    A = m.arrays["hx_areag"]
    with Swm(m, precision="fp64", dt=900.0) as s:
        s.set_scheme(stag="HC")
        s.set_state(h2, u2)
        s.step(10)
        hg, ug = s.get_state()
        d = s.diagnostics()
        s.set_scheme()
        s.set_state(h, u)
        s.step(3)
        hp, up = s.get_state()
    assert rel_l2(hg, ho, A) < 1e-12 and rel_l2(ug, uo) < 1e-11, (rel_l2(hg, ho, A), rel_l2(ug, uo))
    with Swm(m, precision="fp64_ref", dt=900.0) as r:
        r.set_scheme(stag="HC")
        r.set_state(hg, ug)
        dr = r.diagnostics()
    for k in ("mass", "tenergy", "rmsdiv", "penstrophy"):
        assert abs(d[k] - dr[k]) <= 1e-12 * abs(dr[k]), (k, d[k], dr[k])
    with Swm(m, precision="fp64", dt=900.0) as p:
        p.set_state(h2, u2)  # the same h_ref (the area mean of the first state set)
        p.set_state(h, u)
        p.step(3)
        hq, uq = p.get_state()
    assert np.array_equal(hp, hq) and np.array_equal(up, uq)


def test_fast_scheme_jacobian_equals_per_column():
    """The normal-mode Jacobian of a fused FP64 handle with APVM and
    hyperdiffusion: its stencil superset widens with the options, so the
    coloured columns equal the one-column-at-a-time differences of the same
    handle's tendency."""
    say("fast scheme jacobian g2")
    m = Mesh.icos(2)
    h, u = m.spectral(5)
    with Swm(m, precision="fp64", dt=900.0, erad=ERAD_REF) as s:
        s.set_scheme(pv="apvm", pvspar=0.5, k4=1e14, hyperdiffus="const")
        pert = float(np.mean(m.arrays["hx_areag"]))
        J, _ = s.jacobian(h, u, pert)
        x = np.concatenate([h, u])
        m0, f0 = s.tendency(h, u)
        base = np.concatenate([m0, f0])
        for col in (0, 7, m.nv + 3, m.nv + 400):
            xp = x.copy()
            xp[col] += pert
            mp, fp = s.tendency(xp[:m.nv], xp[m.nv:])
            ref = (np.concatenate([mp, fp]) - base) / pert
            assert np.array_equal(J[:, col].toarray().ravel(), ref), col
