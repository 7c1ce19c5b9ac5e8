"""The C oracle against golden vectors written by the reference itself.

oracle/swm_oracle.c restates tendency / ode_rk4 / the diagnostics of
src/swm.f90 and src/swm_operators.f90 loop for loop; with FMA contraction off
on both sides it must reproduce the reference bit for bit.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle

FIELDS = ["h_ed", "uh", "eta", "h_tr", "q_tr", "q_ed", "divuh", "ke_hx", "ghbK", "grad_ghbK",
          "uhq_perp", "divu", "zeta"]


@pytest.mark.parametrize("run", ["ref_run_g3_tc2.bin", "ref_run_g3_spec.bin"])
def test_tendency_bitexact(golden, run):
    mesh = golden("ref_mesh_g3.bin")
    r = golden(run)
    o = Oracle(mesh)
    m, f = o.tendency(r["h0"], r["u0"])
    assert np.array_equal(m, r["masseq"])
    assert np.array_equal(f, r["momeq"])
    for name in FIELDS:
        assert np.array_equal(o.field(name.lower()), r[name]), name


@pytest.mark.parametrize("run", ["ref_run_g3_tc2.bin", "ref_run_g3_spec.bin"])
def test_rk4_and_diagnostics_bitexact(golden, run):
    mesh = golden("ref_mesh_g3.bin")
    r = golden(run)
    o = Oracle(mesh)
    d0 = o.diagnostics(r["h0"], r["u0"], refresh=True)
    # mass, P, K, T, A, RMSdiv of swm.f90:224-227 and :352
    assert np.array_equal(d0[[0, 1, 2, 3, 4, 6]], r["diag0"])
    ser = r["series"].reshape(-1, 7)
    h, u = r["h0"], r["u0"]
    for k in range(ser.shape[0]):
        h, u = o.step(h, u, 900.0, 1)
        # the reference loop reads the edge fields of its last RK stage
        d = o.diagnostics(h, u, refresh=False)
        assert np.array_equal(d[[0, 1, 2, 3, 4, 6]], ser[k, 1:]), k
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


def test_96_steps_bitexact(golden):
    mesh = golden("ref_mesh_g2.bin")
    r = golden("ref_run_g2_spec96.bin")
    o = Oracle(mesh)
    h, u = o.step(r["h0"], r["u0"], 900.0, 96)
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


def test_oracle_detects_planted_mistake(golden):
    """A dropped term or a sign slip in the oracle must show up against the
    golden vectors: flip the sign of one TRiSK weight."""
    mesh = dict(golden("ref_mesh_g3.bin"))
    r = golden("ref_run_g3_tc2.bin")
    w = mesh["hx_trskw"].copy()
    w[1] = -w[1]  # cell 1, weight (1,2)
    mesh["hx_trskw"] = w
    o = Oracle(mesh)
    _, f = o.tendency(r["h0"], r["u0"])
    assert not np.array_equal(f, r["momeq"])


# scheme options of the scheme fixtures (tests/golden/make_golden.py SCHEMES)
SCHEMES = {
    "apvm": {"pv": "apvm", "pvspar": 0.5},
    "k2align": {"k2": 1e5, "diffus": "align"},
    "k4diam": {"k4": 1e15, "hyperdiffus": "diam"},
    "mix": {"k2": 2e5, "diffus": "const", "k4": 3e15, "hyperdiffus": "align", "pv": "apvm",
            "pvspar": 0.3},
    "pxt16": {"wrap": "pxt16"},
    "gass18": {"wrap": "gass18"},
    "pxmix": {"wrap": "pxt16", "pv": "apvm", "pvspar": 0.4, "k4": 1e15, "hyperdiffus": "const"},
    "gassapvm": {"wrap": "gass18", "pv": "apvm", "pvspar": 0.5},
    "hc": {"stag": "HC"},
    "hcpx": {"stag": "HC", "wrap": "pxt16", "k2": 1e5, "diffus": "align"},
    "hcmix": {"stag": "HC", "wrap": "gass18", "pv": "apvm", "pvspar": 0.5},
    "hck4": {"stag": "HC", "pv": "apvm", "pvspar": 0.3, "k4": 1e15, "hyperdiffus": "diam"},
}
SCHEME_FIELDS = ["v_hx", "q_grad_ed", "q_grad_tr", "lapu", "lap_lapu"]


@pytest.mark.parametrize("name", sorted(SCHEMES))
def test_scheme_options_bitexact(golden, name):
    """APVM (grad_tr2ed, Perot reconstructions, apvm), diffusion and
    hyperdiffusion with their coefficient set-up: tendency, every
    intermediate the reference exposes, and 3 RK4 steps, bit for bit."""
    mesh = golden("ref_mesh_g3.bin")
    r = golden(f"ref_run_g3_{name}.bin")
    o = Oracle(mesh, perot=True, scheme=dict(SCHEMES[name], dt=900.0))
    if "dif_coef_hx" in r:
        assert np.array_equal(o.dif_hx, r["dif_coef_hx"]) and np.array_equal(o.dif_tr, r["dif_coef_tr"])
    m, f = o.tendency(r["h0"], r["u0"])
    assert np.array_equal(m, r["masseq"]) and np.array_equal(f, r["momeq"])
    for k in SCHEME_FIELDS + ["h_tr", "q_tr", "q_ed", "ke_hx", "uhq_perp"]:
        if k in r:
            assert np.array_equal(o.field(k), r[k]), k
    h, u = o.step(r["h0"], r["u0"], 900.0, 3)
    assert np.array_equal(h, r["h"]) and np.array_equal(u, r["u"])


def test_perot_reconstruction_bitexact(golden):
    """vector_edhx2hx_perot, computed by the reference every step (dead for
    plain trsk10), against its v_hx."""
    mesh = golden("ref_mesh_g3.bin")
    r = golden("ref_run_g3_spec.bin")
    o = Oracle(mesh, perot=True)
    o.tendency(r["h0"], r["u0"])
    assert np.array_equal(o.field("v_hx"), r["v_hx"])


def test_wrapper_planted_mistakes_fail(golden):
    """Each wrapper's fixture rejects the oracle run with another wrapper."""
    mesh = golden("ref_mesh_g3.bin")
    for name, other in (("pxt16", "trsk10"), ("gass18", "trsk10"), ("pxt16", "gass18")):
        r = golden(f"ref_run_g3_{name}.bin")
        o = Oracle(mesh, scheme={"wrap": other, "dt": 900.0})
        _, f = o.tendency(r["h0"], r["u0"])
        assert not np.array_equal(f, r["momeq"]), (name, other)
    # and the HC fixture rejects HTC
    for name in ("hc", "hcpx"):
        r = golden(f"ref_run_g3_{name}.bin")
        o = Oracle(mesh, scheme=dict(SCHEMES[name], stag="HTC", dt=900.0))
        _, f = o.tendency(r["h0"], r["u0"])
        assert not np.array_equal(f, r["momeq"]), (name, other)


def test_gass18_with_diffusion_refused(golden):
    """The reference stops there (scalar_hx2tr has no Gassmann branch for
    the coefficients); so does the oracle."""
    with pytest.raises(ValueError, match="gass18"):
        Oracle(golden("ref_mesh_g3.bin"), scheme={"wrap": "gass18", "k2": 1e5})


def test_scheme_planted_mistakes_fail(golden):
    """The scheme fixtures catch a wrong APVM parameter or a wrong diffusion
    coefficient in the oracle's inputs. (With K4 > 0 the reference runs both
    Laplacian terms with the square roots of the hyperdiffusion coefficients,
    src/swm.f90:206-208, so in "mix" the K2 kind and value do not enter.)"""
    mesh = golden("ref_mesh_g3.bin")
    r = golden("ref_run_g3_mix.bin")
    for bad in ({"pvspar": 0.31}, {"k4": 3.1e15}, {"hyperdiffus": "const"}, {"k2": 0.0}):
        o = Oracle(mesh, scheme=dict(SCHEMES["mix"], dt=900.0, **bad))
        _, f = o.tendency(r["h0"], r["u0"])
        assert not np.array_equal(f, r["momeq"]), bad
