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
