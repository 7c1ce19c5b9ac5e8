"""Ensembles (many instances in one set of launches) and the autotuner."""
import numpy as np
import pytest

from imodel_amd.swm import Ensemble, Mesh, Swm, cfl_dt
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_ensemble_matches_single_instances(prec, tol):
    """Every instance of an ensemble (different levels, seeds, time steps)
    evolves as the same instance run alone."""
    say("ensemble vs single", prec)
    meshes = [Mesh.icos(3), Mesh.icos(4), Mesh.icos(3)]
    states = [meshes[0].spectral(1), meshes[1].spectral(2), meshes[2].spectral(3)]
    dts = [cfl_dt(m, s[1]) for m, s in zip(meshes, states)]
    href = 2.94e4 / 9.80616
    with Ensemble(meshes, dts, precision=prec, h_ref=href) as E:
        for i, (h, u) in enumerate(states):
            E.set_state(i, h, u)
        E.step(6)
        ens = [E.get_state(i) for i in range(3)]
        dens = [E.diagnostics(i) for i in range(3)]
    for i, m in enumerate(meshes):
        with Swm(m, precision=prec, dt=dts[i], h_ref=href) as s:
            s.set_state(*states[i])
            s.step(6)
            h, u = s.get_state()
            d = s.diagnostics()
        A = m.arrays["hx_areag"]
        assert rel_l2(ens[i][0] - href, h - href, A) < tol
        assert rel_l2(ens[i][1], u) < tol
        assert abs(dens[i]["tenergy"] / d["tenergy"] - 1) < tol


def test_tune_keeps_results():
    say("tune")
    m = Mesh.icos(4)
    h, u = m.spectral(4)
    out = []
    for tune in (False, True):
        with Swm(m, precision="fp32", dt=600.0) as s:
            s.set_state(h, u)
            if tune:
                cfg = s.tune(2)
                assert set(cfg) == {"tri", "cell", "edge"}
                assert all(b in (64, 128, 256, 512) and r in (0, 1) for b, r in cfg.values())
            s.step(5)
            out.append(s.get_state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
