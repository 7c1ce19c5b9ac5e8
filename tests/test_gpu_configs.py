"""The benchmark configurations at their own workloads.

* glevel 8, FP32, one simulated day (768 steps of 112.5 s) against the FP64
  reference-order run (bit-identical to the reference) from the same seeded
  state: height L2 <= 1e-5 (the accuracy target), with h and u L2/Linf and
  the mass drift reported.
* 256 independent instances, glevels drawn uniformly from 4..7, each with
  its own seeded state and a CFL-0.8 time step, autotuned, in one set of
  launches per RK stage: one instance per level equals the same instance run
  alone bit for bit, and the smallest instance follows the oracle.
* the learned correction at glevel 6 (dt 450 s) trained on this solver's
  FP64 glevel-8 runs restricted to glevel 6: the corrected step equals the
  oracle chain and conserves mass; its errors against the plain glevel-6
  and glevel-7 runs are logged (the network does not reach glevel-7
  accuracy), with a bound of twice the plain glevel-6 error that a broken
  correction would exceed.
* glevel 10 (10,485,762 cells, dt 28 s): FP32 steps against the FP64
  reference-order path.

Error norms follow error_calc (src/swm.f90:3492-3543): relative L2 with cell
areas (h) or edge areas 1/2 l d (u), relative max norm max|d| / max|ref|.
"""
import time

import numpy as np
import pytest

from imodel_amd.swm import H0_GEO, Ensemble, Mesh, Swm, cfl_dt
from oracle.oracle import Oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def error_norms(mesh, h, u, hR, uR, h0):
    A = mesh.arrays["hx_areag"]
    E = 0.5 * mesh.arrays["ed_leng"] * mesh.arrays["edhx_leng"]
    m0 = np.sum(A * h0)
    return {
        "h_l2": rel_l2(h, hR, A), "h_linf": float(np.max(np.abs(h - hR)) / np.max(np.abs(hR))),
        "u_l2": rel_l2(u, uR, E), "u_linf": float(np.max(np.abs(u - uR)) / np.max(np.abs(uR))),
        "mass_drift": float((np.sum(A * h) - m0) / m0),
        "mass_drift_ref": float((np.sum(A * hR) - m0) / m0),
    }


def test_g8_one_day_fp32_accuracy():
    """glevel 8, 768 FP32 steps vs 768 FP64 reference-order steps."""
    say("g8 fp32 one day vs fp64_ref")
    m = Mesh.icos(8)
    h0, u0 = m.spectral(1)
    out = {}
    for prec in ("fp32", "fp64_ref"):
        with Swm(m, precision=prec, dt=112.5) as s:
            s.set_state(h0, u0)
            t = time.perf_counter()
            s.step(768)
            out[prec] = s.get_state()
            say(prec, f"768 steps {time.perf_counter() - t:.2f} s", s.diagnostics()["badflag"])
    e = error_norms(m, *out["fp32"], *out["fp64_ref"], h0)
    say("one-day errors", e)
    assert e["h_l2"] <= 1e-5, e
    assert e["h_linf"] <= 1e-4 and e["u_l2"] <= 1e-4 and e["u_linf"] <= 1e-3, e
    assert abs(e["mass_drift"]) < 1e-9 and abs(e["mass_drift_ref"]) < 1e-13, e


@pytest.fixture(scope="module")
def ens256():
    rng = np.random.default_rng(1)
    levels = rng.integers(4, 8, size=256)
    meshes = {g: Mesh.icos(int(g)) for g in range(4, 8)}
    ms = [meshes[int(g)] for g in levels]
    states = [ms[i].spectral(i + 1) for i in range(256)]
    dts = [cfl_dt(m, st[1]) for m, st in zip(ms, states)]
    return levels, ms, states, dts


def test_ensemble_256_levels_4_7(ens256):
    levels, ms, states, dts = ens256
    say("ensemble 256 instances", np.bincount(levels, minlength=8)[4:].tolist(), sum(m.nv for m in ms), "cells")
    t = time.perf_counter()
    with Ensemble(ms, dts, precision="fp32", h_ref=H0_GEO) as E:
        say(f"create {time.perf_counter() - t:.1f} s")
        for i, st in enumerate(states):
            E.set_state(i, *st)
        cfg = E.tune(2)
        say("autotuned", cfg)
        for i, st in enumerate(states):
            E.set_state(i, *st)
        E.step(10)
        pick = [int(np.flatnonzero(levels == g)[0]) for g in range(4, 8)]
        got = {i: E.get_state(i) for i in pick}
        diag = [E.diagnostics(i)["badflag"] for i in range(256)]
    assert sum(diag) == 0
    for i in pick:
        with Swm(ms[i], precision="fp32", dt=dts[i], h_ref=H0_GEO) as s:
            s.set_state(*states[i])
            s.step(10)
            h, u = s.get_state()
        say("instance", i, "glevel", levels[i], "dt", dts[i])
        assert np.array_equal(got[i][0], h) and np.array_equal(got[i][1], u), i
    i = pick[0]
    ho, uo = Oracle(ms[i].arrays).step(*states[i], dts[i], 10)
    A = ms[i].arrays["hx_areag"]
    assert rel_l2(got[i][0], ho, A) < 1e-7 and rel_l2(got[i][1], uo) < 1e-5


def test_learned_g6_from_g8():
    """Train on restricted FP64 glevel-8 runs, then: corrected step = oracle
    chain, mass conserved, error within twice plain glevel 6; the ratios
    against plain glevel 6 and 7 logged."""
    from imodel_amd.learn import Trainer, evaluate, restriction
    from oracle import learn_oracle as lo
    from oracle import mlp_oracle as mo

    say("learned g6 from g8")
    t = time.perf_counter()
    tr = Trainer(gc=6, gf=8, dt_c=450.0, log=lambda *a: print(*a, flush=True))
    tr.calibrate([1000, 1001], nsamples=8, spinup=2)
    tr.collect(list(range(1, 25)), 32, spinup=2)
    fit = tr.fit(5000, 1e-3, 5e-5, 1e-4)
    say(f"trained in {time.perf_counter() - t:.1f} s, loss {fit[0].sum():.3f} -> {fit[-1].sum():.3f}")
    assert fit[-1].sum() < fit[0].sum()

    # one corrected step against the oracle chain with the trained network
    g6 = tr.coarse_mesh
    hs, us = g6.spectral(77)
    tr.coarse.set_state(hs, us)
    tr.coarse.step(2)
    hn, un = tr.coarse.get_state()
    raw = tr.learn.sample(raw=True)
    hp, up = tr.coarse.get_state()
    tr.coarse.set_state(hn, un)
    assert tr.learn.step(1, bound=1e30) == 0
    hc, uc = tr.coarse.get_state()
    mu, sig = tr.learn_scaling
    X = (raw.astype(np.float64) - mu) / sig
    Y, _ = mo.forward(tr.mlp.params().astype(np.float64), X, emulate_bf16=True)
    ho, uo, _ = lo.correct(hp, up, Y, lo.div_matrix(g6.arrays), tr.tsu, tr.tsh, 1e30)
    assert np.linalg.norm((uc - up) - (uo - up)) < 2e-2 * np.linalg.norm(uo - up)
    assert np.linalg.norm((hc - hp) - (ho - hp)) < 2e-2 * np.linalg.norm(ho - hp)

    m7 = Mesh.icos(7)
    mid = (Swm(m7, precision="fp32", dt=225.0, h_ref=H0_GEO),) + restriction(tr.fine_mesh, m7) + \
        restriction(m7, g6)
    r = evaluate(tr, 9999, 96, spinup=2, mid=mid)
    say("learned vs plain g6 vs plain g7 (half a day):", r)
    say("error ratio learned/g7: h %.3f u %.3f; learned/g6: h %.3f u %.3f; wall g7/learned %.2f" % (
        r["learned"]["h"] / r["mid"]["h"], r["learned"]["u"] / r["mid"]["u"],
        r["learned"]["h"] / r["plain"]["h"], r["learned"]["u"] / r["plain"]["u"],
        r["mid"]["wall_s"] / r["learned"]["wall_s"]))
    assert abs(r["learned"]["mass_drift"]) < 1e-8
    assert r["learned"]["h"] < 2 * r["plain"]["h"] and r["learned"]["u"] < 2 * r["plain"]["u"]


def test_g10_fp32_vs_fp64_ref():
    say("g10 fp32 vs fp64_ref")
    t = time.perf_counter()
    m = Mesh.icos(10)
    h0, u0 = m.spectral(1)
    say(f"mesh + state {time.perf_counter() - t:.1f} s")
    out = {}
    for prec in ("fp32", "fp64_ref"):
        t = time.perf_counter()
        with Swm(m, precision=prec, dt=28.0) as s:
            say(prec, f"create {time.perf_counter() - t:.1f} s")
            s.set_state(h0, u0)
            s.step(1)
            s.set_state(h0, u0)
            ms = s.step_timed(4)
            out[prec] = s.get_state()
            say(prec, f"4 steps {ms:.2f} ms = {m.nv * 4 / (ms * 1e-3):.3e} cell-steps/s")
    e = error_norms(m, *out["fp32"], *out["fp64_ref"], h0)
    say("g10 errors after 4 steps", e)
    assert e["h_l2"] < 1e-7 and e["u_l2"] < 1e-5, e
