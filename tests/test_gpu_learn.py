"""Learned coarse-grid correction on the GPU (imx_learn_*): features and the
corrected step against the numpy restatement, the RMS limiter, mass
conservation of the flux-form h correction, and what a network trained
with this solver's own finer run gains over the plain coarse step."""
import numpy as np
import pytest

from imodel_amd.learn import Learn, Trainer, evaluate, restriction
from imodel_amd.mlp import Mlp
from imodel_amd.swm import H0_GEO, Mesh, Swm
from oracle import learn_oracle as lo
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def g3():
    return Mesh.icos(3)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_features_match_oracle(g3, prec):
    say("learn features", prec)
    s = Swm(g3, precision=prec, dt=3600.0, h_ref=H0_GEO)
    L = Learn(s, Mlp(seed=1), g3)
    s.set_state(*g3.spectral(4))
    hn, un = s.get_state()          # the state as stored (FP32-rounded for fp32)
    raw = L.sample(raw=True)
    hp, up = s.get_state()
    ref = lo.features(lo.stencils(g3.arrays), hn, un, hp, up, H0_GEO)
    scale = np.abs(ref).max(axis=0) + 1e-30
    tol = 2e-6 if prec == "fp32" else 1e-6
    assert (np.abs(raw - ref) / scale).max() < tol


def test_corrected_step_matches_oracle(g3):
    """Plain step + MLP correction (BF16 network, flux-form h) vs the oracle
    chain: features -> normalise -> network (BF16 emulated) -> D -> update."""
    say("learn corrected step")
    m = Mlp(seed=9)
    p = m.params().astype(np.float64)
    s = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    L = Learn(s, m, g3)
    h0, u0 = g3.spectral(6)
    s.set_state(h0, u0)
    hn, un = s.get_state()
    raw = L.sample(raw=True)
    hp, up = s.get_state()
    mu, sig = raw.mean(0), raw.std(0)
    sig[sig < 1e-6] = 1.0
    tsu, tsh = 50.0, 0.5
    L.set_scaling(mu, sig, tsu, tsh)
    s.set_state(hn, un)
    nfb = L.step(1, bound=1e30)
    assert nfb == 0
    hc, uc = s.get_state()
    X = (raw.astype(np.float64) - mu) / sig
    Y, _ = mo.forward(p, X, emulate_bf16=True)
    D = lo.div_matrix(g3.arrays)
    ho, uo, _ = lo.correct(hp, up, Y, D, tsu, tsh, 1e30)
    du, dh = uo - up, ho - hp
    assert np.linalg.norm((uc - up) - du) < 2e-2 * np.linalg.norm(du)
    assert np.linalg.norm((hc - hp) - dh) < 2e-2 * np.linalg.norm(dh)


def test_limiter_bound_zero_is_plain(g3):
    say("learn limiter bound 0")
    s = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    p = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    L = Learn(s, Mlp(seed=2), g3)
    h0, u0 = g3.spectral(7)
    s.set_state(h0, u0)
    p.set_state(h0, u0)
    assert L.step(5, bound=0.0) == 5
    p.step(5)
    hs, us = s.get_state()
    hp, up = p.get_state()
    assert np.array_equal(hs, hp) and np.array_equal(us, up)
    assert L.last_rms() > 0.0


def test_learned_step_follows_handle_changes(g3):
    """The learned step's graph holds the coarse handle's step kernels; it
    is captured again after the handle's scheme or launch shapes change.
    With bound 0 the learned step is the handle's plain step, so after
    set_scheme(K2) it equals the K2 step of a handle that had K2 from the
    start, bit for bit, and after tune() still does."""
    say("learn graph follows the handle")
    h0, u0 = g3.spectral(8)
    s = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    L = Learn(s, Mlp(seed=3), g3)
    s.set_state(h0, u0)
    L.step(1, bound=0.0)  # captured with the plain scheme
    hq, uq = s.get_state()
    s.set_scheme(k2=1e5, diffus="const")
    s.set_state(h0, u0)
    L.step(1, bound=0.0)
    h1, u1 = s.get_state()
    p = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    p.set_scheme(k2=1e5, diffus="const")
    p.set_state(h0, u0)
    p.step(1)
    hp, up = p.get_state()
    assert not np.array_equal(up, uq)
    assert np.array_equal(h1, hp) and np.array_equal(u1, up)
    s.tune(1)
    s.set_state(h0, u0)
    L.step(1, bound=0.0)
    h2, u2 = s.get_state()
    assert np.array_equal(h2, hp) and np.array_equal(u2, up)


def test_limiter_falls_back_to_plain_step(g3):
    """A correction above the bound (the default limiter) is dropped: the
    corrected step equals the plain step bit for bit and counts as limited;
    below the bound the correction is applied."""
    say("learn limiter fallback")
    s = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    p = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    L = Learn(s, Mlp(seed=2), g3)
    h0, u0 = g3.spectral(7)
    s.set_state(h0, u0)
    raw = L.sample(raw=True)
    L.set_scaling(raw.mean(0), raw.std(0) + 1e-3, 50.0, 0.5)
    s.set_state(h0, u0)
    assert L.step(1, bound=1e30) == 0
    r = L.last_rms()
    hc, uc = s.get_state()
    s.set_state(h0, u0)
    assert L.step(1, bound=0.99 * r) == 1
    hs, us = s.get_state()
    p.set_state(h0, u0)
    p.step(1)
    hp, up = p.get_state()
    assert np.array_equal(hs, hp) and np.array_equal(us, up)
    assert not np.array_equal(uc, up)


def test_limiter_scales_to_bound(g3):
    """With the scaling limiter, a bound of a quarter of the correction's
    RMS applies a quarter of it."""
    say("learn limiter scale")
    s = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    L = Learn(s, Mlp(seed=2), g3)
    L.set_limiter("scale")
    h0, u0 = g3.spectral(7)
    s.set_state(h0, u0)
    raw = L.sample(raw=True)
    hp, up = s.get_state()
    L.set_scaling(raw.mean(0), raw.std(0) + 1e-3, 50.0, 0.5)
    s.set_state(h0, u0)
    assert L.step(1, bound=1e30) == 0
    r = L.last_rms()
    h1, u1 = s.get_state()
    s.set_state(h0, u0)
    assert L.step(1, bound=0.25 * r) == 1
    h2, u2 = s.get_state()
    du1, du2 = u1.astype(np.float64) - up, u2.astype(np.float64) - up
    dh1, dh2 = h1.astype(np.float64) - hp, h2.astype(np.float64) - hp
    # FP32 state: the increments carry the rounding of u + du
    assert np.linalg.norm(du2 - 0.25 * du1) < 1e-2 * np.linalg.norm(du1)
    assert np.linalg.norm(dh2 - 0.25 * dh1) < 1e-2 * np.linalg.norm(dh1)


def test_correction_conserves_mass(g3):
    """Large corrections (untrained network, small target scales) change h by
    metres per step; total mass moves only by FP32 rounding."""
    say("learn mass")
    s = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    L = Learn(s, Mlp(seed=3), g3)
    h0, u0 = g3.spectral(8)
    s.set_state(h0, u0)
    raw = L.sample(raw=True)
    L.set_scaling(raw.mean(0), raw.std(0) + 1e-3, 1.0, 0.2)
    s.set_state(h0, u0)
    m0 = s.diagnostics()["mass"]
    assert L.step(6, bound=1e30) == 0
    h1, _ = s.get_state()
    m1 = s.diagnostics()["mass"]
    p = Swm(g3, precision="fp32", dt=3600.0, h_ref=H0_GEO)
    p.set_state(h0, u0)
    p.step(6)
    hp, _ = p.get_state()
    assert np.abs(h1 - hp).max() > 1.0                 # the correction is large
    assert abs(m1 - m0) / m0 < 1e-9, (m0, m1)


def test_learned_beats_plain():
    """Train on a glevel-6 FP64 run restricted to glevel 4 (1280 samples),
    then run 48 steps (1 day) from a held-out seed: the corrected glevel-4
    model tracks the restricted truth better than the plain one."""
    say("learn train g4 from g6")
    tr = Trainer(gc=4, gf=6, dt_c=1800.0, log=lambda *a: print(*a, flush=True))
    tr.calibrate([1000, 1001], nsamples=8, spinup=4)
    tr.collect(list(range(1, 41)), 32, spinup=4)
    fit = tr.fit(4000, 1e-3, 5e-5, 1e-4)
    assert fit[-1].sum() < 0.7 * fit[0].sum() or fit[-1].sum() < 1.0
    r = evaluate(tr, 9999, 48, spinup=4)
    say(r)
    assert r["learned"]["h"] < 0.7 * r["plain"]["h"]
    assert r["learned"]["u"] < 0.8 * r["plain"]["u"]
    assert 0 <= r["learned"]["limited"] <= 48
    assert abs(r["learned"]["mass_drift"]) < 1e-8
