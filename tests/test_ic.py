"""Seeded spectral initial state (imx_ic_spectral; not in the reference).

Pinned independently of libimx: the test regenerates the state with its own
SplitMix64 stream and scipy's spherical harmonics (a different Legendre
algorithm), and checks the closed-form properties (div u = 0, amplitudes).
"""
import numpy as np
import pytest

from imodel_amd.swm import H0_GEO, Mesh

try:
    from scipy.special import sph_harm_y as _sh

    def ylm(l, m, colat, lon):
        return _sh(l, m, colat, lon)
except ImportError:  # older scipy
    from scipy.special import sph_harm as _sh2

    def ylm(l, m, colat, lon):
        return _sh2(m, l, lon, colat)

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self):
        return ((self.next() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self):
        u1, u2 = self.uniform(), self.uniform()
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def _coeffs(rng, L):
    c = []
    for l in range(1, L + 1):
        for m in range(l + 1):
            amp = 1.0 / (l * l)
            a = amp * rng.normal()
            b = amp * rng.normal() if m > 0 else 0.0
            c.append((l, m, a, b))
    return c


def _field(coeffs, p):
    colat = np.arccos(np.clip(p[:, 2], -1, 1))
    lon = np.arctan2(p[:, 1], p[:, 0])
    f = np.zeros(len(p))
    for l, m, a, b in coeffs:
        y = ylm(l, m, colat, lon)
        if m == 0:
            f += a * y.real
        else:
            f += np.sqrt(2.0) * (a * y.real + b * y.imag)
    return f


def test_spectral_state_matches_independent_construction():
    m = Mesh.icos(3)
    a = m.arrays
    h, u = m.spectral(7)
    rng = SplitMix64(7)
    ch = _coeffs(rng, 16)
    cp = _coeffs(rng, 16)
    A = a["hx_areag"]
    hp = _field(ch, a["v_p"].reshape(-1, 3))
    hp *= 0.05 * H0_GEO / np.sqrt(np.sum(A * hp ** 2) / np.sum(A))
    assert np.max(np.abs((h - H0_GEO) - hp)) < 1e-9 * np.max(np.abs(hp))
    psi = _field(cp, a["tr_c"].reshape(-1, 3))
    ev = a["ed_v"].reshape(-1, 2) - 1
    es = a["ed_sh"].reshape(-1, 2) - 1
    vp = a["v_p"].reshape(-1, 3)
    tc = a["tr_c"].reshape(-1, 3)
    n = np.cross(vp[ev[:, 0]], vp[ev[:, 1]])
    s = np.where(np.sum(n * (tc[es[:, 1]] - tc[es[:, 0]]), axis=1) >= 0, 1.0, -1.0)
    uu = -s * (psi[es[:, 1]] - psi[es[:, 0]]) / (a["edhx_leng"] * 6.37122e6)
    ratio = u / np.where(uu == 0, 1, uu)
    k = np.median(ratio[np.abs(uu) > 1e-3 * np.abs(uu).max()])
    assert np.max(np.abs(u - k * uu)) < 1e-9 * np.max(np.abs(u))


@pytest.mark.parametrize("seed", [1, 2])
def test_amplitudes_and_determinism(seed):
    m = Mesh.icos(4)
    a = m.arrays
    h, u = m.spectral(seed)
    h2, u2 = m.spectral(seed)
    assert np.array_equal(h, h2) and np.array_equal(u, u2)
    h3, _ = m.spectral(seed + 100)
    assert not np.array_equal(h, h3)
    A = a["hx_areag"]
    rms = np.sqrt(np.sum(A * (h - H0_GEO) ** 2) / np.sum(A))
    assert abs(rms / (0.05 * H0_GEO) - 1) < 1e-12
    # sqrt(2 <K>) = 20 m/s with K as kinetic_energy_hx (trsk)
    ed = a["v_ed"].reshape(-1, 6) - 1
    nnb = a["v_nnb"]
    w = 0.25 * a["ed_leng"] * a["edhx_leng"]
    K = np.zeros(m.nv)
    for i in range(m.nv):
        e = ed[i, :nnb[i]]
        K[i] = np.sum(w[e] * u[e] ** 2) / np.sum(w[e])
    assert abs(np.sqrt(2 * np.sum(A * K) / np.sum(A)) / 20.0 - 1) < 1e-12
