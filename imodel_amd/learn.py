"""Learned coarse-grid correction: restriction operators, training-data
generation with this solver, GPU training through libimx's MLP kernels, and
the corrected time step (imx_learn_*).

Pipeline (coarse glevel gc, fine glevel gf = gc + 2, nested icosahedral
grids so every coarse node is a fine node):
  x_n      = R(fine state at t)                        (restriction below)
  x_plain  = RK4 step of the coarse plain model from x_n
  target   = R(fine state at t + dt_c) - x_plain       (du per edge, dh per cell)
The per-edge MLP maps the 48 stencil features of (x_n, x_plain) to the
target; the corrected model adds its output after every plain step.

Restriction. Cells: area-weighted mean of the fine cells whose nodes are
nearest to the coarse node. Edges: a coarse primal edge i->j is the union of
the 2^(gf-gc) fine primal edges on the same great circle; the coarse
tangential (HTC) velocity is their length-weighted mean, oriented i->j.
"""
import ctypes

import numpy as np

from . import capi
from .mlp import Mlp
from .swm import H0_GEO, Mesh, Swm

NF = 48


def restriction(fine, coarse):
    """Sparse (R_h, R_u) mapping fine (h, u) to coarse (h, u), caller order."""
    from scipy.sparse import csr_matrix
    from scipy.spatial import cKDTree

    fa, ca = fine.arrays, coarse.arrays
    pf = fa["v_p"].reshape(-1, 3)
    pc = ca["v_p"].reshape(-1, 3)
    if not np.array_equal(pf[:coarse.nv], pc):
        raise ValueError("grids are not nested (coarse nodes must be the first fine nodes)")
    _, owner = cKDTree(pc).query(pf)
    A = fa["hx_areag"]
    Rh = csr_matrix((A, (owner, np.arange(fine.nv))), shape=(coarse.nv, fine.nv))
    rs = np.asarray(Rh.sum(axis=1)).ravel()
    Rh = csr_matrix(Rh.multiply(1.0 / rs[:, None]))

    nb = fa["v_nb"].reshape(-1, 6) - 1
    ed = fa["v_ed"].reshape(-1, 6) - 1
    nnb = fa["v_nnb"]
    fev = fa["ed_v"].reshape(-1, 2) - 1
    fl = fa["ed_leng"]
    cev = ca["ed_v"].reshape(-1, 2) - 1
    rows, cols, vals = [], [], []
    for ce, (i, j) in enumerate(cev):
        n = np.cross(pc[i], pc[j])
        n /= np.linalg.norm(n)
        cur, path, tot = i, [], 0.0
        while cur != j:
            best, bd, be = -1, -2.0, -1
            for s in range(nnb[cur]):
                v = nb[cur, s]
                if abs(np.dot(n, pf[v])) > 1e-8:
                    continue
                d = np.dot(pf[v], pf[j])
                if d > bd:
                    best, bd, be = v, d, ed[cur, s]
            if best < 0 or len(path) > 64:
                raise RuntimeError(f"no fine path for coarse edge {ce}")
            sign = 1.0 if fev[be, 0] == cur else -1.0
            path.append((be, sign * fl[be]))
            tot += fl[be]
            cur = best
        for fe, w in path:
            rows.append(ce)
            cols.append(fe)
            vals.append(w / tot)
    Ru = csr_matrix((vals, (rows, cols)), shape=(coarse.ne, fine.ne))
    return Rh, Ru


class Learn:
    """imx_learn: the MLP correction bound to a coarse FP32/FP64 handle."""

    def __init__(self, swm, mlp, mesh):
        self.swm, self.mlp, self.mesh = swm, mlp, mesh
        self._h = ctypes.c_void_p()
        capi.check(capi.lib().imx_learn_create(swm._h, mlp.handle, ctypes.byref(mesh.desc()),
                                               ctypes.byref(self._h)), "imx_learn_create")

    def close(self):
        if self._h:
            capi.lib().imx_learn_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sample(self, raw=False):
        """Save the state, take one plain step, build the features."""
        buf = np.zeros((self.mesh.ne, NF), dtype=np.float32) if raw else None
        capi.check(capi.lib().imx_learn_sample(self._h, capi.f32p(buf) if raw else None), "imx_learn_sample")
        return buf

    def set_scaling(self, mu, sig, tsu, tsh):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        sig = np.ascontiguousarray(sig, dtype=np.float32)
        capi.check(capi.lib().imx_learn_set_scaling(self._h, capi.f32p(mu), capi.f32p(sig), float(tsu),
                                                    float(tsh)), "imx_learn_set_scaling")

    def train(self, du, dh, epochs=1, lr=1e-3, wd=0.0):
        du = np.ascontiguousarray(du, dtype=np.float64)
        dh = np.ascontiguousarray(dh, dtype=np.float64)
        loss = ctypes.c_double()
        capi.check(capi.lib().imx_learn_train(self._h, capi.f64p(du), capi.f64p(dh), int(epochs), float(lr),
                                              float(wd), ctypes.byref(loss)), "imx_learn_train")
        return loss.value

    def step(self, nsteps, bound=1e300):
        nfb = np.zeros(1, dtype=np.int32)
        capi.check(capi.lib().imx_learn_step(self._h, int(nsteps), float(bound), capi.i32p(nfb)),
                   "imx_learn_step")
        return int(nfb[0])

    def last_rms(self):
        r = ctypes.c_double()
        capi.check(capi.lib().imx_learn_last_rms(self._h, ctypes.byref(r)), "imx_learn_last_rms")
        return r.value


class Trainer:
    """Generate (x_n, target) pairs with this solver and train the MLP.

    fine: FP64 fused model at glevel gf (the truth); coarse: FP32 fused model
    at glevel gc with the plain step; both share seeded spectral states.
    """

    def __init__(self, gc=6, gf=8, dt_c=450.0, mlp_seed=1, log=print):
        self.gc, self.gf, self.dt_c = gc, gf, dt_c
        self.ratio = 2 ** (gf - gc)
        self.fine_mesh = Mesh.icos(gf)
        self.coarse_mesh = Mesh.icos(gc)
        self.Rh, self.Ru = restriction(self.fine_mesh, self.coarse_mesh)
        self.mlp = Mlp(seed=mlp_seed)
        self.coarse = Swm(self.coarse_mesh, precision="fp32", dt=dt_c, h_ref=H0_GEO)
        self.fine = Swm(self.fine_mesh, precision="fp64", dt=dt_c / self.ratio)
        self.learn = Learn(self.coarse, self.mlp, self.coarse_mesh)
        self.log = log
        self.scaled = False
        self.train_rms_u = None

    def _restrict(self):
        hf, uf = self.fine.get_state()
        return self.Rh @ hf, self.Ru @ uf

    def calibrate(self, seeds, steps=8):
        """Feature mean/std and target scales from a few samples."""
        raws, dus, dhs = [], [], []
        for sd in seeds:
            h, u = self.fine_mesh.spectral(sd)
            self.fine.set_state(h, u)
            for _ in range(steps):
                hc, uc = self._restrict()
                self.coarse.set_state(hc, uc)
                raws.append(self.learn.sample(raw=True))
                self.fine.step(self.ratio)
                hc2, uc2 = self._restrict()
                hp, up = self.coarse.get_state()
                dus.append(uc2 - up)
                dhs.append(hc2 - hp)
        R = np.concatenate(raws)
        mu, sig = R.mean(0), R.std(0)
        sig[sig < 1e-12] = 1.0
        self.tsu = 1.0 / np.std(np.concatenate(dus))
        self.tsh = 1.0 / np.std(np.concatenate(dhs))
        self.learn.set_scaling(mu, sig, self.tsu, self.tsh)
        self.scaled = True
        return mu, sig

    def train(self, seeds, steps_per_run, epochs=2, lr=1e-3, wd=1e-4):
        losses = []
        for r, sd in enumerate(seeds):
            h, u = self.fine_mesh.spectral(sd)
            self.fine.set_state(h, u)
            for k in range(steps_per_run):
                hc, uc = self._restrict()
                self.coarse.set_state(hc, uc)
                self.learn.sample()
                self.fine.step(self.ratio)
                hc2, uc2 = self._restrict()
                hp, up = self.coarse.get_state()
                losses.append(self.learn.train(uc2 - up, hc2 - hp, epochs=epochs, lr=lr, wd=wd))
            self.log(f"run {r} (seed {sd}): loss {np.mean(losses[-steps_per_run:]):.4e}")
        return losses
