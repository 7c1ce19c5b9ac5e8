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

Restriction (one bisection level, composed over levels). Cells: area-weighted
mean of the fine cell at the coarse node and half of each fine cell at the
bisection points of its edges (those are equidistant from two coarse nodes).
Edges: a coarse primal edge i->j is the union of two fine primal edges on
the same great circle; the coarse tangential (HTC) velocity is their
length-weighted mean, oriented i->j.
"""
import ctypes
import time

import numpy as np

from . import capi
from .mlp import Mlp
from .swm import H0_GEO, Mesh, Swm

NF = 48


def _restriction_step(fine, coarse):
    """One bisection level: fine = refinement of coarse (glevel + 1)."""
    from scipy.sparse import csr_matrix
    from scipy.spatial import cKDTree

    fa, ca = fine.arrays, coarse.arrays
    pf = fa["v_p"].reshape(-1, 3)
    pc = ca["v_p"].reshape(-1, 3)
    if not np.array_equal(pf[:coarse.nv], pc):
        raise ValueError("grids are not nested (coarse nodes must be the first fine nodes)")
    # every fine node is a coarse node or the bisection point m of one
    # coarse edge i->j (equidistant from i and j)
    fev = fa["ed_v"].reshape(-1, 2) - 1
    cev = ca["ed_v"].reshape(-1, 2) - 1
    mid = pc[cev[:, 0]] + pc[cev[:, 1]]
    mid /= np.linalg.norm(mid, axis=1)[:, None]
    dist, m = cKDTree(pf).query(mid)
    if dist.max() > 1e-9 or fine.nv != coarse.nv + coarse.ne or \
            not np.array_equal(np.sort(m), np.arange(coarse.nv, fine.nv)):
        raise ValueError("fine grid is not the bisection of the coarse grid")
    # cells: area-weighted mean of the fine cell at the node and half of each
    # fine cell at the bisection points of its edges
    A = fa["hx_areag"]
    rows = np.concatenate([np.arange(coarse.nv), cev[:, 0], cev[:, 1]])
    cols = np.concatenate([np.arange(coarse.nv), m, m])
    vals = np.concatenate([A[:coarse.nv], 0.5 * A[m], 0.5 * A[m]])
    Rh = csr_matrix((vals, (rows, cols)), shape=(coarse.nv, fine.nv))
    rs = np.asarray(Rh.sum(axis=1)).ravel()
    Rh = csr_matrix(Rh.multiply(1.0 / rs[:, None]))
    # edges: coarse edge i->j = fine edges i->m, m->j
    nf = np.int64(fine.nv)
    key = np.minimum(fev[:, 0], fev[:, 1]) * nf + np.maximum(fev[:, 0], fev[:, 1])
    order = np.argsort(key)
    skey = key[order]

    def find(a, b):
        k = np.minimum(a, b) * nf + np.maximum(a, b)
        pos = np.searchsorted(skey, k)
        if np.any(pos >= len(skey)) or np.any(skey[np.minimum(pos, len(skey) - 1)] != k):
            raise ValueError("missing fine edge on a coarse edge")
        return order[pos]

    e1 = find(cev[:, 0], m)
    e2 = find(m, cev[:, 1])
    fl = fa["ed_leng"]
    s1 = np.where(fev[e1, 0] == cev[:, 0], 1.0, -1.0)
    s2 = np.where(fev[e2, 0] == m, 1.0, -1.0)
    tot = fl[e1] + fl[e2]
    nc = coarse.ne
    rows = np.concatenate([np.arange(nc), np.arange(nc)])
    cols = np.concatenate([e1, e2])
    vals = np.concatenate([s1 * fl[e1] / tot, s2 * fl[e2] / tot])
    Ru = csr_matrix((vals, (rows, cols)), shape=(nc, fine.ne))
    return Rh, Ru


def restriction(fine, coarse, meshes=None):
    """Sparse (R_h, R_u) mapping fine (h, u) to coarse (h, u), caller order.

    Built level by level (each level: area-weighted mean of the fine cell at
    the coarse node and half of the fine cells at the bisection points of its
    edges; length-weighted, orientation-signed mean of the two fine edges
    that bisect each coarse edge) and composed, so
    restricting g -> g-1 -> g-2 equals restricting g -> g-2 exactly.
    `meshes` may supply the intermediate levels {glevel: Mesh}.
    """
    gf, gc = fine.glevel, coarse.glevel
    if gf <= gc:
        raise ValueError("fine glevel must exceed coarse glevel")
    meshes = dict(meshes or {})
    meshes[gf], meshes[gc] = fine, coarse
    Rh = Ru = None
    for g in range(gf, gc, -1):
        if g - 1 not in meshes:
            meshes[g - 1] = Mesh.icos(g - 1)
        h, u = _restriction_step(meshes[g], meshes[g - 1])
        Rh = h if Rh is None else (h @ Rh).tocsr()
        Ru = u if Ru is None else (u @ Ru).tocsr()
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
        """Fit the correction of the last sample; returns the scaled (u, h) losses."""
        du = np.ascontiguousarray(du, dtype=np.float64)
        dh = np.ascontiguousarray(dh, dtype=np.float64)
        loss = np.zeros(2)
        capi.check(capi.lib().imx_learn_train(self._h, capi.f64p(du), capi.f64p(dh), int(epochs), float(lr),
                                              float(wd), capi.f64p(loss)), "imx_learn_train")
        return float(loss[0]), float(loss[1])

    def reserve(self, cap):
        capi.check(capi.lib().imx_learn_reserve(self._h, int(cap)), "imx_learn_reserve")

    def store(self, du, dh):
        du = np.ascontiguousarray(du, dtype=np.float64)
        dh = np.ascontiguousarray(dh, dtype=np.float64)
        slot = np.zeros(1, dtype=np.int32)
        capi.check(capi.lib().imx_learn_store(self._h, capi.f64p(du), capi.f64p(dh), capi.i32p(slot)),
                   "imx_learn_store")
        return int(slot[0])

    def fit(self, nsteps, lr0=1e-3, lr1=None, wd=0.0, seed=1):
        """nsteps AdamW steps on stored samples; returns the scaled (u, h) losses."""
        loss = np.zeros(2)
        lr1 = lr0 if lr1 is None else lr1
        capi.check(capi.lib().imx_learn_fit(self._h, int(nsteps), float(lr0), float(lr1), float(wd), int(seed),
                                            capi.f64p(loss)), "imx_learn_fit")
        return float(loss[0]), float(loss[1])

    def set_limiter(self, mode):
        """What a correction above the RMS bound does: "fallback" (the
        default) keeps the plain step, "scale" scales it down to the bound."""
        capi.check(capi.lib().imx_learn_set_limiter(self._h, {"fallback": 0, "scale": 1}[mode]),
                   "imx_learn_set_limiter")

    def step(self, nsteps, bound=1e300):
        """nsteps corrected steps; a correction above RMS `bound` (scaled
        units) is dropped for the plain step (or scaled down to the bound,
        set_limiter("scale")). Returns the number of limited steps."""
        nlim = np.zeros(1, dtype=np.int32)
        capi.check(capi.lib().imx_learn_step(self._h, int(nsteps), float(bound), capi.i32p(nlim)),
                   "imx_learn_step")
        return int(nlim[0])

    def last_rms(self):
        r = ctypes.c_double()
        capi.check(capi.lib().imx_learn_last_rms(self._h, ctypes.byref(r)), "imx_learn_last_rms")
        return r.value


class Trainer:
    """Generate (x_n, target) pairs with this solver and train the MLP.

    fine: FP64 fused model at glevel gf (the truth); coarse: FP32 fused model
    at glevel gc whose plain step is corrected. Trajectories start from the
    seeded spectral state (test case 100) of the fine grid and are spun up
    for `spinup` coarse steps before sampling.
    """

    def __init__(self, gc=6, gf=8, dt_c=450.0, mlp_seed=1, horizon=1, window=0, skip=2, log=print):
        self.gc, self.gf, self.dt_c = gc, gf, dt_c
        self.horizon, self.window, self.skip = int(horizon), int(window), int(skip)
        self.ratio = 2 ** (gf - gc)
        self.log = log
        self.fine_mesh = Mesh.icos(gf)
        self.coarse_mesh = Mesh.icos(gc)
        self.Rh, self.Ru = restriction(self.fine_mesh, self.coarse_mesh)
        self.log(f"trainer: meshes glevel {gc} and {gf}, restriction built")
        self.mlp = Mlp(seed=mlp_seed)
        self.coarse = Swm(self.coarse_mesh, precision="fp32", dt=dt_c, h_ref=H0_GEO)
        self.fine = Swm(self.fine_mesh, precision="fp64", dt=dt_c / self.ratio, h_ref=H0_GEO)
        self.learn = Learn(self.coarse, self.mlp, self.coarse_mesh)
        self.log("trainer: handles created")
        self.train_rms = None

    def restrict(self):
        hf, uf = self.fine.get_state()
        return self.Rh @ hf, self.Ru @ uf

    def trajectory(self, seed, nsamples, spinup=0, raw=False):
        """Yield (raw features or None, du target, dh target) along one fine run.

        Targets (R = restriction, P = plain coarse step, x_f the fine run):
         * window == 0: the one-step defect from the restricted state,
           R x_f(t+dt) - P(R x_f(t)), with horizon K > 1 the K-step defect
           over K.
         * window W > 0: states from the coarse model's own plain run,
           restarted from R x_f every W steps with the first `skip` steps of
           each window not sampled; the target is the truth increment minus
           the coarse increment, (R x_f(t+dt) - R x_f(t)) - (P(x_c) - x_c).
           Features then come from states the free coarse run visits
           rather than from restricted fine states, whose first coarse step
           contains an adjustment to the restriction that a free run never
           makes.
        """
        K, W, S = self.horizon, self.window, self.skip
        h, u = self.fine_mesh.spectral(seed)
        self.fine.set_state(h, u)
        if spinup:
            self.fine.step(spinup * self.ratio)
        if W > 0:
            nper = W - S
            nwin = -(-nsamples // nper)
            steps = nwin * W
        else:
            steps = nsamples + K - 1
        states = [self.restrict()]
        for _ in range(steps):
            self.fine.step(self.ratio)
            states.append(self.restrict())
        if W <= 0:
            for i in range(nsamples):
                self.coarse.set_state(*states[i])
                feats = self.learn.sample(raw=raw)
                if K > 1:
                    self.coarse.step(K - 1)
                hp, up = self.coarse.get_state()
                hT, uT = states[i + K]
                yield feats, (uT - up) / K, (hT - hp) / K
            return
        done = 0
        for w in range(nwin):
            i0 = w * W
            self.coarse.set_state(*states[i0])
            if S:
                self.coarse.step(S)
            hx, ux = self.coarse.get_state()
            for i in range(i0 + S, i0 + W):
                if done == nsamples:
                    return
                feats = self.learn.sample(raw=raw)
                hp, up = self.coarse.get_state()
                (h0, u0), (h1, u1) = states[i], states[i + 1]
                yield feats, (u1 - u0) - (up - ux), (h1 - h0) - (hp - hx)
                hx, ux = hp, up
                done += 1

    def calibrate(self, seeds, nsamples=8, spinup=0):
        """Feature mean/std, target scales and the scaled training-target RMS."""
        raws, dus, dhs = [], [], []
        for sd in seeds:
            for f, du, dh in self.trajectory(sd, nsamples, spinup, raw=True):
                raws.append(f)
                dus.append(du)
                dhs.append(dh)
        self.log(f"calibrated on {len(raws)} samples")
        R = np.concatenate(raws)
        mu, sig = R.mean(0), R.std(0)
        sig[sig < 1e-6 * max(1.0, float(np.abs(mu).max()))] = 1.0
        self.tsu = 1.0 / float(np.sqrt(np.mean(np.concatenate(dus) ** 2)))
        self.tsh = 1.0 / float(np.sqrt(np.mean(np.concatenate(dhs) ** 2)))
        self.learn.set_scaling(mu, sig, self.tsu, self.tsh)
        self.learn_scaling = (mu.astype(np.float64), sig.astype(np.float64))
        # both scaled targets have unit mean square: RMS of the pair is sqrt(2)
        self.train_rms = float(np.sqrt(2.0))
        return mu, sig

    def collect(self, seeds, nsamples, spinup=0):
        """Store nsamples (features, targets) of each seeded run on the device."""
        self.learn.reserve(len(seeds) * nsamples)
        for k, sd in enumerate(seeds):
            for _, du, dh in self.trajectory(sd, nsamples, spinup):
                self.learn.store(du, dh)
            self.log(f"collected run {k + 1}/{len(seeds)} (seed {sd})")
        self.log(f"stored {len(seeds) * nsamples} samples")

    def fit(self, nsteps, lr0=1e-3, lr1=1e-4, wd=1e-4, seed=1, chunks=10):
        """Shuffled training on the stored samples, logged in chunks."""
        out, per = [], max(1, nsteps // chunks)
        for c in range(chunks):
            a, b = c * per, min(nsteps, (c + 1) * per)
            if a >= b:
                break
            la = lr0 + (lr1 - lr0) * a / max(1, nsteps - 1)
            lb = lr0 + (lr1 - lr0) * (b - 1) / max(1, nsteps - 1)
            out.append(self.learn.fit(b - a, la, lb, wd, seed + c))
            self.log(f"fit {b}/{nsteps}: loss u {out[-1][0]:.4f} h {out[-1][1]:.4f}")
        return np.array(out)

    def train(self, seeds, nsamples, spinup=0, epochs=2, lr=1e-3, wd=1e-4, lr_end=None):
        """Online training: `epochs` steps on each sample as it is produced."""
        losses, k, total = [], 0, len(seeds) * nsamples
        lr_end = lr if lr_end is None else lr_end
        for r, sd in enumerate(seeds):
            for _, du, dh in self.trajectory(sd, nsamples, spinup):
                a = k / max(1, total - 1)
                losses.append(self.learn.train(du, dh, epochs=epochs, lr=lr + (lr_end - lr) * a, wd=wd))
                k += 1
            lu, lh = np.mean(losses[-nsamples:], axis=0)
            self.log(f"run {r} (seed {sd}): loss u {lu:.4f} h {lh:.4f}")
        return np.array(losses)

    def bound(self, factor=3.0):
        return factor * self.train_rms


def rel_l2(a, b, w=None):
    w = np.ones_like(b) if w is None else w
    return float(np.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w * b ** 2)))


def _sync_time(fn):
    t = time.perf_counter()
    fn()
    return time.perf_counter() - t


def evaluate(tr, seed, nsteps, spinup=0, bound=None, mid=None):
    """Errors after nsteps coarse steps against the restricted fine run.

    Compares, from the same restricted start: the plain coarse model, the
    corrected one, and (when `mid` = (Swm at glevel gc+1, R_h, R_u fine->mid,
    R_h, R_u mid->coarse) is given) the plain model one level finer at half
    the time step. h errors are relative to the truth's anomaly
    (h - mean h), area-weighted; u errors are relative L2. Wall times are
    for the whole interval (each call ends with a device synchronisation).
    """
    bound = tr.bound() if bound is None else bound
    h, u = tr.fine_mesh.spectral(seed)
    tr.fine.set_state(h, u)
    if spinup:
        tr.fine.step(spinup * tr.ratio)
    hf0, uf0 = tr.fine.get_state()
    hc0, uc0 = tr.Rh @ hf0, tr.Ru @ uf0
    tr.fine.step(nsteps * tr.ratio)
    ht, ut = tr.restrict()
    A = tr.coarse_mesh.arrays["hx_areag"]
    ha = ht - np.sum(A * ht) / np.sum(A)

    def errs(hc, uc):
        return {"h": float(np.sqrt(np.sum(A * (hc - ht) ** 2) / np.sum(A * ha ** 2))), "u": rel_l2(uc, ut)}

    out = {"seed": seed, "nsteps": nsteps, "dt": tr.dt_c, "bound": bound}
    tr.coarse.set_state(hc0, uc0)
    tr.coarse.step(1)  # graph instantiation outside the timed region
    tr.coarse.set_state(hc0, uc0)
    t = _sync_time(lambda: tr.coarse.step(nsteps))
    hp, up = tr.coarse.get_state()
    out["plain"] = dict(errs(hp, up), wall_s=t)
    tr.coarse.set_state(hc0, uc0)
    tr.learn.step(0, bound)
    nfb = [0]
    t = _sync_time(lambda: nfb.__setitem__(0, tr.learn.step(nsteps, bound)))
    hl, ul = tr.coarse.get_state()
    out["learned"] = dict(errs(hl, ul), wall_s=t, limited=nfb[0])
    M0 = np.sum(A * hc0)
    out["learned"]["mass_drift"] = float((np.sum(A * hl) - M0) / M0)
    out["plain"]["mass_drift"] = float((np.sum(A * hp) - M0) / M0)
    if mid is not None:
        sm, Rfh, Rfu, Rmh, Rmu = mid
        sm.set_state(Rfh @ hf0, Rfu @ uf0)
        sm.step(1)
        sm.set_state(Rfh @ hf0, Rfu @ uf0)
        r = int(round(tr.dt_c / sm.opts.dt))
        t = _sync_time(lambda: sm.step(nsteps * r))
        hm, um = sm.get_state()
        out["mid"] = dict(errs(Rmh @ hm, Rmu @ um), wall_s=t)
    return out
