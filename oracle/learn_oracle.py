"""TEST INFRASTRUCTURE ONLY: numpy restatement of the learned-correction
coupling of libimx (imodel_amd/csrc/imx_learn.hip), in the caller's
numbering: the 48 edge-stencil features of a (state, plain-step state)
pair, the flux divergence D that turns the network's second output into
the cell-depth correction, and the corrected step's arithmetic.

No counterpart in the reference (the learned correction is new). What pins
it (tests/test_learn_oracle.py): the ring edges and their signs through
the mimetic identity sum_{ring of c} s_k u_k l~_k + ttgout_ce tnrccw_ce u_e
l~_e = tnrccw_ce A_c erad div_hx(u)_c, which vanishes for u = discrete curl
of a streamfunction; pentagon padding counts; the apex cells being the
third vertices of the edge's two triangles on either side of the edge
plane; D through sum_c A_c (D F)_c = 0 and D F = -kappa erad div_hx(F)
with div_hx taken from the reference-pinned C oracle
(src/swm_operators.f90:613-661).
"""
import numpy as np

NSE, NSC, NF = 11, 10, 48


def stencils(a):
    """Per caller edge: ring edges (11, -1 pad), their signs, stencil cells
    (10, -1 pad), 6 geometry features. `a`: mesh arrays (1-based tables)."""
    nnb = a["v_nnb"]
    nb = a["v_nb"].reshape(-1, 6) - 1
    ved = a["v_ed"].reshape(-1, 6) - 1
    ttg = a["hx_ttgout"].reshape(-1, 6)
    tnr = a["hx_tnrccw"].reshape(-1, 6)
    ev = a["ed_v"].reshape(-1, 2) - 1
    esh = a["ed_sh"].reshape(-1, 2) - 1
    tv = a["tr_v"].reshape(-1, 3) - 1
    p = a["v_p"].reshape(-1, 3)
    ne = len(ev)
    st_e = -np.ones((ne, NSE), dtype=np.int64)
    st_s = np.zeros((ne, NSE))
    st_c = -np.ones((ne, NSC), dtype=np.int64)
    for e in range(ne):
        cells = ev[e]
        slots = [int(np.nonzero(ved[c, :nnb[c]] == e)[0][0]) for c in cells]
        st_e[e, 0], st_s[e, 0] = e, 1.0
        k = 1
        for c, i in zip(cells, slots):
            n = nnb[c]
            for t in range(1, 6):
                if t < n:
                    j = (i + t) % n
                    st_e[e, k] = ved[c, j]
                    st_s[e, k] = -1.0 if ttg[c, j] * tnr[c, i] < 0 else 1.0
                k += 1
        nrm = np.cross(p[cells[0]], p[cells[1]])
        apex = [-1, -1]
        for t in esh[e]:
            for v in tv[t]:
                if v not in cells:
                    apex[0 if np.dot(nrm, p[v]) > 0 else 1] = v
        st_c[e, :4] = [cells[0], cells[1], apex[0], apex[1]]
        kc = 4
        for c, i in zip(cells, slots):
            n = nnb[c]
            for t in range(2, 5):
                if t <= n - 2:
                    st_c[e, kc] = nb[c, (i + t) % n]
                kc += 1
    z = a["ed_c"].reshape(-1, 3)[:, 2]
    l, dl, A = a["ed_leng"], a["edhx_leng"], a["hx_areag"]
    geo = np.stack([z, np.sqrt(np.maximum(0.0, 1 - z * z)), l / l.mean(), dl / dl.mean(),
                    A[ev[:, 0]] / A.mean(), A[ev[:, 1]] / A.mean()], 1)
    return st_e, st_s, st_c, geo


def features(st, hn, un, hp, up, href):
    """raw[ne][48] of the state (hn, un) and its plain step (hp, up)."""
    st_e, st_s, st_c, geo = st
    ne = len(st_e)
    raw = np.zeros((ne, NF))
    me = st_e >= 0
    ie = np.where(me, st_e, 0)
    raw[:, :NSE] = np.where(me, st_s * un[ie], 0.0)
    raw[:, NSE:2 * NSE] = np.where(me, st_s * (up[ie] - un[ie]), 0.0)
    mc = st_c >= 0
    ic = np.where(mc, st_c, 0)
    raw[:, 2 * NSE:2 * NSE + NSC] = np.where(mc, hn[ic] - href, 0.0)
    raw[:, 2 * NSE + NSC:2 * NSE + 2 * NSC] = np.where(mc, hp[ic] - hn[ic], 0.0)
    raw[:, 2 * NSE + 2 * NSC:] = geo
    return raw


def div_matrix(a):
    """D (nv x ne): (D F)_c = -kappa / A_c sum_{e in c} ttgout_ce l~_e F_e,
    kappa = mean(A) / mean(l~)."""
    from scipy.sparse import csr_matrix
    nnb = a["v_nnb"]
    ved = a["v_ed"].reshape(-1, 6) - 1
    ttg = a["hx_ttgout"].reshape(-1, 6)
    A, dl = a["hx_areag"], a["edhx_leng"]
    kappa = A.mean() / dl.mean()
    rows, cols, vals = [], [], []
    for c in range(len(nnb)):
        for j in range(nnb[c]):
            e = ved[c, j]
            rows.append(c)
            cols.append(e)
            vals.append(-kappa * np.sign(ttg[c, j]) * dl[e] / A[c])
    return csr_matrix((vals, (rows, cols)), shape=(len(nnb), len(dl)))


def correct(h, u, Y, D, tsu, tsh, bound):
    """One correction: (h, u) + s (D Y[:,1] / tsh, Y[:,0] / tsu) with
    s = 1 while the RMS sqrt(mean Y0^2 + mean (D Y1)^2) is within bound,
    s = bound / RMS above it, and the plain (h, u) when s is 0 or the RMS
    is not finite. Returns (h, u, rms)."""
    dhc = D @ Y[:, 1]
    rms = float(np.sqrt(np.mean(Y[:, 0] ** 2) + np.mean(dhc ** 2)))
    s = 1.0 if rms <= bound else (bound / rms if np.isfinite(rms) else 0.0)
    if not s > 0.0:
        return h, u, rms
    return h + s * (dhc / tsh), u + s * (Y[:, 0] / tsu), rms
