"""Python host mirroring the reference ``swm`` module over libimx.

Reference call surface (src/swm.f90): ``tendency(h, u, masseq, momeq)``
(:2416-2598), ``ode_rk4(t, h, u, h_new, u_new, dt)`` (:2600-2683) and the
swm_tests time loop (:238-370). Errors are raised as ``capi.ImxError``
carrying libimx's message (the reference prints and stops).
"""
import ctypes
import os
import tempfile

import numpy as np

from . import capi, imxbin

PREC = {"fp64_ref": capi.PREC_FP64_REF, "fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32}
ERAD_REF = 6.37122e6   # constants.f90:55
GRAV = 9.80616         # constants.f90:61
OMEGA = 7.292e-5       # constants.f90:67
H0_GEO = 2.94e4 / GRAV  # mean depth for gH0 = 2.94e4 m^2/s^2

_SHAPES = {
    "v_nnb": ("nv", 1), "v_nb": ("nv", 6), "v_ed": ("nv", 6), "v_tr": ("nv", 6),
    "hx_ttgout": ("nv", 6), "hx_tnrccw": ("nv", 6), "v_p": ("nv", 3), "hx_areag": ("nv", 1),
    "ed_v": ("ne", 2), "ed_sh": ("ne", 2), "ed_leng": ("ne", 1), "edhx_leng": ("ne", 1),
    "ed_c": ("ne", 3), "edhx_c": ("ne", 3), "tr_v": ("nt", 3), "tr_ed": ("nt", 3),
    "tr_tg": ("nt", 3), "tr_areag": ("nt", 1), "tr_c": ("nt", 3), "tr_clat": ("nt", 1),
    "tr_trhx_area": ("nt", 3), "hx_hxtr_area": ("nv", 6), "hx_trskw": ("nv", 36),
}


class Mesh:
    """A mesh in the reference's flat layout (dict of IMXBIN record arrays)."""

    def __init__(self, arrays, handle=None):
        self.arrays = arrays
        self._handle = handle
        self.nv, self.ne, self.nt = (int(x) for x in arrays["dims"][:3])
        self.glevel = int(arrays["dims"][3]) if len(arrays["dims"]) > 3 else -1
        self._desc = None

    @classmethod
    def icos(cls, glevel, pos="eqs", gridprop=True):
        """Icosahedral non-optimised grid (meshbuild for icos/nopt)."""
        L = capi.lib()
        h = ctypes.c_void_p()
        capi.check(L.imx_mesh_build_icos(int(glevel), capi.POS_EQS if pos == "eqs" else capi.POS_POL,
                                         ctypes.byref(h)), "imx_mesh_build_icos")
        return cls._from_handle(h, gridprop)

    @classmethod
    def build_read(cls, path, gridprop=True):
        """meshbuild for mesh kind "read": nodes from a .xyz (or .gmt lon/lat)
        file and the neighbour rings from the .ngb file beside it."""
        h = ctypes.c_void_p()
        capi.check(capi.lib().imx_mesh_build_read(str(path).encode(), ctypes.byref(h)), "imx_mesh_build_read")
        return cls._from_handle(h, gridprop)

    @classmethod
    def read(cls, path, gridprop=True):
        """A mesh from an IMXBIN file in imx_mesh_write's layout (the
        reference harness's mesh dumps share it)."""
        h = ctypes.c_void_p()
        capi.check(capi.lib().imx_mesh_read(str(path).encode(), ctypes.byref(h)), "imx_mesh_read")
        return cls._from_handle(h, gridprop)

    @classmethod
    def _from_handle(cls, h, gridprop):
        L = capi.lib()
        if gridprop:
            capi.check(L.imx_mesh_gridprop(h), "imx_mesh_gridprop")
        d = capi.MeshDesc()
        capi.check(L.imx_mesh_desc_of(h, ctypes.byref(d)), "imx_mesh_desc_of")
        n = {"nv": d.nv, "ne": d.ne, "nt": d.nt}
        arrays = {"dims": np.array([d.nv, d.ne, d.nt, d.glevel], dtype=np.int32)}
        for name, (ent, w) in _SHAPES.items():
            ptr = getattr(d, name)
            if not ptr:
                continue
            arrays[name] = np.ctypeslib.as_array(ptr, shape=(n[ent] * w,)).copy()
        m = cls(arrays, handle=h)
        return m

    @classmethod
    def load(cls, path):
        return cls(imxbin.read(path))

    @classmethod
    def load_ref(cls, griddir, name, gridprop=True):
        """Grid files written by the reference's meshwrite (<griddir>/<name>_*.dat)."""
        h = ctypes.c_void_p()
        capi.check(capi.lib().imx_mesh_load_ref(griddir.encode(), name.encode(), ctypes.byref(h)),
                   "imx_mesh_load_ref")
        return cls._from_handle(h, gridprop)

    def write_ref(self, griddir, name):
        """Write the reference's 11 grid files (meshload reads them)."""
        if self._handle is None:
            raise ValueError("write_ref needs a library mesh (Mesh.icos or Mesh.load_ref)")
        capi.check(capi.lib().imx_mesh_write_ref(self._handle, griddir.encode(), name.encode()),
                   "imx_mesh_write_ref")

    def write(self, path):
        """Full IMXBIN record set (library meshes carry every record)."""
        if self._handle is not None:
            capi.check(capi.lib().imx_mesh_write(self._handle, path.encode()), "imx_mesh_write")
        else:
            imxbin.write(path, self.arrays)

    def full_arrays(self):
        """Every IMXBIN record, including those outside imx_mesh_desc."""
        if self._handle is None:
            return self.arrays
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "m.bin")
            self.write(p)
            return imxbin.read(p)

    def desc(self):
        if self._desc is None:
            self._desc = capi.make_desc(self.arrays)
        return self._desc[0]

    def remap_ll(self, f, pos=0, nlon=1440, nlat=720):
        """plot_scalarfield's lat-lon sampling: (nlat*nlon, 3) float32 rows
        (lon, lat, value), the stream the reference writes."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros((nlat * nlon, 3), dtype=np.float32)
        capi.check(capi.lib().imx_remap_ll(ctypes.byref(self.desc()), int(pos), capi.f64p(f), int(nlon), int(nlat),
                                           capi.f32p(out)), "imx_remap_ll")
        return out

    def tc2(self, erad=ERAD_REF, stag="HTC"):
        """Williamson TC2 state (initialize_fields case 2) on a library mesh,
        u for HTC (along ed%tg at the edge crossings) or HC (along edhx%nr
        at the Voronoi edge midpoints)."""
        if self._handle is None:
            raise ValueError("tc2 needs a library-built mesh")
        h = np.zeros(self.nv)
        u = np.zeros(self.ne)
        fn = capi.lib().imx_ic_tc2_hc if stag == "HC" else capi.lib().imx_ic_tc2
        capi.check(fn(self._handle, erad, capi.f64p(h), capi.f64p(u)), "imx_ic_tc2")
        return h, u

    def testcase(self, tc, stag="HTC", erad=ERAD_REF, hollgw=2.0):
        """initialize_fields of test case tc (0, 1, 2, 5, 6, 11, 12, 21, 22,
        23, 32-35, 40-42, 51, 52, 53) on a library mesh: (h, u, bt), u for
        HTC or HC staggering; hollgw is par/swm.par's Hollingsworth depth
        (test cases 32-35 and 51). The Coriolis parameter a case runs with
        is testcase_coriolis(tc)."""
        if self._handle is None:
            raise ValueError("testcase needs a library-built mesh")
        h, u, bt = np.zeros(self.nv), np.zeros(self.ne), np.zeros(self.nv)
        capi.check(capi.lib().imx_ic_testcase(self._handle, int(tc), capi.STAG[stag], float(erad), float(hollgw),
                                              capi.f64p(h), capi.f64p(u), capi.f64p(bt)), "imx_ic_testcase")
        return h, u, bt

    @staticmethod
    def testcase_coriolis(tc):
        """(fsphere, fcte) of test case tc, for Swm.set_coriolis."""
        fs = np.zeros(1, dtype=np.int32)
        fc = np.zeros(1)
        capi.check(capi.lib().imx_ic_testcase_coriolis(int(tc), capi.i32p(fs), capi.f64p(fc)),
                   "imx_ic_testcase_coriolis")
        return int(fs[0]), float(fc[0])

    def spectral(self, seed, lmax=16, h0=H0_GEO, hfrac=0.05, urms=20.0, erad=ERAD_REF):
        """Seeded spectral state (h, u)."""
        h = np.zeros(self.nv)
        u = np.zeros(self.ne)
        capi.check(capi.lib().imx_ic_spectral(ctypes.byref(self.desc()), int(seed), int(lmax), h0, hfrac,
                                              urms, erad, capi.f64p(h), capi.f64p(u)), "imx_ic_spectral")
        return h, u

    def gridprop(self):
        """initialize_gridprop products for this mesh (trhx, hxtr, trskw)."""
        trhx = np.zeros(3 * self.nt)
        hxtr = np.zeros(6 * self.nv)
        trskw = np.zeros(36 * self.nv)
        capi.check(capi.lib().imx_gridprop(ctypes.byref(self.desc()), capi.f64p(trhx), capi.f64p(hxtr),
                                           capi.f64p(trskw)), "imx_gridprop")
        return trhx, hxtr, trskw

    def __del__(self):
        if getattr(self, "_handle", None) is not None:
            try:
                capi.lib().imx_mesh_destroy(self._handle)
            except Exception:
                pass
            self._handle = None


class Swm:
    """One shallow-water instance resident on the GPU (libimx handle)."""

    def __init__(self, mesh, precision="fp32", dt=900.0, erad=ERAD_REF, grav=GRAV, omega=OMEGA,
                 testcase=2, use_sfc=True, use_graph=True, h_ref=0.0):
        L = capi.lib()
        self.mesh = mesh
        self.precision = precision
        o = capi.SwmOpts()
        L.imx_swm_default_opts(ctypes.byref(o))
        o.precision = PREC[precision]
        o.testcase = int(testcase)
        o.use_sfc = int(bool(use_sfc))
        o.use_graph = int(bool(use_graph))
        o.erad, o.grav, o.omega, o.dt, o.h_ref = erad, grav, omega, float(dt), float(h_ref)
        self.opts = o
        self._h = ctypes.c_void_p()
        capi.check(L.imx_swm_create(ctypes.byref(mesh.desc()), ctypes.byref(o), ctypes.byref(self._h)),
                   "imx_swm_create")
        self.nv, self.ne, self.nt = mesh.nv, mesh.ne, mesh.nt

    def close(self):
        if self._h:
            capi.lib().imx_swm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, h, u):
        h = np.ascontiguousarray(h, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        capi.check(capi.lib().imx_swm_set_state(self._h, capi.f64p(h), capi.f64p(u)), "imx_swm_set_state")

    def set_topography(self, bt):
        """Bottom topography b [nv] (None: b = 0); h stays the depth."""
        if bt is None:
            capi.check(capi.lib().imx_swm_set_topography(self._h, None), "imx_swm_set_topography")
            return
        bt = np.ascontiguousarray(bt, dtype=np.float64)
        capi.check(capi.lib().imx_swm_set_topography(self._h, capi.f64p(bt)), "imx_swm_set_topography")

    def set_coriolis(self, fsphere=0, fcte=0.0):
        """The Coriolis term: fsphere 0 (f = 2 Omega sin(lat)), 1 (eta = fcte
        alone, the linearised equations) or 2 (f = fcte), as vort_tr."""
        fsphere = capi.FSPHERE.get(fsphere, fsphere)
        capi.check(capi.lib().imx_swm_set_coriolis(self._h, int(fsphere), float(fcte)), "imx_swm_set_coriolis")

    def get_state(self):
        h = np.zeros(self.nv)
        u = np.zeros(self.ne)
        capi.check(capi.lib().imx_swm_get_state(self._h, capi.f64p(h), capi.f64p(u)), "imx_swm_get_state")
        return h, u

    def set_scheme(self, stag="HTC", wrap="trsk10", pv="none", pvspar=0.5, k2=0.0, diffus="const", k4=0.0,
                   hyperdiffus="const", sint="trsk", cor="trsk", rec="trsk", gasscoef=1.0, nopv=False):
        """Scheme options of par/swm.par: the staggering ("HTC" | "HC"),
        the method wrapper ("trsk10" | "pxt16" | "gass18" | "none"), APVM
        (pv="apvm", parameter pvspar), diffusion K2_max with coefficient
        function diffus, hyperdiffusion K4_max with hyperdiffus ("const" |
        "align" | "diam"). With wrap="none" the individual methods: scalar
        interpolation sint ("trsk" | "bary" | "gass"), Coriolis cor ("trsk" |
        "pered" | "gass" | "dtred"), kinetic energy rec ("trsk" | "perhx" |
        "gass" with gasscoef | "dubos" | "melv") and the level model nopv.
        Every option runs on FP64_REF handles; fused (FP32/FP64) handles take
        every triple of sint "trsk"/"bary", cor "trsk"/"pered"/"dtred" and
        any rec, sint "gass" with cor "gass"/"trsk" or cor "gass" with sint
        "trsk"/"bary" (no diffusion) with any rec, and nopv with trsk10's methods, with APVM, diffusion and
        hyperdiffusion, HTC or HC, and refuse the others."""
        if rec == "dubos":  # Gassmann's K with no trsk part (swmpars, src/swm_data.f90:513-516)
            rec, gasscoef = "gass", 0.0
        sc = capi.Scheme(stag=capi.STAG[stag], wrap=capi.WRAP[wrap], apvm=int(pv == "apvm"), pvspar=float(pvspar), k2=float(k2),
                         diffus=capi.COEF[diffus], k4=float(k4), hyperdiffus=capi.COEF[hyperdiffus],
                         sinterp=capi.SINT[sint], coriolis=capi.COR[cor], recon=capi.REC[rec],
                         gasscoef=float(gasscoef), nopv=int(bool(nopv)))
        capi.check(capi.lib().imx_swm_set_scheme(self._h, ctypes.byref(self.mesh.desc()), ctypes.byref(sc)),
                   "imx_swm_set_scheme")

    def tendency(self, h, u):
        """(masseq, momeq) = RHS of the mass and momentum equations."""
        h = np.ascontiguousarray(h, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        m = np.zeros(self.nv)
        f = np.zeros(self.ne)
        capi.check(capi.lib().imx_swm_tendency(self._h, capi.f64p(h), capi.f64p(u), capi.f64p(m),
                                               capi.f64p(f)), "imx_swm_tendency")
        return m, f

    def jacobian(self, h, u, pert=None):
        """Linearised operator (swm_normal_mode_analysis): scipy CSR matrix
        M[r, i] = (T(x + pert e_i) - T(x)) / pert, unknowns (h, u); pert
        defaults to the mean cell area on the unit sphere (the reference's
        mesh%meanhxarea). Returns (M, number of tendency evaluations)."""
        from scipy.sparse import csr_matrix
        h = np.ascontiguousarray(h, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        if pert is None:
            pert = float(np.mean(self.mesh.arrays["hx_areag"]))
        nnz = ctypes.c_int64()
        ncol = np.zeros(1, dtype=np.int32)
        L = capi.lib()
        d = ctypes.byref(self.mesh.desc())
        capi.check(L.imx_swm_jacobian(self._h, d, capi.f64p(h), capi.f64p(u), pert, 0, None, None, None,
                                      ctypes.byref(nnz), capi.i32p(ncol)), "imx_swm_jacobian")
        r = np.zeros(nnz.value, dtype=np.int32)
        c = np.zeros(nnz.value, dtype=np.int32)
        v = np.zeros(nnz.value)
        capi.check(L.imx_swm_jacobian(self._h, d, capi.f64p(h), capi.f64p(u), pert, nnz.value, capi.i32p(r),
                                      capi.i32p(c), capi.f64p(v), ctypes.byref(nnz), capi.i32p(ncol)),
                   "imx_swm_jacobian")
        n = self.nv + self.ne
        return csr_matrix((v, (r - 1, c - 1)), shape=(n, n)), int(ncol[0]) + 1

    def field(self, name):
        pos = capi.FIELD_POS[name]
        n = {"v": self.nv, "e": self.ne, "t": self.nt}[pos[0]]
        out = np.zeros((n, 3) if pos.endswith("3") else n)
        capi.check(capi.lib().imx_swm_get_field(self._h, capi.FIELDS.index(name), capi.f64p(out)),
                   "imx_swm_get_field")
        return out

    def lte(self, hollgw=2.0):
        """Local truncation error analysis (swm_horiz_loc_trunc_er) of the
        handle's test case, 2 or 32-35 (layer depth hollgw): sets the case as
        the state, evaluates the tendency and returns
        ({row: (max relative, relative L2, max)}, {mimetic sum: value}).
        Needs an FP64_REF handle on a library mesh."""
        if self.mesh._handle is None:
            raise ValueError("lte needs a library mesh (Mesh.icos, Mesh.read, Mesh.load_ref)")
        err = np.zeros(3 * len(capi.LTE_ROWS))
        mim = np.zeros(len(capi.LTE_MIM))
        capi.check(capi.lib().imx_swm_lte_hol(self._h, self.mesh._handle, float(hollgw), capi.f64p(err),
                                              capi.f64p(mim)), "imx_swm_lte_hol")
        rows = {k: tuple(err[3 * i:3 * i + 3]) for i, k in enumerate(capi.LTE_ROWS)}
        return rows, dict(zip(capi.LTE_MIM, mim))

    def step(self, nsteps=1):
        capi.check(capi.lib().imx_swm_step(self._h, int(nsteps)), "imx_swm_step")

    def step_timed(self, nsteps):
        ms = ctypes.c_double()
        capi.check(capi.lib().imx_swm_step_timed(self._h, int(nsteps), ctypes.byref(ms)), "imx_swm_step_timed")
        return ms.value

    def profile(self, nsteps):
        ms = np.zeros(16)
        nk = ctypes.c_int32()
        capi.check(capi.lib().imx_swm_profile(self._h, int(nsteps), capi.f64p(ms), ctypes.byref(nk)),
                   "imx_swm_profile")
        return {capi.lib().imx_swm_kernel_name(self._h, k).decode(): ms[k] for k in range(nk.value)}

    def diagnostics(self):
        out = np.zeros(capi.NDIAG)
        capi.check(capi.lib().imx_swm_diagnostics(self._h, capi.f64p(out)), "imx_swm_diagnostics")
        return dict(zip(capi.DIAG_NAMES, out))

    def info(self):
        out = (ctypes.c_int64 * 8)()
        capi.check(capi.lib().imx_swm_info(self._h, out), "imx_swm_info")
        return list(out)

    def tune(self, nsteps=3):
        """Autotune block size and XCD remap of each fused kernel."""
        out = np.zeros(6, dtype=np.int32)
        capi.check(capi.lib().imx_swm_tune(self._h, int(nsteps), capi.i32p(out)), "imx_swm_tune")
        return {k: (int(out[2 * i]), int(out[2 * i + 1])) for i, k in enumerate(("tri", "cell", "edge"))}


class Ensemble:
    """Independent instances (meshes, time steps) advanced together on the GPU."""

    def __init__(self, meshes, dts, precision="fp32", erad=ERAD_REF, grav=GRAV, omega=OMEGA,
                 testcase=2, use_sfc=True, use_graph=True, h_ref=0.0):
        L = capi.lib()
        self.meshes = list(meshes)
        n = len(self.meshes)
        descs = (capi.MeshDesc * n)(*[m.desc() for m in self.meshes])
        self._dts = np.ascontiguousarray(dts, dtype=np.float64)
        o = capi.SwmOpts()
        L.imx_swm_default_opts(ctypes.byref(o))
        o.precision = PREC[precision]
        o.testcase = int(testcase)
        o.use_sfc, o.use_graph = int(bool(use_sfc)), int(bool(use_graph))
        o.erad, o.grav, o.omega, o.dt, o.h_ref = erad, grav, omega, float(self._dts[0]), float(h_ref)
        self._h = ctypes.c_void_p()
        capi.check(L.imx_ens_create(n, descs, capi.f64p(self._dts), ctypes.byref(o), ctypes.byref(self._h)),
                   "imx_ens_create")
        self.n = n

    def close(self):
        if self._h:
            capi.lib().imx_ens_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, i, h, u):
        h = np.ascontiguousarray(h, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        capi.check(capi.lib().imx_ens_set_state(self._h, i, capi.f64p(h), capi.f64p(u)), "imx_ens_set_state")

    def get_state(self, i):
        m = self.meshes[i]
        h = np.zeros(m.nv)
        u = np.zeros(m.ne)
        capi.check(capi.lib().imx_ens_get_state(self._h, i, capi.f64p(h), capi.f64p(u)), "imx_ens_get_state")
        return h, u

    def step(self, nsteps=1):
        capi.check(capi.lib().imx_ens_step(self._h, int(nsteps)), "imx_ens_step")

    def step_timed(self, nsteps):
        ms = ctypes.c_double()
        capi.check(capi.lib().imx_ens_step_timed(self._h, int(nsteps), ctypes.byref(ms)), "imx_ens_step_timed")
        return ms.value

    def diagnostics(self, i):
        out = np.zeros(capi.NDIAG)
        capi.check(capi.lib().imx_ens_diagnostics(self._h, i, capi.f64p(out)), "imx_ens_diagnostics")
        return dict(zip(capi.DIAG_NAMES, out))

    def tune(self, nsteps=2):
        out = np.zeros(6, dtype=np.int32)
        capi.check(capi.lib().imx_ens_tune(self._h, int(nsteps), capi.i32p(out)), "imx_ens_tune")
        return {k: (int(out[2 * i]), int(out[2 * i + 1])) for i, k in enumerate(("tri", "cell", "edge"))}

    def info(self):
        out = (ctypes.c_int64 * 8)()
        capi.check(capi.lib().imx_ens_info(self._h, out), "imx_ens_info")
        return list(out)


def cfl_dt(mesh, u, cfl=0.8, erad=ERAD_REF, gh0=2.94e4):
    """Time step at Courant number cfl for c = sqrt(g H0) + max|u| over the
    shortest distance between cell centres (min ed%leng)."""
    c = np.sqrt(gh0) + np.max(np.abs(u))
    return cfl * np.min(mesh.arrays["ed_leng"]) * erad / c


def tendency(swm, h, u):
    """subroutine tendency(h, u, masseq, momeq) (src/swm.f90:2416)."""
    return swm.tendency(h, u)


def ode_rk4(swm, h, u):
    """subroutine ode_rk4(t, h, u, h_new, u_new, dt) (src/swm.f90:2600): one step."""
    swm.set_state(h, u)
    swm.step(1)
    return swm.get_state()
