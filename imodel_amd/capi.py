"""ctypes view of the libimx C ABI (include/imx_swm.h).

The shared library is built in-tree by ``__graft_entry__.build()`` into
``imodel_amd/libimx.so``. Loading fails loudly when it is missing: there is
no CPU fallback for any product entry point.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libimx.so")

IMX_MAXNB = 6
PREC_FP64_REF, PREC_FP64, PREC_FP32 = 0, 1, 2
POS_EQS, POS_POL = 0, 1
NDIAG = 8
DIAG_NAMES = ("mass", "penergy", "kenergy", "tenergy", "aenergy", "penstrophy",
              "rmsdiv", "badflag")
FIELDS = ("h_ed", "uh", "eta", "h_tr", "q_tr", "q_ed", "divuh", "ke_hx", "ghbk",
          "grad_ghbk", "uhq_perp", "divu", "zeta", "q_grad_ed", "lapu", "lap_lapu",
          "q_hx", "vhq_tr", "vh_hx", "vhq_hx", "q_grad_tr")
FIELD_POS = {"h_ed": "e", "uh": "e", "eta": "t", "h_tr": "t", "q_tr": "t",
             "q_ed": "e", "divuh": "v", "ke_hx": "v", "ghbk": "v",
             "grad_ghbk": "e", "uhq_perp": "e", "divu": "v", "zeta": "t",
             "q_grad_ed": "e", "lapu": "e", "lap_lapu": "e",
             "q_hx": "v", "vhq_tr": "t3", "vh_hx": "v3", "vhq_hx": "v3", "q_grad_tr": "t3"}
# rows of imx_swm_lte (the table swm_horiz_loc_trunc_er prints) and its
# mimetic sums
LTE_ROWS = ("h_ed", "divuh", "kenergy", "abs_vort", "h_tr", "q_tr", "q_hx", "q_ed", "grad_ghbK",
            "grad_ghbK_exact", "vhq_tr", "q_grad_tr", "vh_hx", "vhq_hx", "uhq_perp", "lapu", "masseq",
            "momeq")
LTE_MIM = ("hdivu", "prodrule", "divuh", "proddiff", "crossrule", "coriolis", "coriolis_rel")
COEF = {"const": 0, "align": 1, "diam": 2}
WRAP = {"trsk10": 0, "pxt16": 1, "gass18": 2, "none": 3}
SINT = {"trsk": 0, "bary": 1, "gass": 2}
COR = {"trsk": 0, "pered": 1, "gass": 2, "dtred": 3}
REC = {"trsk": 0, "perhx": 1, "gass": 2, "melv": 3}
STAG = {"HTC": 0, "HC": 1}
FSPHERE = {"variable": 0, "linear": 1, "constant": 2}

_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)
_F32P = ctypes.POINTER(ctypes.c_float)


class Scheme(ctypes.Structure):
    """imx_scheme: APVM, diffusion and hyperdiffusion options (par/swm.par)."""
    _fields_ = [("stag", ctypes.c_int32), ("wrap", ctypes.c_int32), ("apvm", ctypes.c_int32),
                ("pvspar", ctypes.c_double),
                ("k2", ctypes.c_double),
                ("diffus", ctypes.c_int32), ("k4", ctypes.c_double), ("hyperdiffus", ctypes.c_int32),
                ("sinterp", ctypes.c_int32), ("coriolis", ctypes.c_int32), ("recon", ctypes.c_int32),
                ("gasscoef", ctypes.c_double), ("nopv", ctypes.c_int32)]


class MeshDesc(ctypes.Structure):
    """imx_mesh_desc: flat arrays of the reference grid_structure (1-based)."""
    _fields_ = [
        ("nv", ctypes.c_int32), ("ne", ctypes.c_int32), ("nt", ctypes.c_int32),
        ("glevel", ctypes.c_int32),
        ("v_nnb", _I32P), ("v_nb", _I32P), ("v_ed", _I32P), ("v_tr", _I32P),
        ("hx_ttgout", _I32P), ("hx_tnrccw", _I32P), ("v_p", _F64P),
        ("hx_areag", _F64P),
        ("ed_v", _I32P), ("ed_sh", _I32P), ("ed_leng", _F64P), ("edhx_leng", _F64P),
        ("ed_c", _F64P), ("edhx_c", _F64P),
        ("tr_v", _I32P), ("tr_ed", _I32P), ("tr_tg", _I32P), ("tr_areag", _F64P),
        ("tr_c", _F64P), ("tr_clat", _F64P),
        ("tr_trhx_area", _F64P), ("hx_hxtr_area", _F64P), ("hx_trskw", _F64P),
    ]


DESC_ARRAYS = [f for f, _ in MeshDesc._fields_[4:]]
OPTIONAL_ARRAYS = ("tr_trhx_area", "hx_hxtr_area", "hx_trskw")


def make_desc(arrays):
    """Build a MeshDesc over a dict of numpy arrays (IMXBIN record names).

    Returns (desc, keepalive); keep ``keepalive`` referenced while the
    descriptor is in use.
    """
    nv, ne, nt = (int(x) for x in arrays["dims"][:3])
    glevel = int(arrays["dims"][3]) if len(arrays["dims"]) > 3 else -1
    keep = {}
    d = MeshDesc(nv=nv, ne=ne, nt=nt, glevel=glevel)
    for name, ctype in MeshDesc._fields_[4:]:
        if name not in arrays:
            if name in OPTIONAL_ARRAYS:
                continue
            raise KeyError(f"mesh record {name!r} missing")
        if ctype is _I32P:
            a = np.ascontiguousarray(arrays[name], dtype=np.int32)
        else:
            a = np.ascontiguousarray(arrays[name], dtype=np.float64)
        keep[name] = a
        setattr(d, name, a.ctypes.data_as(ctype))
    return d, keep


class SwmOpts(ctypes.Structure):
    """imx_swm_opts."""
    _fields_ = [("precision", ctypes.c_int32), ("testcase", ctypes.c_int32),
                ("use_sfc", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("erad", ctypes.c_double), ("grav", ctypes.c_double),
                ("omega", ctypes.c_double), ("dt", ctypes.c_double),
                ("h_ref", ctypes.c_double)]


_lib = None


def lib():
    """The loaded libimx (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libimx not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        sig = {
            "imx_last_error": (ctypes.c_char_p, []),
            "imx_mesh_build_icos": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(vp)]),
            "imx_mesh_desc_of": (ctypes.c_int, [vp, ctypes.POINTER(MeshDesc)]),
            "imx_mesh_gridprop": (ctypes.c_int, [vp]),
            "imx_mesh_write": (ctypes.c_int, [vp, ctypes.c_char_p]),
            "imx_mesh_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
            "imx_mesh_build_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
            "imx_swm_lte": (ctypes.c_int, [vp, vp, _F64P, _F64P]),
            "imx_swm_lte_hol": (ctypes.c_int, [vp, vp, ctypes.c_double, _F64P, _F64P]),
            "imx_mesh_load_ref": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]),
            "imx_mesh_write_ref": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p]),
            "imx_make_dir": (ctypes.c_int, [ctypes.c_char_p]),
            "imx_swm_jacobian": (ctypes.c_int, [vp, ctypes.POINTER(MeshDesc), _F64P, _F64P, ctypes.c_double,
                                                ctypes.c_int64, _I32P, _I32P, _F64P,
                                                ctypes.POINTER(ctypes.c_int64), _I32P]),
            "imx_remap_ll": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_int32, _F64P, ctypes.c_int32,
                                            ctypes.c_int32, _F32P]),
            "imx_mesh_destroy": (None, [vp]),
            "imx_gridprop": (ctypes.c_int, [ctypes.POINTER(MeshDesc), _F64P, _F64P, _F64P]),
            "imx_ic_tc2": (ctypes.c_int, [vp, ctypes.c_double, _F64P, _F64P]),
            "imx_ic_tc2_hc": (ctypes.c_int, [vp, ctypes.c_double, _F64P, _F64P]),
            "imx_ic_tc2_pv": (ctypes.c_int, [vp, ctypes.c_double, _F64P]),
            "imx_ic_testcase": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                               ctypes.c_double, _F64P, _F64P, _F64P]),
            "imx_swm_set_topography": (ctypes.c_int, [vp, _F64P]),
            "imx_swm_set_coriolis": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_double]),
            "imx_ic_testcase_coriolis": (ctypes.c_int, [ctypes.c_int32, _I32P, _F64P]),
            "imx_ic_spectral": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_uint64,
                                               ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, _F64P, _F64P]),
            "imx_swm_default_opts": (None, [ctypes.POINTER(SwmOpts)]),
            "imx_swm_create": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(SwmOpts),
                                              ctypes.POINTER(vp)]),
            "imx_swm_set_state": (ctypes.c_int, [vp, _F64P, _F64P]),
            "imx_swm_get_state": (ctypes.c_int, [vp, _F64P, _F64P]),
            "imx_swm_tendency": (ctypes.c_int, [vp, _F64P, _F64P, _F64P, _F64P]),
            "imx_swm_get_field": (ctypes.c_int, [vp, ctypes.c_int32, _F64P]),
            "imx_swm_set_scheme": (ctypes.c_int, [vp, ctypes.POINTER(MeshDesc), ctypes.POINTER(Scheme)]),
            "imx_swm_step": (ctypes.c_int, [vp, ctypes.c_int32]),
            "imx_swm_step_timed": (ctypes.c_int, [vp, ctypes.c_int32, _F64P]),
            "imx_swm_profile": (ctypes.c_int, [vp, ctypes.c_int32, _F64P, _I32P]),
            "imx_swm_kernel_name": (ctypes.c_char_p, [vp, ctypes.c_int32]),
            "imx_swm_diagnostics": (ctypes.c_int, [vp, _F64P]),
            "imx_swm_info": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
            "imx_swm_destroy": (None, [vp]),
            "imx_copy_probe": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _F64P]),
            "imx_swm_tune": (ctypes.c_int, [vp, ctypes.c_int32, _I32P]),
            "imx_ens_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(MeshDesc), _F64P,
                                              ctypes.POINTER(SwmOpts), ctypes.POINTER(vp)]),
            "imx_ens_set_state": (ctypes.c_int, [vp, ctypes.c_int32, _F64P, _F64P]),
            "imx_ens_get_state": (ctypes.c_int, [vp, ctypes.c_int32, _F64P, _F64P]),
            "imx_ens_step": (ctypes.c_int, [vp, ctypes.c_int32]),
            "imx_ens_step_timed": (ctypes.c_int, [vp, ctypes.c_int32, _F64P]),
            "imx_ens_diagnostics": (ctypes.c_int, [vp, ctypes.c_int32, _F64P]),
            "imx_ens_tune": (ctypes.c_int, [vp, ctypes.c_int32, _I32P]),
            "imx_ens_info": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
            "imx_ens_destroy": (None, [vp]),
            "imx_mlp_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(vp)]),
            "imx_mlp_nparams": (ctypes.c_int, []),
            "imx_mlp_get_params": (ctypes.c_int, [vp, _F32P]),
            "imx_mlp_set_params": (ctypes.c_int, [vp, _F32P]),
            "imx_mlp_init": (ctypes.c_int, [vp, ctypes.c_uint64]),
            "imx_mlp_forward": (ctypes.c_int, [vp, ctypes.c_int64, _F32P, _F32P]),
            "imx_mlp_train": (ctypes.c_int, [vp, ctypes.c_int64, _F32P, _F32P, _F32P, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_int32, _F64P, _F32P, ctypes.c_int32]),
            "imx_mlp_forward_timed": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int32, _F64P]),
            "imx_mlp_destroy": (None, [vp]),
            "imx_learn_create": (ctypes.c_int, [vp, vp, ctypes.POINTER(MeshDesc), ctypes.POINTER(vp)]),
            "imx_learn_sample": (ctypes.c_int, [vp, _F32P]),
            "imx_learn_set_scaling": (ctypes.c_int, [vp, _F32P, _F32P, ctypes.c_float, ctypes.c_float]),
            "imx_learn_train": (ctypes.c_int, [vp, _F64P, _F64P, ctypes.c_int32, ctypes.c_float,
                                               ctypes.c_float, _F64P]),
            "imx_learn_reserve": (ctypes.c_int, [vp, ctypes.c_int32]),
    "imx_learn_store": (ctypes.c_int, [vp, _F64P, _F64P, _I32P]),
    "imx_learn_fit": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_uint64, _F64P]),
    "imx_learn_step": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_double, _I32P]),
            "imx_learn_last_rms": (ctypes.c_int, [vp, _F64P]),
            "imx_learn_set_limiter": (ctypes.c_int, [vp, ctypes.c_int32]),
            "imx_learn_destroy": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class ImxError(RuntimeError):
    """A libimx call returned a non-zero code."""


def check(rc, what=""):
    if rc != 0:
        msg = lib().imx_last_error().decode(errors="replace")
        raise ImxError(f"{what}: libimx error {rc}: {msg}")


def f64p(a):
    return a.ctypes.data_as(_F64P)


def i32p(a):
    return a.ctypes.data_as(_I32P)


def f32p(a):
    return a.ctypes.data_as(_F32P)
