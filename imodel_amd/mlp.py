"""Python host of the learned-correction MLP kernels in libimx (imx_mlp_*)."""
import ctypes

import numpy as np

from . import capi

IN, OUT = 48, 2


class Mlp:
    """48 -> 128 -> 128 -> 128 -> 2 GELU MLP on the GPU (BF16 MFMA, FP32 master)."""

    def __init__(self, seed=1):
        self._h = ctypes.c_void_p()
        capi.check(capi.lib().imx_mlp_create(int(seed), ctypes.byref(self._h)), "imx_mlp_create")
        self.nparams = capi.lib().imx_mlp_nparams()

    def close(self):
        if self._h:
            capi.lib().imx_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        capi.check(capi.lib().imx_mlp_get_params(self._h, capi.f32p(p)), "imx_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        capi.check(capi.lib().imx_mlp_set_params(self._h, capi.f32p(p)), "imx_mlp_set_params")

    def init(self, seed):
        """Parameters as Mlp(seed) starts, AdamW state cleared."""
        capi.check(capi.lib().imx_mlp_init(self._h, int(seed)), "imx_mlp_init")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], OUT), dtype=np.float32)
        capi.check(capi.lib().imx_mlp_forward(self._h, X.shape[0], capi.f32p(X), capi.f32p(Y)),
                   "imx_mlp_forward")
        return Y

    def train_step(self, X, T, w=None, lr=1e-3, wd=0.0, apply=True, gmask=0):
        """One AdamW step on L = sum w |y - t|^2; returns (loss, grad).

        Output o with bit o of gmask set takes T[:, o] as dL/dy itself."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        w = np.ascontiguousarray(np.ones(X.shape[0]) if w is None else w, dtype=np.float32)
        loss = ctypes.c_double()
        g = np.zeros(self.nparams, dtype=np.float32)
        capi.check(capi.lib().imx_mlp_train(self._h, X.shape[0], capi.f32p(X), capi.f32p(T), capi.f32p(w),
                                            float(lr), float(wd), int(bool(apply)), ctypes.byref(loss),
                                            capi.f32p(g), int(gmask)), "imx_mlp_train")
        return loss.value, g

    def forward_timed(self, n, reps=20):
        ms = ctypes.c_double()
        capi.check(capi.lib().imx_mlp_forward_timed(self._h, int(n), int(reps), ctypes.byref(ms)),
                   "imx_mlp_forward_timed")
        return ms.value
