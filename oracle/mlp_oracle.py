"""TEST INFRASTRUCTURE ONLY: FP64 numpy reference of the learned-correction
MLP (48 -> 128 -> 128 -> 128 -> 2, activation GELU in its tanh form,
g(x) = x sigma(2 c (x + 0.044715 x^3)), c = sqrt(2/pi)), its loss
L = sum_e w_e |y_e - t_e|^2 and its parameter gradients, with optional
emulation of the BF16 storage the GPU kernels use (weights, layer inputs and
pre-activation gradients rounded to BF16, round-to-nearest-even).
Pinned by central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np

IN, H, OUT = 48, 128, 2
SHAPES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
          ("b3", (H,)), ("W4", (H, OUT)), ("b4", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, s in SHAPES:
        k = int(np.prod(s))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(s)
        o += k
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


# GELU, tanh form (Hendrycks & Gimpel 2016): x/2 (1 + tanh(c (x + a x^3)))
# = x sigma(2 c (x + a x^3)), c = sqrt(2/pi), a = 0.044715
GELU_C2 = 2.0 * np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def _s(x):
    t = GELU_C2 * (x + GELU_A * x ** 3)
    return 0.5 * (1.0 + np.tanh(0.5 * t))  # sigma(t), overflow-free


def gelu(x):
    return x * _s(x)


def gelu_d(x):
    s = _s(x)
    return s + x * s * (1.0 - s) * GELU_C2 * (1.0 + 3.0 * GELU_A * x * x)


def forward(p, X, emulate_bf16=False):
    P = unpack(p)
    r = bf16 if emulate_bf16 else (lambda a: a)
    a0 = r(np.asarray(X, dtype=np.float64))
    z1 = a0 @ r(P["W1"]) + P["b1"]
    a1 = r(gelu(z1))
    z2 = a1 @ r(P["W2"]) + P["b2"]
    a2 = r(gelu(z2))
    z3 = a2 @ r(P["W3"]) + P["b3"]
    a3 = r(gelu(z3))
    y = a3 @ r(P["W4"]) + P["b4"]
    return y, (a0, z1, a1, z2, a2, z3, a3)


def loss_and_grad(p, X, T, w, emulate_bf16=False):
    P = unpack(p)
    r = bf16 if emulate_bf16 else (lambda a: a)
    y, (a0, z1, a1, z2, a2, z3, a3) = forward(p, X, emulate_bf16)
    res = y - T
    L = float(np.sum(w[:, None] * res ** 2))
    dz4 = r(2.0 * w[:, None] * res)
    g = {"W4": a3.T @ dz4, "b4": dz4.sum(0)}
    dz3 = r((dz4 @ r(P["W4"]).T) * gelu_d(z3))
    g["W3"], g["b3"] = a2.T @ dz3, dz3.sum(0)
    dz2 = r((dz3 @ r(P["W3"]).T) * gelu_d(z2))
    g["W2"], g["b2"] = a1.T @ dz2, dz2.sum(0)
    dz1 = r((dz2 @ r(P["W2"]).T) * gelu_d(z1))
    g["W1"], g["b1"] = a0.T @ dz1, dz1.sum(0)
    return L, pack(g)
