"""CPU reference of the luminosity surrogate network.  TEST INFRASTRUCTURE
ONLY: the checker for the HIP training kernels' forward pass and gradients.

The network has no counterpart in upcgen (the reference integrates every
L(M, Y)); this restates the library's own definition:
  inputs [B][32] -> 3 x (Linear 128 + SiLU) -> Linear 1
Parameter vector layout (float32): W1[128][32], b1[128], W2[128][128], b2,
W3[128][128], b3, W4[128], b4.  Loss: mean((pred - t)^2).

With bf16_ops=True the products round their operands to bf16 (round to
nearest even) exactly where the GPU kernels do: both operands of every
hidden-layer product (forward, dW and dH products) and W4 in the output
layer.  With split=True (the library's default precision) every
hidden-layer product is hi.hi + hi.lo + lo.hi, hi = bf16(x) and lo =
bf16(x - hi) of each FP32 operand, and W4 is used unrounded.
"""
import numpy as np

IN, H = 32, 128
OFF = {}
_o = 0
for _name, _n in (("W1", H * IN), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H), ("W4", H),
                  ("b4", 1)):
    OFF[_name] = (_o, _o + _n)
    _o += _n
NPARAMS = _o


def bf16(x):
    """Round float32 values to bf16 (nearest even) and back to float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = {k: p[a:b] for k, (a, b) in OFF.items()}
    g["W1"] = g["W1"].reshape(H, IN)
    g["W2"] = g["W2"].reshape(H, H)
    g["W3"] = g["W3"].reshape(H, H)
    return g


def _r(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, dtype=np.float64)


def _hl(x):
    x = np.asarray(x, dtype=np.float32)
    hi = bf16(x)
    lo = bf16(x - hi)  # exact in float32: hi is x to 8 significant bits
    return hi.astype(np.float64), lo.astype(np.float64)


def _mm(A, B, bf16_ops, split):
    """A @ B with the GPU's operand rounding (A, B hold FP32 values)."""
    if split:
        ah, al = _hl(A)
        bh, bl = _hl(B)
        return ah @ bh + ah @ bl + al @ bh
    return _r(A, bf16_ops) @ _r(B, bf16_ops)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, bf16_ops=True, split=False):
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    m = lambda a, b: _mm(a, b, bf16_ops, split)
    Z1 = m(X, w["W1"].T) + w["b1"]
    H1 = _silu(Z1)
    Z2 = m(H1, w["W2"].T) + w["b2"]
    H2 = _silu(Z2)
    Z3 = m(H2, w["W3"].T) + w["b3"]
    H3 = _silu(Z3)
    pred = H3 @ _r(w["W4"], bf16_ops and not split) + w["b4"][0]
    return pred, (X, Z1, H1, Z2, H2, Z3, H3)


def loss(p, X, T, bf16_ops=True, split=False):
    pred, _ = forward(p, X, bf16_ops, split)
    return float(np.mean((pred - np.asarray(T, dtype=np.float64)) ** 2))


def grad(p, X, T, bf16_ops=True, split=False):
    """Gradient of mean((pred - T)^2) as a flat float64 vector."""
    w = unpack(p)
    m = lambda a, b: _mm(a, b, bf16_ops, split)
    pred, (X, Z1, H1, Z2, H2, Z3, H3) = forward(p, X, bf16_ops, split)
    B = X.shape[0]
    gout = 2.0 * (pred - np.asarray(T, dtype=np.float64)) / B
    g = np.zeros(NPARAMS)

    def put(name, v):
        a, b = OFF[name]
        g[a:b] = np.ravel(v)

    put("W4", H3.T @ gout)
    put("b4", [gout.sum()])
    dZ3 = gout[:, None] * _r(w["W4"], bf16_ops and not split)[None, :] * _dsilu(Z3)
    put("W3", m(dZ3.T, H2))
    put("b3", dZ3.sum(axis=0))
    dZ2 = m(dZ3, w["W3"]) * _dsilu(Z2)
    put("W2", m(dZ2.T, H1))
    put("b2", dZ2.sum(axis=0))
    dZ1 = m(dZ2, w["W2"]) * _dsilu(Z1)
    put("W1", m(dZ1.T, X))
    put("b1", dZ1.sum(axis=0))
    return g
