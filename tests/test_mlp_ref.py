"""The CPU reference of the surrogate network, pinned by finite differences
(float64, no rounding) and by known bf16 roundings."""
import numpy as np

from oracle import mlp_ref as R


def _params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.15, R.NPARAMS)
    a, b = R.OFF["W1"]
    W1 = p[a:b].reshape(R.H, R.IN)
    W1[:, 11:] = 0.0
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(1)
    X = np.zeros((24, R.IN))
    X[:, :11] = rng.normal(0, 1, (24, 11))
    T = rng.normal(0, 1, 24)
    g = R.grad(p, X, T, bf16_ops=False)
    idx = np.concatenate([rng.integers(a, b, 6) for a, b in R.OFF.values()])
    for i in idx:
        if R.OFF["W1"][0] <= i < R.OFF["W1"][1] and (i - R.OFF["W1"][0]) % R.IN >= 11:
            continue
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss(pp, X, T, False) - R.loss(pm, X, T, False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.0e-3, -2.5, 1 + 2 ** -8], dtype=np.float32)
    r = R.bf16(x)
    # 1 + 2^-8 is the midpoint between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[5] == 1.0
    assert r[1] == np.float32(1.0078125) or r[1] == np.float32(1.0)
    assert r[4] == -2.5
    assert abs(r[3] / 3.0e-3 - 1) < 2 ** -8
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)
