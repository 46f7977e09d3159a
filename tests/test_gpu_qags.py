"""The two QAGS kernels of the form-factor flux run the same dqagse
arithmetic (QUADPACK; gsl_integration_qags at epsabs = epsrel = 1e-4, limit
1000, src/UpcCrossSection.cpp:209).  The fast kernel holds up to 16 intervals
per lane, with interval end points packed as dyadic (level, index) pairs of
[0, 10] and the order list in 4-bit fields; the other keeps the reference's
1000-interval FP64 workspace in global memory.  Lowering the fast kernel's
cap hands integrals to the second kernel part-way, so every cap must give
bit-identical results, and both must match the oracle's QAGS."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(x):
    import torch
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device="cuda")


def test_overflow_kernel_is_bit_identical(gpu, oracle_ctx):
    o = oracle_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    rng = np.random.default_rng(23)
    b = rng.uniform(0.3, 13.36, 3000)
    k = 10 ** rng.uniform(-2.4, 3.0, 3000)
    with gpu.Context(flux_point=0, breakup_mode=2, nm=3, ny=3) as g:
        ref = None
        for cap in (16, 8, 3, 1):
            g.set_qags_cap(cap)
            f, ier = g.flux_form(_t(b), _t(k))
            f, ier = f.cpu().numpy(), ier.cpu().numpy()
            calls, evals = g.info.qags_calls, g.info.qags_evals
            print(f"cap {cap}: {calls} integrals, {evals} evaluations")
            if ref is None:
                ref = (f, ier, evals)
                assert np.all(ier == 0)
                sel = rng.choice(len(b), 150, replace=False)
                want = np.array([o.flux_form(b[i], k[i])[0] for i in sel])
                assert np.max(np.abs(f[sel] - want) / np.abs(want)) < 1e-11
            else:
                assert np.array_equal(f.view(np.uint64), ref[0].view(np.uint64)), cap
                assert np.array_equal(ier, ref[1]), cap
                assert evals == ref[2], cap
        g.set_qags_cap(16)


def test_overflow_kernel_grid(gpu):
    kw = dict(flux_point=0, breakup_mode=2, nm=21, ny=11)
    with gpu.Context(**kw) as g:
        L16 = g.lumi_grid().cpu().numpy()
        g.set_qags_cap(2)
        L2 = g.lumi_grid().cpu().numpy()
    assert np.array_equal(L16.view(np.uint64), L2.view(np.uint64))


def test_qags_cap_range(gpu):
    with gpu.Context(flux_point=0, breakup_mode=2, nm=3, ny=3) as g:
        for bad in (0, 17):
            with pytest.raises(gpu.UpcError):
                g.set_qags_cap(bad)


def test_device_j1_vs_mpmath(gpu):
    """The device J1 (Chebyshev series below x = 8, Hankel modulus/phase
    above, coefficients from gen_coeffs.py) against 50-digit mpmath values
    over the integrand's range x = b q / hbar c <= 700."""
    kat = np.load(os.path.join(os.path.dirname(__file__), "golden", "bessel_kat.npz"))
    x, want = kat["xj"], kat["j1"]
    got = gpu.bessel_j1(x).cpu().numpy()
    env = np.where(x < 2, 1.0, np.sqrt(2 / (np.pi * np.maximum(x, 1e-300))))
    err = np.abs(got - want) / env
    print(f"max |J1 - mpmath| / envelope = {err.max():.2e} at x = {x[err.argmax()]:.4g}")
    assert err.max() <= 5e-16
