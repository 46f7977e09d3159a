"""upc_flux_form_noalloc (the per-point form-factor flux on a file-scope
device scratch array, for callers that cannot allocate) against
upc_flux_form, both called through the C ABI: identical values and QAGS
statuses on two calls in a row, the second with fewer points, and sizes past
the stated maximum refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


def _points(n, seed):
    rng = np.random.default_rng(seed)
    # mostly inside 2R (QAGS), some beyond (point flux); photon energies of the
    # 1001 x 121 grid's rows, small to very large
    b = np.where(rng.random(n) < 0.85, rng.uniform(0.33, 13.36, n), rng.uniform(13.37, 60.0, n))
    k = 10 ** rng.uniform(-2.3, 3.6, n)
    return b, k


def _call(lib, fn, ctx, b, k):
    import torch
    n = b.numel()
    out = torch.full((n,), -7.0, dtype=torch.float64, device="cuda")
    ier = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(lib, fn)(ctx._ctx, C.c_void_p(b.data_ptr()), C.c_void_p(k.data_ptr()), n,
                          C.c_void_p(out.data_ptr()), C.c_void_p(ier.data_ptr()), stream)
    assert rc == 0, lib.upc_last_error().decode()
    torch.cuda.synchronize()
    return out.cpu().numpy(), ier.cpu().numpy()


def test_noalloc_matches_allocating_entry_point(gpu):
    import torch
    lib = gpu.load()
    ctx = gpu.Context(flux_point=0, breakup_mode=2, nm=3, ny=3)
    try:
        for n, seed in ((200_000, 1), (1_234, 2)):
            bn, kn = _points(n, seed)
            b = torch.tensor(bn, dtype=torch.float64, device="cuda")
            k = torch.tensor(kn, dtype=torch.float64, device="cuda")
            ref, ref_ier = _call(lib, "upc_flux_form", ctx, b, k)
            got, got_ier = _call(lib, "upc_flux_form_noalloc", ctx, b, k)
            print(f"n {n}: {int((bn <= 2 * 6.68).sum())} QAGS points, "
                  f"{int((ref_ier != 0).sum())} with ier != 0; max |diff| "
                  f"{np.nanmax(np.abs(got - ref)):.3e}")
            assert np.array_equal(got, ref, equal_nan=True)
            assert np.array_equal(got_ier, ref_ier)
    finally:
        ctx.close()


def test_noalloc_refuses_oversize_and_zero_length(gpu):
    import torch
    lib = gpu.load()
    ctx = gpu.Context(flux_point=0, breakup_mode=2, nm=3, ny=3)
    try:
        n = MAX_POINTS + 1
        b = torch.ones(n, dtype=torch.float64, device="cuda")
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        rc = lib.upc_flux_form_noalloc(ctx._ctx, C.c_void_p(b.data_ptr()), C.c_void_p(b.data_ptr()), n,
                                       C.c_void_p(out.data_ptr()), None, None)
        assert rc == -1 and b"at most" in lib.upc_last_error()
        assert lib.upc_flux_form_noalloc(ctx._ctx, None, None, 0, None, None, None) == 0
    finally:
        ctx.close()
