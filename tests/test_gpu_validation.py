"""Argument checks that must fire before any kernel is launched: the C ABI
takes plain device pointers and sizes, so a narrower dtype, a shorter paired
input or a smaller output buffer would otherwise be read or written past its
end on the device.  None of these calls launches a kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu_ctx):
    return gpu_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)


def test_rejects_float32_inputs(gpu, ctx):
    import torch
    b = torch.ones(8, dtype=torch.float32, device="cuda")
    k = torch.ones(8, dtype=torch.float64, device="cuda")
    for call in (lambda: ctx.flux_point(b, k), lambda: ctx.flux_form(b, k), lambda: ctx.lumi_points(b, k),
                 lambda: ctx.photon_flux(b, k), lambda: ctx.breakup_prob(b), lambda: ctx.form_factor(b),
                 lambda: ctx.nuclear_overlap(b)):
        with pytest.raises(gpu.UpcError, match="dtype"):
            call()


def test_rejects_unequal_pairs(gpu, ctx):
    import torch
    a = torch.ones(8, dtype=torch.float64, device="cuda")
    short = torch.ones(5, dtype=torch.float64, device="cuda")
    for call in (lambda: ctx.flux_point(a, short), lambda: ctx.flux_form(a, short),
                 lambda: ctx.lumi_points(a, short), lambda: ctx.photon_flux(a, short)):
        with pytest.raises(gpu.UpcError, match="same number"):
            call()


def test_rejects_small_or_host_outputs(gpu, ctx):
    import torch
    small = torch.empty((3, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(gpu.UpcError, match="at least"):
        ctx.lumi_grid(out=small)
    with pytest.raises(gpu.UpcError, match="device tensor"):
        ctx.lumi_grid(out=torch.empty((3, 3), dtype=torch.float64))
    with pytest.raises(gpu.UpcError, match="at least"):
        ctx.nuc_cs_ym(torch.ones(4, dtype=torch.float64, device="cuda"))
    with pytest.raises(gpu.UpcError, match="at least"):
        ctx.nuc_cs_y(3.1, torch.ones(2, dtype=torch.float64, device="cuda"),
                     torch.ones(3, dtype=torch.float64, device="cuda"))


def test_rejects_short_sampler_tables(gpu, ctx):
    import torch
    L = ctx.lumi_grid()
    cs, _, _ = ctx.nuc_cs_ym(L)
    sum_ym = ctx.pdf_init(cs.reshape(1, -1))[0].contiguous()
    sum_z = ctx.pdf_init(ctx.cs_zm(0)).contiguous()
    with pytest.raises(gpu.UpcError, match="sum_z"):
        ctx.sample_events(sum_ym, sum_z=sum_z[:2], n=10)
    with pytest.raises(gpu.UpcError, match="sum_ym"):
        ctx.sample_events(sum_ym[:4], sum_z=sum_z, n=10)
    with pytest.raises(gpu.UpcError, match="needs sum_z"):
        ctx.sample_events(sum_ym, n=10)


def test_grid_beyond_2_pow_24_flux_rows(gpu, oracle_ctx):
    """An asymmetric Y range gives every point its own k2 row: 4096 x 2049
    points have 16,785,408 flux rows, past 2^24 where row * 128 + node no
    longer fits 31 bits.  Task records, row offsets and flux slots are 64-bit;
    with the point-charge flux (no QAGS tasks, 16 GB of flux rows) the grid's
    last rows must still match the oracle."""
    import torch
    nm, ny, ymin, ymax = 4096, 2049, -5.0, 6.0
    c = gpu.Context(flux_point=1, breakup_mode=1, nm=nm, ny=ny, ymin=ymin, ymax=ymax)
    try:
        L = c.lumi_grid()
        torch.cuda.synchronize()
        o = oracle_ctx(flux_point=1, breakup_mode=1, nm=3, ny=3)
        dm, dy = (50 - 3.56) / nm, (ymax - ymin) / ny
        for im, iy in ((nm - 1, ny - 1), (nm - 1, 0), (nm // 2, ny // 3), (0, ny - 1)):
            ref = o.lumi(3.56 + dm * im, ymin + dy * iy) * dm * dy
            got = float(L[im, iy])
            assert abs(got / ref - 1) < 1e-9, (im, iy, got, ref)
    finally:
        c.close()
