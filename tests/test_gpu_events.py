"""End-to-end event sampling on the full ditau workload: Pb-Pb 5.02 TeV XNXN,
form-factor flux, PROC_ID 15, sigma(Y, M) 1001 x 121 from the MFMA
luminosity and sigma(z, M) 1001 x 100; 10^8 (Y, M, z) draws from the FP64
prefix-sum CDFs with the counter-based Philox stream (seed 12345).

Bar: chi-square per projection (Y, M, z) against the probabilities of the
FP64-form tables (themselves within 1e-9 of the oracle) below the 0.9999
quantile.  At 10^8 draws a per-bin probability error of a few 1e-3 would
fail; the MFMA luminosity's 5e-6 does not show.
"""
import numpy as np
import pytest

import upcgen_amd as U

pytestmark = pytest.mark.gpu

CFG = dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0, flux_point=0, breakup_mode=2, nm=1001, ny=121, nz=100,
           proc_id=15)


def _chi2(h, p, n):
    keep = p * n > 50
    e = p[keep] * n
    return float(np.sum((h[keep] - e) ** 2 / e)), int(keep.sum())


def test_ditau_events_1e8(gpu_ctx):
    import torch
    from scipy.stats import chi2
    g = gpu_ctx(**CFG)
    # Four (Y, M) bins of this grid are NaN: QK21 nodes of their form-factor
    # integrals fall past the form-factor table's last knot, where the
    # reference's spline raises GSL_EDOM and its process stops.  A caller that
    # samples on regardless has to empty them; so does this test.
    # tables the draws use: MFMA luminosity
    Lm = g.lumi_grid(U.FORM_MFMA)
    assert int(torch.isnan(Lm).sum()) == 4
    cs_m, _, _ = g.nuc_cs_ym(torch.nan_to_num(Lm, nan=0.0))
    sum_ym = g.pdf_init(cs_m.reshape(1, -1))[0].contiguous()
    zt = g.cs_zm(0)
    sum_z = g.pdf_init(zt).contiguous()
    # expected probabilities: FP64 luminosity
    L64 = g.lumi_grid(U.FORM_FP64)
    cs64, _, _ = g.nuc_cs_ym(torch.nan_to_num(L64, nan=0.0))
    c = cs64.cpu().numpy()  # [ny][nm]
    py = c.sum(axis=1) / c.sum()
    pm = c.sum(axis=0) / c.sum()
    zr = zt.cpu().numpy()  # [nm][nz]
    pz = pm @ (zr / zr.sum(axis=1, keepdims=True))
    ny, nm, nz = CFG["ny"], CFG["nm"], CFG["nz"]
    hy = torch.zeros(ny, dtype=torch.float64, device="cuda")
    hm = torch.zeros(nm, dtype=torch.float64, device="cuda")
    hz = torch.zeros(nz, dtype=torch.float64, device="cuda")
    n_total, chunk = 100_000_000, 25_000_000
    for off in range(0, n_total, chunk):
        y, m, z = g.sample_events(sum_ym, sum_z=sum_z, n=chunk, seed=12345, offset=off)
        iy = ((y + 6.0) / (12.0 / ny)).floor().long().clamp(0, ny - 1)
        im = ((m - 3.56) / ((50.0 - 3.56) / nm)).floor().long().clamp(0, nm - 1)
        iz = ((z + 1.0) / (2.0 / nz)).floor().long().clamp(0, nz - 1)
        hy += torch.bincount(iy, minlength=ny).double()
        hm += torch.bincount(im, minlength=nm).double()
        hz += torch.bincount(iz, minlength=nz).double()
        del y, m, z, iy, im, iz
    for name, h, p in (("Y", hy, py), ("M", hm, pm), ("z", hz, pz)):
        stat, dof = _chi2(h.cpu().numpy(), p, n_total)
        print(f"{name}: chi2 {stat:.1f} over {dof} bins (0.9999 quantile {chi2.ppf(0.9999, dof - 1):.1f})")
        assert stat < chi2.ppf(0.9999, dof - 1), name
