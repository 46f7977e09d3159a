"""GPU parity of the HIP library (through the C ABI) against the CPU oracle.

Bar: FP64 forms within 1e-9 relative per (M, Y) bin (the stated tolerance of
the plain form); tables and per-point functions within the tolerances written
at each assertion.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu

PB = dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0)


def _t(x):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda")


def _np(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_tables_match_oracle(gpu_ctx, oracle_ctx):
    g = gpu_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    o = oracle_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    inf = g.info
    assert inf.rho0 == pytest.approx(o.rho0, rel=1e-14)
    assert inf.csNN == pytest.approx(o.cs_nn, rel=1e-14)
    assert inf.factor == o.factor and inf.g1 == o.g1
    assert inf.p20 == pytest.approx(o.breakup_spline(20.0), rel=1e-12)
    rng = np.random.default_rng(7)
    b = np.concatenate([rng.uniform(0, 20, 400), [0.0, 20.0, 13.36, 19.999999]])
    gaa = _np(g.nuclear_overlap(_t(b), 0))
    ta = _np(g.nuclear_overlap(_t(b), 1))
    og = np.array([o.gaa(x) for x in b])
    ot = np.array([o.ta(x) for x in b])
    assert np.max(np.abs(gaa - og)) < 1e-13
    assert rel(ta, ot) < 1e-12
    q2 = np.concatenate([10 ** rng.uniform(-9, np.log10(1.99), 500), [1e-9, 1.999998]])
    # calcFormFac cancels like 1/(QR)^2 as Q -> 0: one-ulp differences of sin,
    # cos, sinh, cosh between libraries grow to ~1e-15/(QR)^2 relative there
    qr2 = q2 / 0.1973269718 ** 2 * 6.68 ** 2
    tol = 5e-14 + 2e-15 / qr2
    ff_an = _np(g.form_factor(_t(q2), False))
    ref_an = np.array([o.formfac(x) for x in q2])
    assert np.all(np.abs(ff_an - ref_an) <= tol * 208)
    ff_sp = _np(g.form_factor(_t(q2), True))
    ref_sp = np.array([o.formfac_spline(x) for x in q2])
    # the spline carries the noise of its interval's knots (2e-6 apart): the
    # whole first interval interpolates from the knot at Q2 = 1e-9
    dq2 = (2.0 - 1e-9) / 1e6
    q2k = 1e-9 + np.floor((q2 - 1e-9) / dq2) * dq2
    tol_sp = 20 * (5e-14 + 2e-15 / (q2k / 0.1973269718 ** 2 * 6.68 ** 2))
    assert np.all(np.abs(ff_sp - ref_sp) <= tol_sp * 208)
    print(f"form factor max abs diff: analytic {np.max(np.abs(ff_an - ref_an)):.2e}, "
          f"spline {np.max(np.abs(ff_sp - ref_sp)):.2e}")
    bb = np.concatenate([10 ** rng.uniform(-5, np.log10(20), 300), [20.0, 1e-6]])
    for mode in (2, 3, 4):
        assert rel(_np(g.breakup_prob(_t(bb), mode)), [o.breakup_prob(x, mode) for x in bb]) < 1e-12
    assert rel(_np(g.breakup_prob(_t(bb), tabulated=True)), [o.breakup_spline(x) for x in bb]) < 1e-11


def test_flux_matches_oracle(gpu_ctx, oracle_ctx):
    g = gpu_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    o = oracle_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    rng = np.random.default_rng(11)
    b = 10 ** rng.uniform(-0.5, 5, 600)
    k = 10 ** rng.uniform(-2.4, 4, 600)
    pt = _np(g.flux_point(_t(b), _t(k)))
    opt = np.array([o.flux_point(x, y) for x, y in zip(b, k)])
    nz = opt > 1e-250
    assert rel(pt[nz], opt[nz]) < 1e-13
    # form-factor flux: QAGS inside 2R, point flux outside
    bf = np.concatenate([rng.uniform(0.3, 13.36, 400), rng.uniform(13.37, 40, 50)])
    kf = 10 ** rng.uniform(-2.4, 4, 450)
    ff, ier = g.flux_form(_t(bf), _t(kf))
    ff, ier = _np(ff), _np(ier)
    oref = [o.flux_form(x, y) for x, y in zip(bf, kf)]
    off = np.array([v for v, _, _ in oref])
    assert np.all(ier == 0) and all(e == 0 for _, _, e in oref)
    assert rel(ff, off) < 1e-11


def _lumi_check(L, ref, tol=1e-9):
    r = rel(L, ref)
    print(f"max rel err {r:.3e}")
    assert np.all(np.isfinite(L))
    assert r <= tol
    return r


def test_lumi_grid_config1_full(gpu_ctx):
    gold = np.load(os.path.join(GOLD, "oracle_grids.npz"))
    g = gpu_ctx(flux_point=1, breakup_mode=1, nm=101, ny=41, nz=50, proc_id=13)
    L = _np(g.lumi_grid())
    _lumi_check(L, gold["cfg1_point_mode1"])


def test_lumi_grid_formfactor_fixtures(gpu_ctx):
    gold = np.load(os.path.join(GOLD, "oracle_grids.npz"))
    g2 = gpu_ctx(flux_point=0, breakup_mode=2, nm=7, ny=5)
    _lumi_check(_np(g2.lumi_grid()), gold["ff_mode2_7x5"])
    assert g2.info.qags_bad == 0 and g2.info.qags_calls > 0
    g3 = gpu_ctx(flux_point=0, breakup_mode=3, use_polarized=1, nm=7, ny=5)
    ls, lp = g3.lumi_grid()
    _lumi_check(_np(ls), gold["ff_mode3_pol_s_7x5"])
    _lumi_check(_np(lp), gold["ff_mode3_pol_p_7x5"])


@pytest.mark.parametrize("kw", [
    dict(PB, breakup_mode=2),
    dict(PB, breakup_mode=1),
    dict(PB, breakup_mode=4),
    dict(Z=54, A=129, R=5.4889, a=0.54, sqrts=5440.0, breakup_mode=3),
    dict(Z=79, A=197, R=6.3691, a=0.54, sqrts=200.0, breakup_mode=2),
    dict(Z=8, A=16, R=2.4809, a=0.54, sqrts=7000.0, breakup_mode=2),
])
def test_contraction_skip_radius(gpu_ctx, kw):
    """The FP64 contraction takes P G as 0 below info.bskip, where the G_AA
    spline must be below p20 2^-70 in magnitude on every interval the
    estimate can name (knots approached from both sides included); and the
    radius must not be vacuous (at least half of 2R, where the hadronic
    overlap shuts G off)."""
    g = gpu_ctx(flux_point=0, nm=3, ny=3, **kw)
    inf = g.info
    bs, p20 = inf.bskip, inf.p20
    print(f"bskip {bs:.4f} fm, p20 {p20:.4g}")
    assert 0.5 * 2 * kw["R"] <= bs < 20.0
    db = 20.0 / 199
    knots = np.arange(int(bs / db) + 1) * db
    b = np.concatenate([np.linspace(0.0, bs, 40001, endpoint=False), knots * (1 - 4e-16), knots,
                        knots * (1 + 4e-16)])
    b = b[(b >= 0) & (b < bs)]
    G = _np(g.nuclear_overlap(_t(b), 0))
    assert np.max(np.abs(G)) < p20 * 2.0 ** -70


@pytest.mark.parametrize("kw", [
    dict(flux_point=1, breakup_mode=4, use_polarized=0),
    dict(flux_point=0, breakup_mode=1, use_polarized=1),
    dict(flux_point=1, breakup_mode=2, use_polarized=1),
    dict(Z=54, A=129, R=5.4889, a=0.54, sqrts=5440.0, flux_point=0, breakup_mode=3, use_polarized=0),
    dict(Z=79, A=197, R=6.3691, a=0.54, sqrts=200.0, flux_point=1, breakup_mode=2, use_polarized=1,
         mmin=1.0, mmax=10.0, ymin=-3.0, ymax=3.0),
])
def test_lumi_grid_modes(gpu_ctx, oracle_ctx, kw):
    kw = dict(kw)
    kw.setdefault("nm", 6)
    kw.setdefault("ny", 5)
    g = gpu_ctx(**kw)
    o = oracle_ctx(**kw)
    ref, nbad = o.lumi_grid()
    assert nbad == 0
    out = g.lumi_grid()
    if kw.get("use_polarized"):
        _lumi_check(_np(out[0]), ref[0])
        _lumi_check(_np(out[1]), ref[1])
    else:
        _lumi_check(_np(out), ref)


@pytest.mark.parametrize("nm,ny,pol", [(3, 4, 0), (2, 2, 1), (3, 1, 0), (1, 6, 1)])
def test_lumi_grid_even_and_single(gpu, gpu_ctx, oracle_ctx, nm, ny, pol):
    """Grids where a Y bin is its own mirror (even ny: iy = ny / 2 shares its
    k2 row with itself), a single Y bin and a single M bin: the FP64 form
    against the oracle at 1e-9, the MFMA forms at 1e-5."""
    kw = dict(flux_point=0, breakup_mode=2, use_polarized=pol, nm=nm, ny=ny)
    g = gpu_ctx(**kw)
    o = oracle_ctx(**kw)
    ref, nbad = o.lumi_grid()
    assert nbad == 0
    refs = ref if pol else (ref,)
    out = g.lumi_grid()
    for a, r in zip(out if pol else (out,), refs):
        _lumi_check(_np(a), r)
    for form in (gpu.FORM_MFMA, gpu.FORM_MFMA_FP32):
        m = g.lumi_grid(form)
        for a, r in zip(m if pol else (m,), refs):
            assert _lumi_check(_np(a), r, tol=1e-5) <= 1e-5


def test_zero_length_point_calls(gpu_ctx):
    """n = 0 is a valid call for every per-point entry point: nothing is
    launched and nothing fails."""
    g = gpu_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    e = _t(np.zeros(0))
    assert g.lumi_points(e, e).numel() == 0
    f, ier = g.flux_form(e, e)
    assert f.numel() == 0 and ier.numel() == 0
    assert g.flux_point(e, e).numel() == 0
    assert g.photon_flux(e, e).numel() == 0
    assert g.nuclear_overlap(e).numel() == 0
    assert g.breakup_prob(e).numel() == 0
    assert g.form_factor(e).numel() == 0


def test_lumi_points_arbitrary(gpu_ctx, oracle_ctx):
    g = gpu_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    o = oracle_ctx(flux_point=0, breakup_mode=2, nm=3, ny=3)
    rng = np.random.default_rng(3)
    M = rng.uniform(3.56, 50, 12)
    Y = rng.uniform(-6, 6, 12)
    L = _np(g.lumi_points(_t(M), _t(Y)))
    _lumi_check(L, [o.lumi(m, y) for m, y in zip(M, Y)])


def test_cross_section_tables(gpu_ctx, oracle_ctx):
    import oracle as O
    for pol in (0, 1):
        kw = dict(flux_point=1, breakup_mode=1, use_polarized=pol, nm=23, ny=7, nz=30, proc_id=15)
        g, o = gpu_ctx(**kw), oracle_ctx(**kw)
        flags = (1, 2) if pol else (0,)
        for f in flags:
            assert rel(_np(g.cs_zm(f)), o.cs_zm(f)) < 1e-12
        lumi = g.lumi_grid()
        if pol:
            cs, ratio, tot = g.nuc_cs_ym(lumi)
            ocs, oratio, otot = o.nuc_cs_ym(lumi_s=_np(lumi[0]), lumi_p=_np(lumi[1]))
            assert rel(_np(ratio), oratio) < 1e-12
        else:
            cs, _, tot = g.nuc_cs_ym(lumi)
            ocs, _, otot = o.nuc_cs_ym(lumi=_np(lumi))
        assert rel(_np(cs), ocs) < 1e-13
        assert tot == pytest.approx(otot, rel=1e-12)
        # sampler tables: gsl_histogram_pdf_init semantics row by row
        sums = _np(g.pdf_init(cs.reshape(1, -1)))[0]
        osum, st = O.pdf_init(_np(cs))
        assert st == 0 and np.max(np.abs(sums - osum)) < 1e-14
        zt = g.cs_zm(flags[0])
        zs = _np(g.pdf_init(zt))
        for r in range(zt.shape[0]):
            orow, _ = O.pdf_init(_np(zt[r]))
            assert np.max(np.abs(zs[r] - orow)) < 1e-14


def test_pdf_rejects_negative_bins(gpu):
    import torch
    with pytest.raises(gpu.UpcError):
        gpu.Context.pdf_init(torch.tensor([1.0, -1.0, 2.0], dtype=torch.float64, device="cuda"))


def test_sampling_matches_tables(gpu_ctx):
    """Drawn (Y, M, z) follow the tabulated densities (chi-square per projection)."""
    from scipy.stats import chi2
    kw = dict(flux_point=1, breakup_mode=1, nm=40, ny=21, nz=20, proc_id=15)
    g = gpu_ctx(**kw)
    lumi = g.lumi_grid()
    cs, _, _ = g.nuc_cs_ym(lumi)
    sum_ym = g.pdf_init(cs.reshape(1, -1))[0].contiguous()
    zt = g.cs_zm(0)
    sum_z = g.pdf_init(zt).contiguous()
    n = 2_000_000
    y, m, z = g.sample_events(sum_ym, sum_z=sum_z, n=n, seed=12345)
    y, m, z = _np(y), _np(m), _np(z)
    assert np.all(np.isfinite(y)) and np.all(np.isfinite(m)) and np.all(np.isfinite(z))
    csn = _np(cs)
    py = csn.sum(axis=1) / csn.sum()
    hy, _ = np.histogram(y, bins=21, range=(-6, 6))
    e = py * n
    stat = np.sum((hy - e) ** 2 / e)
    assert stat < chi2.ppf(0.999, 20)
    pm = csn.sum(axis=0) / csn.sum()
    hm, _ = np.histogram(m, bins=40, range=(3.56, 50))
    e = pm * n
    stat = np.sum((hm - e) ** 2 / e)
    assert stat < chi2.ppf(0.999, 39)


@pytest.mark.parametrize("kw", [
    dict(flux_point=0, breakup_mode=2),
    dict(flux_point=1, breakup_mode=1),
    dict(flux_point=0, breakup_mode=4),
    dict(Z=54, A=129, R=5.4889, a=0.54, sqrts=5440.0, flux_point=0, breakup_mode=3),
])
def test_vm_photon_flux(gpu_ctx, oracle_ctx, kw):
    """calcPhotonFlux (vector-meson photon flux) at J/psi and Upsilon masses."""
    kw = dict(kw, nm=3, ny=9, ymin=-4.0, ymax=4.0)
    g = gpu_ctx(**kw)
    o = oracle_ctx(**kw)
    M = np.array([3.0969] * 11 + [9.3987] * 11)
    Y = np.tile(np.linspace(-5, 5, 11), 2)
    got = _np(g.photon_flux(_t(M), _t(Y)))
    ref = np.array([o.photon_flux(m, y) for m, y in zip(M, Y)])
    assert np.all(ref > 0)
    assert rel(got, ref) < 1e-9
    rng = np.random.default_rng(8)
    cs_pos, cs_neg = rng.uniform(0.5, 2, 9), rng.uniform(0.5, 2, 9)
    cs, ratio, tot = g.nuc_cs_y(3.0969, _t(cs_pos), _t(cs_neg))
    rcs, rratio, rtot = o.nuc_cs_y(3.0969, cs_pos, cs_neg)
    assert rel(_np(cs), rcs) < 1e-9 and rel(_np(ratio), rratio) < 1e-9
    assert tot == pytest.approx(rtot, rel=1e-9)
