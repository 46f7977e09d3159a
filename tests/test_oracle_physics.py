"""Physics pins of the oracle (CPU): closed forms and invariants that the
reference's arithmetic must satisfy, chosen so that a dropped term, a wrong
sign/index or a missing normalisation fails one of them."""
import math
import os

import numpy as np
import pytest
import scipy.integrate as si

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
HC = 0.1973269718
ALPHA = 1.0 / 137.035999074


def test_ws_rho0_matches_fermi_integral(oracle_ctx):
    o = oracle_ctx(flux_point=1, breakup_mode=1, nm=2, ny=2)
    R, a, A = 6.68, 0.447, 208
    # int_0^inf r^2/(1+e^{(r-R)/a}) dr = R^3/3 (1 + pi^2 a^2/R^2) + O(e^{-R/a})
    fermi = R ** 3 / 3 * (1 + (math.pi * a / R) ** 2)
    assert o.rho0 == pytest.approx(A / (4 * math.pi * fermi), rel=2e-6)


def test_thickness_at_zero_is_twice_the_z_integral(oracle_ctx):
    o = oracle_ctx(flux_point=1, breakup_mode=1, nm=2, ny=2)
    R, a = 6.68, 0.447
    # T_A(0) = 2 rho0 int_0^inf dz/(1+e^{(z-R)/a}) = 2 rho0 a ln(1 + e^{R/a})
    assert o.ta(0.0) == pytest.approx(2 * o.rho0 * a * math.log1p(math.exp(R / a)), rel=1e-5)


def test_thickness_and_overlap_normalisation(oracle_ctx):
    o = oracle_ctx(flux_point=1, breakup_mode=1, nm=2, ny=2)
    b = np.linspace(0, 20, 8001)
    ta = np.array([o.ta(x) for x in b])
    # int d^2b T_A(b) = A
    assert np.trapezoid(2 * np.pi * b * ta, b) == pytest.approx(208, rel=2e-4)
    g = np.array([o.gaa(x) for x in b])
    assert g[0] == 0.0 and abs(g[-1] - 1) < 2e-4 and np.all(np.diff(g) >= -1e-12)
    # -ln G(b) / sigma_NN is twice the overlap T_AA(b) = int d^2s T_A(s) T_A(|b - s|):
    # the reference's radial integrand is 2 s T_A(s) while its phi sum already
    # spans 0..2pi (src/UpcCrossSection.cpp:397-405), a factor kept as written
    from scipy.interpolate import CubicSpline
    tas = CubicSpline(b, ta)
    for bb in (14.0, 16.0, 18.0):
        taa = si.dblquad(lambda phi, s: s * tas(s) * tas(min(20.0, math.sqrt(bb * bb + s * s - 2 * bb * s * math.cos(phi)))),
                         0, 20, 0, 2 * math.pi, epsabs=1e-10, epsrel=1e-6)[0]
        assert -math.log(o.gaa(bb)) / o.cs_nn == pytest.approx(2 * taa, rel=0.05), bb


def test_form_factor_is_woods_saxon_transform(oracle_ctx):
    # calcFormFac = symmetrised-Fermi transform + the first e^{-R/a} Woods-Saxon
    # correction term: the Woods-Saxon transform up to e^{-2R/a} ~ 1e-13
    o = oracle_ctx(flux_point=0, breakup_mode=1, nm=2, ny=2)
    kat = np.load(os.path.join(GOLD, "formfac_kat.npz"))
    for q2, f in zip(kat["q2"], kat["f_over_rho0"]):
        assert o.formfac(q2) == pytest.approx(o.rho0 * f, rel=1e-11, abs=1e-13), q2
    assert o.formfac(1e-9) / 208 == pytest.approx(1.0, abs=2e-7)
    # the 10^6-knot spline reproduces the analytic form factor
    for q2 in (1e-3, 0.05, 0.7, 1.99):
        assert o.formfac_spline(q2) == pytest.approx(o.formfac(q2), rel=1e-10, abs=1e-12)
    # within a few knots of Q2min the natural end condition (f'' = 0) shows
    d = abs(o.formfac_spline(3.3e-6) / o.formfac(3.3e-6) - 1)
    assert 1e-11 < d < 1e-8


def test_breakup_modes_partition_unity(oracle_ctx):
    o = oracle_ctx(flux_point=0, breakup_mode=2, nm=2, ny=2)
    for b in (0.5, 5, 13.36, 19.9, 20, 35, 300):
        p2, p3, p4 = (o.breakup_prob(b, m) for m in (2, 3, 4))
        assert p2 + p3 + p4 == pytest.approx(1.0, abs=1e-14)
        assert 0 <= p2 <= 1 and 0 <= p3 <= 1 and 0 <= p4 <= 1
    assert o.breakup_prob(1e-3, 2) == pytest.approx(1.0, abs=1e-12)
    assert o.breakup_prob(900, 3) == pytest.approx(1.0, abs=1e-3)
    assert o.breakup_prob(900, 3) < o.breakup_prob(1000, 3) < 1.0
    # monotone: XNXN falls with b
    ps = [o.breakup_prob(b, 2) for b in np.linspace(10, 60, 26)]
    assert all(x > y for x, y in zip(ps, ps[1:]))
    # the spline reproduces the direct sum at and between knots
    for b in (12.3456, 17.0005, 19.9999, 20.0):
        assert o.breakup_spline(b) == pytest.approx(o.breakup_prob(b, 2), rel=1e-9)


def test_point_flux_is_weizsacker_williams(oracle_ctx):
    import mpmath as mp
    o = oracle_ctx(flux_point=1, breakup_mode=1, nm=2, ny=2)
    g = 5020 / (2 * 0.9382720813)
    for b, k in ((7.0, 0.01), (30.0, 1.0), (500.0, 0.05), (10.0, 30.0)):
        x = b * k / (g * HC)
        ww = 82 ** 2 * ALPHA / math.pi ** 2 * (k / (g * HC)) ** 2 / k * (
            float(mp.besselk(1, x)) ** 2 + float(mp.besselk(0, x)) ** 2 / g ** 2)
        assert o.flux_point(b, k) == pytest.approx(ww, rel=1e-13)


def test_form_factor_flux_tends_to_point_flux(oracle_ctx):
    # with F = A the integral is w K1(w b): outside the nucleus the two agree
    o = oracle_ctx(flux_point=0, breakup_mode=1, nm=2, ny=2)
    for b, k in ((13.0, 0.003), (12.5, 0.05), (13.3, 1.0)):
        ff, nev, ier = o.flux_form(b, k)
        assert ier == 0
        pt = o.flux_point(b, k)
        assert ff == pytest.approx(pt, rel=2e-3)
    # inside the nucleus the charge is partly enclosed: the flux is suppressed
    ff, _, _ = o.flux_form(2.0, 0.05)
    assert ff < 0.5 * o.flux_point(2.0, 0.05)


def test_lumi_symmetric_in_rapidity(oracle_ctx):
    for kw in (dict(flux_point=1, breakup_mode=1), dict(flux_point=0, breakup_mode=2)):
        o = oracle_ctx(nm=2, ny=2, **kw)
        for M, Y in ((5.0, 1.3), (20.0, 3.7)):
            assert o.lumi(M, Y) == pytest.approx(o.lumi(M, -Y), rel=1e-13)


def test_lumi_scales_as_z4():
    a = O.Oracle(Z=82, flux_point=1, breakup_mode=1, nm=2, ny=2, num_threads=1)
    b = O.Oracle(Z=41, flux_point=1, breakup_mode=1, nm=2, ny=2, num_threads=1)
    for M, Y in ((4.0, 0.0), (30.0, -2.0)):
        assert a.lumi(M, Y) / b.lumi(M, Y) == pytest.approx(16.0, rel=1e-13)


def test_polarised_split_and_sign_convention(oracle_ctx):
    o = oracle_ctx(flux_point=1, breakup_mode=1, nm=2, ny=2)
    for M, Y in ((6.0, 0.5), (15.0, -1.5)):
        s, p = o.lumi_pol(M, Y)
        L = o.lumi(M, Y)
        # cos^2 + sin^2 = 1, but the polarised sum uses b^2 = b1^2+b2^2-2b1b2cos(phi)
        # and the unpolarised +2b1b2cos(phi) on nodes not symmetric about pi/2
        assert s > 0 and p > 0
        assert (s + p) / L == pytest.approx(1.0, rel=0.05)
        assert (s + p) != pytest.approx(L, rel=1e-12)


def test_grid_fixture_regression(oracle_ctx):
    gold = np.load(os.path.join(GOLD, "oracle_grids.npz"))
    o = oracle_ctx(flux_point=1, breakup_mode=1, nm=101, ny=41, nz=50, proc_id=13)
    g, nbad = o.lumi_grid()
    assert nbad == 0
    assert np.array_equal(g, gold["cfg1_point_mode1"])
    # lower bin edges: y_iy and -y_iy = y_{ny-iy}
    ny = 41
    assert np.allclose(g[:, 1:], g[:, ny - np.arange(1, ny)], rtol=1e-12, atol=0)


def test_breit_wheeler_total(oracle_ctx):
    for pdg, ml in ((11, 0.000510998946), (13, 0.1056583745), (15, 1.77686)):
        for m in (4.0, 10.0, 45.0):
            s = m * m
            beta = math.sqrt(1 - 4 * ml * ml / s)
            bw = 2 * math.pi * ALPHA ** 2 / s * ((3 - beta ** 4) * math.log((1 + beta) / (1 - beta))
                                                  - 2 * beta * (2 - beta ** 2))
            assert O.dilep_cs_m(pdg, 0.0, m) == pytest.approx(bw * HC ** 2 * 1e7, rel=1e-12)
            # dsigma/dz integrates to 16 sigma: the reference's z table is normalised
            # 16x high, harmless because only its per-M shape feeds the z samplers
            if pdg != 11:
                zint = si.quad(lambda z: O.dilep_cs_zm(pdg, 0.0, z, m), -1, 1, epsabs=0, epsrel=1e-12,
                               limit=500, points=[-0.999, 0.999])[0]
                assert zint * HC ** 2 * 1e7 == pytest.approx(16 * O.dilep_cs_m(pdg, 0.0, m), rel=1e-9)
            # scalar + pseudoscalar = 2 x unpolarised (fm^2 vs nb)
            sp = O.dilep_cs_m_pol_s(pdg, m) + O.dilep_cs_m_pol_ps(pdg, m)
            # (the reference's atanh(beta) form loses ~8 digits to cancellation for electrons)
            assert sp * 1e7 == pytest.approx(2 * O.dilep_cs_m(pdg, 0.0, m), rel=1e-7 if pdg == 11 else 1e-12)
    assert O.dilep_cs_m_pol_s(15, 3.0) == 0.0


def test_nuclear_cross_section_and_pdf(oracle_ctx):
    o = oracle_ctx(flux_point=1, breakup_mode=1, nm=101, ny=41, nz=50, proc_id=13)
    g, _ = o.lumi_grid()
    cs, _, tot = o.nuc_cs_ym(lumi=g)
    dm = (50 - 3.56) / 101
    m = 3.56 + dm * np.arange(101)
    sig = np.array([O.dilep_cs_m(13, 0.0, x) for x in m])
    assert np.allclose(cs, (sig[:, None] * g).T, rtol=1e-15, atol=0)
    assert tot == pytest.approx(cs.sum() * 1e-6, rel=1e-12)
    s, st = O.pdf_init(cs)
    assert st == 0 and s[0] == 0 and abs(s[-1] - 1) < 1e-12 and np.all(np.diff(s) >= 0)
    _, st = O.pdf_init(np.array([1.0, -1e-300, 2.0]))
    assert st == -1


def test_photon_flux_is_the_b_integral(oracle_ctx):
    """calcPhotonFlux is a 120-node midpoint rule in log b for
    2 pi k int P(min(b,20)) G(b) N(b, k) b db over [R, max(5 g hc / k, 5R)]:
    an adaptive quadrature of that integral (from the oracle's pinned spline
    and flux pieces) agrees to the rule's discretisation error."""
    from scipy.integrate import quad
    R, hc = 6.68, 0.1973269718
    o = oracle_ctx(flux_point=1, breakup_mode=2, nm=3, ny=3)
    g = o.g1
    for M, Y in ((3.1, 0.0), (9.4, 2.0), (3.1, -3.0)):
        k = M / 2 * np.exp(Y)
        bmax = max(5 * g * hc / k, 5 * R)

        def f(u):
            b = np.exp(u)
            P = o.breakup_spline(min(b, 20.0))
            G = o.gaa(b) if b < 20 else 1.0
            return P * G * o.flux_point(b, k) * b * b

        pts = [np.log(2 * R), np.log(20.0)]
        val, _ = quad(f, np.log(R), np.log(bmax), points=[p for p in pts if np.log(R) < p < np.log(bmax)],
                      limit=400, epsrel=1e-10)
        ref = 2 * np.pi * k * val
        got = o.photon_flux(M, Y)
        assert got == pytest.approx(ref, rel=1e-3), (M, Y, got, ref)  # measured 2e-5 .. 5e-4
        # and a wrong normalisation (pi instead of 2 pi) would be far outside
        assert abs(got / ref - 1) < 0.1
