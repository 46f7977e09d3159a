"""Pins of the oracle's calcBreakupProb and of its QAGS on the flux integrand
(CPU).

calcBreakupProb (src/UpcCrossSection.cpp:752-1019) is a trapezoid sum over a
photo-absorption table of about 600 energies built from literal data arrays.
Three checks pin it:
  - the table: the oracle's ee[] / se[] equal the table built from the
    literals as parsed out of the reference file
    (tests/golden/breakup_tables.npz, written by extract_breakup_tables.py),
    entry by entry, so no number was mistyped;
  - the normalisation: at impact parameters where only the giant dipole
    resonance segment is below the energy cut-off, p_xn(b) must equal
    mpmath's quadrature of zcon sigma_GDR(e) e K1(e b / (hc gamma_t))^2 over
    the same energy range (the trapezoid rule at 0.05 MeV is good to ~1e-5
    there); a wrong zcon, scon or Lorentzian fails it by orders of magnitude;
  - the K1 polynomial stays the one of TMath (A&S 9.8.7/9.8.8): with exact K1
    the sum moves by ~1e-7, above the test's rounding but below its
    quadrature tolerance, so that part is pinned by test_oracle_numerics.

The QAGS trace: scipy.integrate.quad runs QUADPACK's Fortran dqagse.  On the
flux integrand itself (the oracle's fluxFormIntegrand, the same Python
callable for both) the oracle's dqagse must return the same result, error,
evaluation count, status and final interval lists bit for bit, over (b, k)
nodes of the 1001 x 121 grid's rows, including ones with many intervals and
the ones whose nodes leave the form-factor table (NaN, ier 2); the oracle's
flux integral stops at the first pass that holds such a node.
"""
import math
import os

import mpmath as mp
import numpy as np
import pytest
import scipy.integrate as si

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PB = dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0)


@pytest.fixture(scope="module")
def lit():
    return np.load(os.path.join(GOLD, "breakup_tables.npz"))


@pytest.fixture(scope="module")
def tables():
    o = O.Oracle(flux_point=0, breakup_mode=2, nm=2, ny=2, num_threads=1, **PB)
    return o, o.breakup_tables()


def _table_from_literals(d, gamma_t):
    """ee[], se[] as calcBreakupProb's first call assembles them for lead,
    from the parsed literals (segment index ranges as in the reference's
    loops: GDR Lorentzian up to 25 MeV, Lepretre e3/s3 j = 1..27, Carlos
    e1/s1 j = 1..22, Armstrong sigt/sigtn j = 9..70 every 25 MeV, Michalowski
    e2/s2 j = 1..11, then the Regge tail)."""
    zp, ap = 82, 208
    o0, o1, g1, si1 = float(d["pb_o0"]), float(d["pb_o1"]), float(d["pb_g1"]), float(d["pb_si1"])
    delo = float(d["delo"])
    scon = .1 * g1 * g1 * si1
    ee, se = [], []
    ne = int((25. - o0) / delo) + 1
    for i in range(1, ne + 1):
        e = o0 + (i - 1) * delo
        ee.append(e)
        se.append(scon * e * e / (((o1 * o1 - e * e) * (o1 * o1 - e * e)) + e * e * g1 * g1))
    for j in range(1, 28):
        ee.append(float(d["e3"][j]))
        se.append(.1 * ap * float(d["s3"][j]) / 208.)
    for j in range(1, 23):
        ee.append(float(d["e1"][j]))
        se.append(.1 * ap * float(d["s1"][j]) / 208.)
    for j in range(9, 71):
        ee.append(ee[-1] + 25.)
        se.append(.1 * (zp * float(d["sigt"][j]) + (ap - zp) * float(d["sigtn"][j])))
    for j in range(1, 12):
        ee.append(float(d["e2"][j]))
        se.append(.1 * ap * float(d["s2"][j]))
    x, y, eps, eta, em = (float(d[k]) for k in ("x", "y", "eps", "eta", "em"))
    exx = math.pow(10, .05)
    s = .002 * em * ee[-1]
    ictr = 150 if gamma_t > 2. * 150. * 150. else 100
    for _ in range(ictr):
        s = s * exx
        ee.append(1000. * .5 * (s - em * em) / em)
        se.append(.1 * .65 * ap * (x * math.pow(s, eps) + y * math.pow(s, -eta)))
    ee.append(99999999999.)
    se.append(0.0)
    return np.array(ee), np.array(se)


def test_literal_arrays_have_declared_sizes(lit):
    for name in ("e1", "s1", "e2", "s2", "e3", "s3", "sa", "sen", "sigt", "sigtn"):
        assert len(lit[name]) == int(lit[name + "_declared"]), name
    # the lead branch and the constants the oracle uses
    assert (float(lit["pb_si1"]), float(lit["pb_g1"]), float(lit["pb_o1"]), float(lit["pb_o0"])) == \
        (640., 4.05, 13.42, 7.4)
    assert float(lit["pi"]) == 3.14159 and float(lit["alpha_inv"]) == 137.04


def test_breakup_table_matches_reference_literals(lit, tables):
    o, t = tables
    g1 = 5020.0 / (2 * 0.9382720813)
    gamma_t = 2 * g1 * g1 - 1
    ee, se = _table_from_literals(lit, gamma_t)
    assert len(t["ee"]) == len(ee)
    assert np.array_equal(t["ee"][:-1], ee[:-1]) and t["ee"][-1] == ee[-1]
    # every segment but the Regge tail is exact; the tail's pow() may differ
    # from Python's by an ulp
    n_lit = len(ee) - 1 - (150 if gamma_t > 45000 else 100)
    assert np.array_equal(t["se"][:n_lit], se[:n_lit])
    assert np.max(np.abs(t["se"][n_lit:-1] / se[n_lit:-1] - 1)) < 4e-16
    # one-neutron table: 100 x the lead GDR one-neutron data, energies o0 + 0.1 (i - 1)
    assert np.array_equal(t["sa"], 100. * lit["sen"][1:161])
    assert np.array_equal(t["eee"][:160], 7.4 + .1 * np.arange(160)) and t["eee"][160] == 24.1
    # zcon = (Z / (gamma_t pi hc))^2 / 137.04 with the truncated pi of the reference
    pi, hc, ai = float(lit["pi"]), float(lit["hbarcmev"]), float(lit["alpha_inv"])
    assert t["zcon"] == 82 / (gamma_t * pi * hc) * 82 / (gamma_t * pi * hc) / ai
    assert t["o0"] == 7.4


@pytest.mark.parametrize("omax_target", [12.0, 18.8, 24.7])
def test_gdr_segment_against_quadrature(lit, tables, omax_target):
    """Where 4 gamma_t hc / b < 25 MeV only GDR energies enter p_xn(b): the
    trapezoid sum must match the integral of its own integrand."""
    o, t = tables
    g1 = 5020.0 / (2 * 0.9382720813)
    gamma_t = 2 * g1 * g1 - 1
    hc = 197.3269718
    b = 4 * gamma_t * hc / omax_target
    ee = t["ee"]
    kmax = int(np.searchsorted(ee, omax_target)) - 1  # last ee < omax
    e_hi = ee[kmax]
    o0, o1, gw, si1 = 7.4, 13.42, 4.05, 640.
    scon = .1 * gw * gw * si1
    zcon = t["zcon"]
    mp.mp.dps = 30

    def f(e):
        sig = scon * e * e / ((o1 * o1 - e * e) ** 2 + e * e * gw * gw)
        return zcon * sig * e * mp.besselk(1, e * b / (hc * gamma_t)) ** 2

    exact = float(mp.quad(f, mp.linspace(o0, e_hi, 9)))
    pxn = o.breakup_pxn(b)
    print(f"omax {omax_target} MeV, b {b:.4e} fm: p_xn {pxn:.10e}, quadrature {exact:.10e}, "
          f"rel {pxn / exact - 1:.2e}")
    assert pxn == pytest.approx(exact, rel=1e-4)
    # and the returned probabilities are the reference's functions of it
    assert o.breakup_prob(b, 3) == pytest.approx(math.exp(-2 * pxn), rel=1e-15)
    assert o.breakup_prob(b, 2) == pytest.approx((1 - math.exp(-pxn)) ** 2, rel=1e-12)


def test_breakup_cutoff_below_gdr_threshold(tables):
    o, _ = tables
    g1 = 5020.0 / (2 * 0.9382720813)
    gamma_t = 2 * g1 * g1 - 1
    b = 4 * gamma_t * 197.3269718 / 7.0  # omax = 7 MeV < o0
    assert o.breakup_pxn(b) == -1.0 and o.breakup_prob(b, 3) == 0.0


def _grid_nodes(nm=1001, ny=121, R=6.68, picks=None):
    g = 5020.0 / (2 * 0.9382720813)
    out = []
    for im, iy, j in picks:
        dm, dy = (50 - 3.56) / nm, 12.0 / ny
        M, Y = 3.56 + dm * im, -6.0 + dy * iy
        k = M / 2. * math.exp(Y)
        bmin, bmax = 0.05 * R, max(5. * g * 0.1973269718 / k, 5. * R)
        ld = (math.log(bmax) - math.log(bmin)) / 120
        b = (bmin * math.exp((j + 1.) * ld) + bmin * math.exp(j * ld)) / 2.
        out.append((b, k))
    return out


# (im, iy, j) of config-2 flux rows: small and large k, first and last QAGS
# nodes, rows with long interval lists, and two nodes whose QK21 abscissae
# leave the form-factor table (the reference's GSL_EDOM: NaN, ier 2)
TRACE_PICKS = [(0, 0, 0), (0, 60, 40), (0, 120, 90), (500, 60, 60), (1000, 0, 10), (1000, 120, 95),
               (250, 30, 5), (750, 90, 70), (213, 117, 62), (497, 101, 28), (497, 101, 0), (120, 100, 50),
               (900, 110, 80), (40, 10, 30), (640, 64, 64), (333, 111, 33)]


@pytest.mark.parametrize("pick", range(len(TRACE_PICKS)))
def test_qags_trace_matches_quadpack_on_flux_integrand(tables, pick):
    o, _ = tables
    (b, k), = _grid_nodes(picks=[TRACE_PICKS[pick]])
    assert b <= 2 * 6.68

    def f(q):
        return o.flux_integrand(b, k, q)

    calls = []

    def f_logged(q):
        v = f(q)
        calls.append(v)
        return v

    res, err, nev, ier, (al, bl, rl, el) = O.qags(f, 0.0, 10.0, 1e-4, 1e-4)
    s = si.quad(f_logged, 0.0, 10.0, epsabs=1e-4, epsrel=1e-4, limit=1000, full_output=1)
    info = s[2]
    n = info["last"]
    print(f"b {b:.4f} k {k:.4g}: result {res!r} neval {nev} intervals {n} ier {ier}")
    assert nev == info["neval"] and len(al) == n
    same = (lambda u, v: np.array_equal(u, v, equal_nan=True))
    assert same(np.array([res, err]), np.array([s[0], s[1]]))
    for mine, key in ((al, "alist"), (bl, "blist"), (rl, "rlist"), (el, "elist")):
        assert same(mine, info[key][:n]), key
    # The oracle's flux value and status are this QAGS run's, except that an
    # integral stops after the first QK21 pass holding a NaN node (past the
    # last form-factor knot, the reference's GSL_EDOM): dqagse evaluates the
    # whole range (pass 0), then two passes per bisection step, 21 nodes
    # each, in the order scipy called the integrand.
    v, nev2, ier2 = o.flux_form(b, k)
    assert len(calls) == nev
    bad = np.flatnonzero(np.isnan(np.array(calls)))
    if bad.size == 0:
        assert nev2 == nev and (ier2 != 0) == (ier != 0)
        assert v == v
    else:
        p = int(bad[0]) // 21
        assert nev2 == 21 + 42 * ((p + 1) // 2) and ier2 == 100 and np.isnan(v)
