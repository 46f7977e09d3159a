"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is a C restatement of upcgen's
luminosity / nuclear cross-section path (src/UpcCrossSection.cpp,
src/UpcTwoPhotonDilep.cpp, include/UpcSampler.h of the reference).  It is the
checker for the HIP product and the timed CPU baseline.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it; the
product package (upcgen_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Params(C.Structure):
    _fields_ = [
        ("Z", C.c_int), ("A", C.c_int),
        ("R", C.c_double), ("a", C.c_double), ("sqrts", C.c_double),
        ("flux_point", C.c_int), ("breakup_mode", C.c_int), ("use_polarized", C.c_int),
        ("nm", C.c_int), ("mmin", C.c_double), ("mmax", C.c_double),
        ("ny", C.c_int), ("ymin", C.c_double), ("ymax", C.c_double),
        ("nz", C.c_int), ("zmin", C.c_double), ("zmax", C.c_double),
        ("proc_id", C.c_int), ("lep_a", C.c_double),
        ("num_threads", C.c_int),
    ]


DEFAULTS = dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0, flux_point=1, breakup_mode=1,
                use_polarized=0, nm=1001, mmin=3.56, mmax=50.0, ny=121, ymin=-6.0, ymax=6.0,
                nz=100, zmin=-1.0, zmax=1.0, proc_id=13, lep_a=0.0,
                num_threads=os.cpu_count() or 1)

_lib = None
QFN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
_DP = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i, vp = C.c_double, C.c_int, C.c_void_p
        for name in ("orc_bessel_k0", "orc_bessel_k1", "orc_bessel_j1", "orc_bessel_k1_as",
                     "orc_bessel_i1_as"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d]
        L.orc_simpson.restype = d
        L.orc_simpson.argtypes = [i, _DP, d]
        L.orc_qags_trace.restype = i
        L.orc_qags_trace.argtypes = [QFN, vp, d, d, d, d, i, _DP, _DP, C.POINTER(i), C.POINTER(i),
                                     _DP, _DP, _DP, _DP]
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(Params)]
        L.orc_destroy.argtypes = [vp]
        for name in ("orc_rho0", "orc_factor", "orc_g1", "orc_cs_nn"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [vp]
        for name in ("orc_gaa", "orc_ta", "orc_formfac", "orc_formfac_spline", "orc_breakup_spline"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [vp, d]
        L.orc_breakup_prob.restype = d
        L.orc_breakup_prob.argtypes = [vp, d, i]
        L.orc_breakup_pxn.restype = d
        L.orc_breakup_pxn.argtypes = [vp, d]
        L.orc_breakup_tables.restype = i
        L.orc_breakup_tables.argtypes = [vp, _DP, _DP, _DP, _DP, _DP, _DP]
        L.orc_table_times.argtypes = [vp, _DP]
        L.orc_flux_integrand.restype = d
        L.orc_flux_integrand.argtypes = [vp, d, d, d]
        L.orc_flux_point.restype = d
        L.orc_flux_point.argtypes = [vp, d, d]
        L.orc_flux_form.restype = d
        L.orc_flux_form.argtypes = [vp, d, d, C.POINTER(i), C.POINTER(i)]
        L.orc_lumi.restype = d
        L.orc_lumi.argtypes = [vp, d, d]
        L.orc_lumi_pol.argtypes = [vp, d, d, _DP, _DP]
        L.orc_lumi_grid.restype = C.c_long
        L.orc_lumi_grid.argtypes = [vp, _DP, _DP, _DP]
        L.orc_lumi_grid_counts.restype = C.c_long
        L.orc_lumi_grid_counts.argtypes = [vp, _DP, _DP, _DP, C.POINTER(C.c_long)]
        for name in ("orc_dilep_cs_m_pol_s", "orc_dilep_cs_m_pol_ps"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [i, d]
        for name in ("orc_dilep_cs_zm_pol_s", "orc_dilep_cs_zm_pol_ps"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [i, d, d]
        L.orc_dilep_cs_m.restype = d
        L.orc_dilep_cs_m.argtypes = [i, d, d]
        L.orc_dilep_cs_zm.restype = d
        L.orc_dilep_cs_zm.argtypes = [i, d, d, d]
        L.orc_fill_cs_zm.argtypes = [vp, _DP, i]
        L.orc_nuc_cs_ym.argtypes = [vp, _DP, _DP, _DP, _DP, _DP, _DP]
        L.orc_photon_pt_hist.argtypes = [vp, d, d, _DP]
        L.orc_set_qags_tol.argtypes = [vp, d, d]
        L.orc_photon_flux.argtypes = [vp, d, d]
        L.orc_photon_flux.restype = d
        L.orc_nuc_cs_y.argtypes = [vp, d, _DP, _DP, _DP, _DP, _DP]
        L.orc_pdf_init.restype = i
        L.orc_pdf_init.argtypes = [_DP, i, _DP]
        L.orc_pdf_sample.restype = d
        L.orc_pdf_sample.argtypes = [_DP, _DP, i, d, C.POINTER(i)]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def make_params(**kw) -> Params:
    v = dict(DEFAULTS)
    v.update(kw)
    return Params(**v)


class Oracle:
    """One oracle context = one fresh reference process with these parameters."""

    def __init__(self, **kw):
        self.params = make_params(**kw)
        self._ctx = lib().orc_create(C.byref(self.params))

    def close(self):
        if self._ctx:
            lib().orc_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # scalars
    rho0 = property(lambda s: lib().orc_rho0(s._ctx))
    factor = property(lambda s: lib().orc_factor(s._ctx))
    g1 = property(lambda s: lib().orc_g1(s._ctx))
    cs_nn = property(lambda s: lib().orc_cs_nn(s._ctx))

    def gaa(self, b): return lib().orc_gaa(self._ctx, b)
    def ta(self, b): return lib().orc_ta(self._ctx, b)
    def formfac(self, q2): return lib().orc_formfac(self._ctx, q2)
    def formfac_spline(self, q2): return lib().orc_formfac_spline(self._ctx, q2)
    def breakup_prob(self, b, mode): return lib().orc_breakup_prob(self._ctx, b, mode)
    def breakup_spline(self, b): return lib().orc_breakup_spline(self._ctx, b)
    def breakup_pxn(self, b): return lib().orc_breakup_pxn(self._ctx, b)

    def table_times(self):
        """Wall seconds of the context's table phases: G_AA, form factor, breakup."""
        out = np.zeros(3)
        lib().orc_table_times(self._ctx, _ptr(out))
        return dict(gaa=out[0], form_factor=out[1], breakup=out[2])
    def flux_integrand(self, b, k, q): return lib().orc_flux_integrand(self._ctx, b, k, q)

    def breakup_tables(self):
        """calcBreakupProb's tables: dict(ee, se) [1..n], eee, sa [1..161/160], zcon, o0."""
        ee, se = np.zeros(10002), np.zeros(10002)
        eee, sa = np.zeros(162), np.zeros(162)
        zc, o0 = C.c_double(), C.c_double()
        n = lib().orc_breakup_tables(self._ctx, _ptr(ee), _ptr(se), _ptr(eee), _ptr(sa), C.byref(zc),
                                     C.byref(o0))
        return dict(ee=ee[1:n + 1], se=se[1:n + 1], eee=eee[1:162], sa=sa[1:161], zcon=zc.value, o0=o0.value)
    def flux_point(self, b, k): return lib().orc_flux_point(self._ctx, b, k)

    def flux_form(self, b, k):
        nev, ier = C.c_int(0), C.c_int(0)
        v = lib().orc_flux_form(self._ctx, b, k, C.byref(nev), C.byref(ier))
        return v, nev.value, ier.value

    def lumi(self, M, Y): return lib().orc_lumi(self._ctx, M, Y)

    def lumi_pol(self, M, Y):
        s, p = C.c_double(), C.c_double()
        lib().orc_lumi_pol(self._ctx, M, Y, C.byref(s), C.byref(p))
        return s.value, p.value

    def lumi_grid(self, counts=False):
        """The grid (or (Ls, Lp)) and the number of QAGS calls with ier != 0.
        counts=True also returns an int64 array [nm][ny][4]: the integrand
        evaluations of each bin's k1 and k2 flux rows and the number of
        their integrals with ier != 0."""
        nm, ny = self.params.nm, self.params.ny
        cnt = np.zeros((nm, ny, 4), dtype=np.int64)
        cp = cnt.ctypes.data_as(C.POINTER(C.c_long))
        dummy = np.zeros(1)
        if self.params.use_polarized:
            s = np.zeros((nm, ny)); p = np.zeros((nm, ny))
            nbad = lib().orc_lumi_grid_counts(self._ctx, _ptr(dummy), _ptr(s), _ptr(p), cp)
            res = (s, p)
        else:
            res = np.zeros((nm, ny))
            nbad = lib().orc_lumi_grid_counts(self._ctx, _ptr(res), _ptr(dummy), _ptr(dummy), cp)
        return (res, nbad, cnt) if counts else (res, nbad)

    def cs_zm(self, flag=0):
        out = np.zeros((self.params.nm, self.params.nz))
        lib().orc_fill_cs_zm(self._ctx, _ptr(out), flag)
        return out

    def photon_flux(self, M, Y):
        return lib().orc_photon_flux(self._ctx, M, Y)

    def nuc_cs_y(self, mpart, cs_pos, cs_neg):
        ny = self.params.ny
        cs_pos = np.ascontiguousarray(cs_pos, dtype=np.float64)
        cs_neg = np.ascontiguousarray(cs_neg, dtype=np.float64)
        cs, ratio, tot = np.empty(ny), np.empty(ny), C.c_double(0.0)
        lib().orc_nuc_cs_y(self._ctx, mpart, _ptr(cs_pos), _ptr(cs_neg), _ptr(cs), _ptr(ratio), C.byref(tot))
        return cs, ratio, tot.value

    def set_qags_tol(self, epsabs, epsrel):
        lib().orc_set_qags_tol(self._ctx, epsabs, epsrel)

    def photon_pt_hist(self, E, gtot):
        out = np.zeros(5000)
        lib().orc_photon_pt_hist(self._ctx, E, gtot, _ptr(out))
        return out

    def nuc_cs_ym(self, lumi=None, lumi_s=None, lumi_p=None):
        nm, ny = self.params.nm, self.params.ny
        d = np.zeros(1)
        cs = np.zeros((ny, nm)); ratio = np.zeros((ny, nm)); tot = C.c_double(0.0)
        la = np.ascontiguousarray(lumi) if lumi is not None else d
        ls = np.ascontiguousarray(lumi_s) if lumi_s is not None else d
        lp = np.ascontiguousarray(lumi_p) if lumi_p is not None else d
        lib().orc_nuc_cs_ym(self._ctx, _ptr(la), _ptr(ls), _ptr(lp), _ptr(cs), _ptr(ratio),
                            C.byref(tot))
        return cs, ratio, tot.value


def bessel_k0(x): return lib().orc_bessel_k0(x)
def bessel_k1(x): return lib().orc_bessel_k1(x)
def bessel_j1(x): return lib().orc_bessel_j1(x)
def bessel_k1_as(x): return lib().orc_bessel_k1_as(x)
def bessel_i1_as(x): return lib().orc_bessel_i1_as(x)


def simpson(v, h):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_simpson(len(v), _ptr(v), h)


def qags(f, a, b, epsabs, epsrel, limit=1000):
    """QUADPACK dqagse on a Python callable; returns (result, abserr, neval, ier, intervals)."""
    cb = QFN(lambda x, _p: f(x))
    res, err = C.c_double(), C.c_double()
    nev, last = C.c_int(), C.c_int()
    al, bl, rl, el = (np.zeros(limit) for _ in range(4))
    ier = lib().orc_qags_trace(cb, None, a, b, epsabs, epsrel, limit, C.byref(res), C.byref(err),
                               C.byref(nev), C.byref(last), _ptr(al), _ptr(bl), _ptr(rl), _ptr(el))
    n = last.value
    return res.value, err.value, nev.value, ier, (al[:n], bl[:n], rl[:n], el[:n])


def dilep_cs_m(pdg, lep_a, m): return lib().orc_dilep_cs_m(pdg, lep_a, m)
def dilep_cs_zm(pdg, lep_a, z, m): return lib().orc_dilep_cs_zm(pdg, lep_a, z, m)
def dilep_cs_m_pol_s(pdg, m): return lib().orc_dilep_cs_m_pol_s(pdg, m)
def dilep_cs_m_pol_ps(pdg, m): return lib().orc_dilep_cs_m_pol_ps(pdg, m)
def dilep_cs_zm_pol_s(pdg, z, m): return lib().orc_dilep_cs_zm_pol_s(pdg, z, m)
def dilep_cs_zm_pol_ps(pdg, z, m): return lib().orc_dilep_cs_zm_pol_ps(pdg, z, m)


class _Spline(C.Structure):
    _fields_ = [("n", C.c_int), ("x", _DP), ("y", _DP), ("c", _DP)]


def spline_eval(x, y, xq):
    """Natural cubic spline (gsl_interp_cspline semantics) through (x, y) at xq."""
    L = lib()
    L.orc_spline_init.argtypes = [C.POINTER(_Spline), _DP, _DP, C.c_int]
    L.orc_spline_init.restype = C.c_int
    L.orc_spline_eval.argtypes = [C.POINTER(_Spline), C.c_double, C.POINTER(C.c_int)]
    L.orc_spline_eval.restype = C.c_double
    L.orc_spline_free.argtypes = [C.POINTER(_Spline)]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    s = _Spline()
    L.orc_spline_init(C.byref(s), _ptr(x), _ptr(y), len(x))
    err = C.c_int(0)
    out = np.array([L.orc_spline_eval(C.byref(s), float(v), C.byref(err)) for v in np.atleast_1d(xq)])
    L.orc_spline_free(C.byref(s))
    return out


def pdf_init(bins):
    bins = np.ascontiguousarray(bins, dtype=np.float64).ravel()
    s = np.zeros(len(bins) + 1)
    st = lib().orc_pdf_init(_ptr(bins), len(bins), _ptr(s))
    return s, st
