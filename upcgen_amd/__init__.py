"""upcgen_amd: MI355X-native two-photon luminosity / nuclear cross-section engine.

Python bindings of the C ABI in include/upc_lumi.h (libupcgen_amd.so, HIP
kernels for gfx950).  torch provides device buffers and streams; every
computation runs in the HIP library.  There is no CPU fallback: loading fails
loudly when the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

__all__ = ["Config", "Info", "Context", "load", "FORM_FP64", "FORM_MFMA", "FORM_SURROGATE", "FORM_MFMA_FP32",
           "UpcError", "LIB_PATH", "Batch", "Tune", "BatchStats", "mixed_systems_jobs", "wood_saxon_radius",
           "bessel_j1"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libupcgen_amd.so")

FORM_FP64, FORM_MFMA, FORM_SURROGATE, FORM_MFMA_FP32 = 0, 1, 2, 3


class UpcError(RuntimeError):
    pass


class Config(C.Structure):
    """upc_config: the parameters.in keys that reach the luminosity path."""
    _fields_ = [
        ("Z", C.c_int), ("A", C.c_int),
        ("R", C.c_double), ("a", C.c_double), ("sqrts", C.c_double),
        ("flux_point", C.c_int), ("breakup_mode", C.c_int), ("use_polarized", C.c_int),
        ("nm", C.c_int), ("mmin", C.c_double), ("mmax", C.c_double),
        ("ny", C.c_int), ("ymin", C.c_double), ("ymax", C.c_double),
        ("nz", C.c_int), ("zmin", C.c_double), ("zmax", C.c_double),
        ("proc_id", C.c_int), ("lep_a", C.c_double),
    ]

    # include/UpcCrossSection.h:72-164 defaults, PROC_ID 13
    DEFAULTS = dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0, flux_point=1, breakup_mode=1,
                    use_polarized=0, nm=1001, mmin=3.56, mmax=50.0, ny=121, ymin=-6.0, ymax=6.0,
                    nz=100, zmin=-1.0, zmax=1.0, proc_id=13, lep_a=0.0)

    @classmethod
    def make(cls, **kw) -> "Config":
        v = dict(cls.DEFAULTS)
        unknown = set(kw) - set(v)
        if unknown:
            raise TypeError(f"unknown config keys: {sorted(unknown)}")
        v.update(kw)
        return cls(**v)

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Info(C.Structure):
    _fields_ = [("g1", C.c_double), ("g2", C.c_double), ("factor", C.c_double),
                ("mNucl", C.c_double), ("rho0", C.c_double), ("csNN", C.c_double),
                ("p20", C.c_double), ("qags_calls", C.c_longlong), ("qags_bad", C.c_longlong),
                ("qags_evals", C.c_longlong), ("mfma_ksteps", C.c_longlong), ("mfma_spacing", C.c_double),
                ("surrogate_points", C.c_longlong), ("surrogate_fallback", C.c_longlong), ("bskip", C.c_double),
                ("mfma_fp64_items", C.c_longlong)]


class EventOpts(C.Structure):
    _fields_ = [("nonzero_gam_pt", C.c_int), ("do_pt_cut", C.c_int), ("pt_min", C.c_double),
                ("do_eta_cut", C.c_int), ("eta_min", C.c_double), ("eta_max", C.c_double)]


class Timings(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_longlong * 8)]

PHASES = ("tables", "flux", "qags", "lumi", "xsec", "pdf", "project")


class Tune(C.Structure):
    _fields_ = [("qags_blocks_per_cu", C.c_int), ("lumi_waves", C.c_int),
                ("qags_candidates", C.c_int * 3), ("qags_ms", C.c_double * 3),
                ("lumi_candidates", C.c_int * 3), ("lumi_ms", C.c_double * 3),
                ("form", C.c_int), ("mfma_tile_group", C.c_int),
                ("mfma_candidates", C.c_int * 3), ("mfma_ms", C.c_double * 3)]


class BatchStats(C.Structure):
    _fields_ = [("njobs", C.c_int), ("points", C.c_longlong), ("rows", C.c_longlong),
                ("qags_calls", C.c_longlong), ("qags_evals", C.c_longlong), ("qags_bad", C.c_longlong),
                ("qags_blocks_per_cu", C.c_int), ("lumi_waves", C.c_int), ("tasks", C.c_longlong),
                ("work_bytes", C.c_longlong), ("form", C.c_int), ("mfma_tile_group", C.c_int),
                ("mfma_fp64_items", C.c_longlong)]


# Every symbol declared in include/upc_lumi.h with its ctypes signature.
_vp, _i, _d, _ll, _ull = C.c_void_p, C.c_int, C.c_double, C.c_longlong, C.c_ulonglong
SIGNATURES = {
    "upc_last_error": (C.c_char_p, []),
    "upc_device_count": (_i, [C.POINTER(_i)]),
    "upc_ctx_create": (_i, [C.POINTER(Config), _vp, C.POINTER(_vp)]),
    "upc_ctx_destroy": (_i, [_vp]),
    "upc_ctx_info": (_i, [_vp, C.POINTER(Info)]),
    "upc_ctx_rebuild_tables": (_i, [_vp, _vp]),
    "upc_ctx_set_timing": (_i, [_vp, _i]),
    "upc_ctx_timings": (_i, [_vp, C.POINTER(Timings), _i]),
    "upc_ctx_set_mfma_spacing": (_i, [_vp, _d]),
    "upc_ctx_set_qags_cap": (_i, [_vp, _i]),
    "upc_flux_point": (_i, [_vp, _vp, _vp, _i, _vp, _vp]),
    "upc_flux_form": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "upc_flux_form_noalloc": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "upc_bessel_j1": (_i, [_vp, _i, _vp, _vp]),
    "upc_form_factor": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "upc_breakup_prob": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "upc_nuclear_overlap": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "upc_lumi_points": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "upc_lumi_grid": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "upc_photon_flux": (_i, [_vp, _vp, _vp, _i, _vp, _vp]),
    "upc_nuc_cs_y": (_i, [_vp, _d, _vp, _vp, _vp, _vp, C.POINTER(_d), _vp]),
    "upc_cs_zm": (_i, [_vp, _d, _d, _i, _d, _d, _i, _i, _vp, _vp]),
    "upc_cs_m": (_i, [_vp, _i, _vp, _vp]),
    "upc_elem_cs": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "upc_nuc_cs_ym": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_d), _vp]),
    "upc_pdf_init": (_i, [_vp, _i, _i, _vp, _vp]),
    "upc_photon_pt_tables": (_i, [_vp, _d, _vp]),
    "upc_photon_pt": (_i, [_vp, _vp, _ll, _ull, _vp, _vp]),
    "upc_generate_pairs": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(EventOpts), _ull, _ull, _ll, _vp,
                                _vp, _vp, _vp]),
    "upc_sample_events": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _ull, _ull, _ll, _vp, _vp, _vp, _vp]),
    "upc_batch_create": (_i, [C.POINTER(Config), _i, _vp, C.POINTER(_vp)]),
    "upc_batch_destroy": (_i, [_vp]),
    "upc_batch_set_outputs": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "upc_batch_run": (_i, [_vp, _i, _vp]),
    "upc_batch_autotune": (_i, [_vp, _vp, C.POINTER(Tune)]),
    "upc_batch_set_tune": (_i, [_vp, _i, _i]),
    "upc_batch_set_form": (_i, [_vp, _i]),
    "upc_ctx_set_mfma_cancel_limit": (_i, [_vp, _d]),
    "upc_batch_set_mfma_tiles": (_i, [_vp, _i]),
    "upc_batch_info": (_i, [_vp, _vp, C.POINTER(BatchStats)]),
    "upc_batch_job_failures": (_i, [_vp, _vp, C.POINTER(_ll)]),
    "upc_sur_create": (_i, [_ull, C.POINTER(_vp)]),
    "upc_sur_destroy": (_i, [_vp]),
    "upc_sur_param_count": (_i, [C.POINTER(_i)]),
    "upc_sur_get_params": (_i, [_vp, _i, _vp]),
    "upc_sur_set_params": (_i, [_vp, _i, _vp]),
    "upc_sur_set_norm": (_i, [_vp, _d, _d]),
    "upc_sur_get_norm": (_i, [_vp, C.POINTER(_d), C.POINTER(_d)]),
    "upc_sur_set_precision": (_i, [_vp, _i]),
    "upc_sur_set_err_scale": (_i, [_vp, _d]),
    "upc_sur_get_err_scale": (_i, [_vp, C.POINTER(_d)]),
    "upc_sur_features": (_i, [C.POINTER(Config), _vp, _vp, _ll, _vp, _vp, _vp]),
    "upc_sur_forward": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "upc_sur_grad": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "upc_sur_train": (_i, [_vp, _vp, _vp, _ll, _i, _i, _d, _i, _vp, _vp]),
    "upc_ctx_set_surrogate_anchors": (_i, [_vp, _i, _i]),
    "upc_ctx_set_surrogate": (_i, [_vp, _vp, _d]),
}

_lib = None


def load() -> C.CDLL:
    """Load libupcgen_amd.so (build it first with upcgen_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UpcError(f"{LIB_PATH} is missing: run `python -m upcgen_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise UpcError(f"{what} failed ({rc}): {load().upc_last_error().decode()}")


def _dptr(t, dtype=None, numel=None, device=None, name="tensor"):
    """Device pointer of a tensor after checking what the C call will assume:
    a contiguous device tensor, of `dtype` (float64 unless stated), holding
    at least `numel` elements, on `device` when given.  A mismatch raises
    UpcError before any kernel runs (the C ABI takes plain pointers and would
    read or write past the end of a shorter or narrower buffer)."""
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise UpcError(f"{name}: device tensor expected")
    if not t.is_contiguous():
        raise UpcError(f"{name}: contiguous tensor expected")
    want = torch.float64 if dtype is None else dtype
    if t.dtype != want:
        raise UpcError(f"{name}: dtype {t.dtype}, expected {want}")
    if numel is not None and t.numel() < numel:
        raise UpcError(f"{name}: {t.numel()} elements, at least {numel} expected")
    if device is not None and t.device.index != device:
        raise UpcError(f"{name}: on cuda:{t.device.index}, the context lives on cuda:{device}")
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


class Context:
    """One configuration = one upc_ctx (UpcCrossSection after init())."""

    def __init__(self, config: Config | None = None, stream=None, **kw):
        import torch
        if not torch.cuda.is_available():
            raise UpcError("no HIP device visible: upcgen_amd runs on the GPU only")
        self.config = config if config is not None else Config.make(**kw)
        self.device = torch.cuda.current_device()
        self._ctx = C.c_void_p()
        _check(load().upc_ctx_create(C.byref(self.config), _stream(stream), C.byref(self._ctx)),
               "upc_ctx_create")

    def close(self):
        if self._ctx:
            load().upc_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def info(self) -> Info:
        inf = Info()
        _check(load().upc_ctx_info(self._ctx, C.byref(inf)), "upc_ctx_info")
        return inf

    def set_mfma_spacing(self, h_fm):
        """Node spacing [fm] of the MFMA forms' shared b grid."""
        _check(load().upc_ctx_set_mfma_spacing(self._ctx, h_fm), "upc_ctx_set_mfma_spacing")

    def set_mfma_cancel_limit(self, limit):
        """MFMA forms: bins whose far-field term exceeds limit x |L| are
        recomputed by the FP64 contraction (default 2; <= 0: off)."""
        _check(load().upc_ctx_set_mfma_cancel_limit(self._ctx, float(limit)), "upc_ctx_set_mfma_cancel_limit")

    def set_qags_cap(self, cap):
        """Intervals the fast QAGS kernel holds before handing an integral to
        the 1000-interval kernel (1..16); results do not depend on it."""
        _check(load().upc_ctx_set_qags_cap(self._ctx, int(cap)), "upc_ctx_set_qags_cap")

    def set_timing(self, enable=True):
        _check(load().upc_ctx_set_timing(self._ctx, int(enable)), "upc_ctx_set_timing")

    def timings(self, reset=True):
        """Summed device milliseconds per phase (HIP events on the launch stream)."""
        t = Timings()
        _check(load().upc_ctx_timings(self._ctx, C.byref(t), int(reset)), "upc_ctx_timings")
        return {name: (t.ms[i], t.launches[i]) for i, name in enumerate(PHASES)}

    def rebuild_tables(self, stream=None):
        _check(load().upc_ctx_rebuild_tables(self._ctx, _stream(stream)), "upc_ctx_rebuild_tables")

    # --- batched per-point functions (torch float64 cuda tensors) ---
    def _empty(self, n, dtype=None):
        import torch
        return torch.empty(n, dtype=dtype or torch.float64, device=f"cuda:{self.device}")

    def _in(self, t, numel=None, name="input", dtype=None):
        return _dptr(t, dtype=dtype, numel=numel, device=self.device, name=name)

    def _pair(self, a, b, names):
        """Two per-point inputs of equal length: (n, ptr_a, ptr_b)."""
        n = a.numel()
        if b is None or b.numel() != n:
            raise UpcError(f"{names[0]} and {names[1]} must have the same number of elements")
        return n, self._in(a, n, names[0]), self._in(b, n, names[1])

    def flux_point(self, b, k, stream=None):
        n, pb, pk = self._pair(b, k, ("b", "k"))
        out = self._empty(n)
        _check(load().upc_flux_point(self._ctx, pb, pk, n, _dptr(out), _stream(stream)), "upc_flux_point")
        return out

    def flux_form(self, b, k, stream=None):
        import torch
        n, pb, pk = self._pair(b, k, ("b", "k"))
        out = self._empty(n)
        ier = self._empty(n, torch.int32)
        _check(load().upc_flux_form(self._ctx, pb, pk, n, _dptr(out), _dptr(ier, torch.int32),
                                    _stream(stream)), "upc_flux_form")
        return out, ier

    def form_factor(self, q2, tabulated=False, stream=None):
        out = self._empty(q2.numel())
        _check(load().upc_form_factor(self._ctx, self._in(q2, name="q2"), q2.numel(), int(tabulated), _dptr(out),
                                      _stream(stream)), "upc_form_factor")
        return out

    def breakup_prob(self, b, mode=None, tabulated=False, stream=None):
        out = self._empty(b.numel())
        m = self.config.breakup_mode if mode is None else mode
        _check(load().upc_breakup_prob(self._ctx, self._in(b, name="b"), b.numel(), m, int(tabulated), _dptr(out),
                                       _stream(stream)), "upc_breakup_prob")
        return out

    def nuclear_overlap(self, b, which=0, stream=None):
        out = self._empty(b.numel())
        _check(load().upc_nuclear_overlap(self._ctx, self._in(b, name="b"), b.numel(), which, _dptr(out),
                                          _stream(stream)), "upc_nuclear_overlap")
        return out

    def lumi_points(self, M, Y, stream=None):
        n, pm, py = self._pair(M, Y, ("M", "Y"))
        if self.config.use_polarized:
            ls, lp = self._empty(n), self._empty(n)
            _check(load().upc_lumi_points(self._ctx, pm, py, n, None, _dptr(ls), _dptr(lp),
                                          _stream(stream)), "upc_lumi_points")
            return ls, lp
        out = self._empty(n)
        _check(load().upc_lumi_points(self._ctx, pm, py, n, _dptr(out), None, None,
                                      _stream(stream)), "upc_lumi_points")
        return out

    def photon_flux(self, M, Y, stream=None):
        """calcPhotonFlux(M, Y) at n points (vector-meson photon flux)."""
        n, pm, py = self._pair(M, Y, ("M", "Y"))
        out = self._empty(n)
        _check(load().upc_photon_flux(self._ctx, pm, py, n, _dptr(out), _stream(stream)), "upc_photon_flux")
        return out

    def nuc_cs_y(self, mpart, cs_pos, cs_neg, stream=None):
        """calcNucCrossSectionY with sigma(+-y) given: (cs[ny], ratio[ny], totCS)."""
        ny = self.config.ny
        cs, ratio = self._empty(ny), self._empty(ny)
        tot = C.c_double(0.0)
        _check(load().upc_nuc_cs_y(self._ctx, mpart, self._in(cs_pos, ny, "cs_pos"), self._in(cs_neg, ny, "cs_neg"),
                                   _dptr(cs), _dptr(ratio),
                                   C.byref(tot), _stream(stream)), "upc_nuc_cs_y")
        return cs, ratio, tot.value

    # --- grids ---
    def lumi_grid(self, form=FORM_FP64, out=None, stream=None):
        """prepareTwoPhotonLumi grid [nm][ny] (x dm x dy). Polarised -> (Ls, Lp)."""
        import torch
        nm, ny = self.config.nm, self.config.ny
        dev = f"cuda:{self.device}"
        if self.config.use_polarized:
            ls, lp = out if out is not None else (
                torch.empty((nm, ny), dtype=torch.float64, device=dev),
                torch.empty((nm, ny), dtype=torch.float64, device=dev))
            _check(load().upc_lumi_grid(self._ctx, form, None, self._in(ls, nm * ny, "out Ls"),
                                        self._in(lp, nm * ny, "out Lp"), _stream(stream)), "upc_lumi_grid")
            return ls, lp
        L = out if out is not None else torch.empty((nm, ny), dtype=torch.float64, device=dev)
        _check(load().upc_lumi_grid(self._ctx, form, self._in(L, nm * ny, "out"), None, None, _stream(stream)),
               "upc_lumi_grid")
        return L

    def cs_zm(self, flag=0, stream=None, zmin=None, zmax=None, nz=None, mmin=None, mmax=None, nm=None):
        """fillCrossSectionZM over the configured grid (or the ranges given): [nm][nz]."""
        import torch
        c = self.config
        zmin = c.zmin if zmin is None else zmin
        zmax = c.zmax if zmax is None else zmax
        nz = c.nz if nz is None else nz
        mmin = c.mmin if mmin is None else mmin
        mmax = c.mmax if mmax is None else mmax
        nm = c.nm if nm is None else nm
        out = torch.empty((nm, nz), dtype=torch.float64, device="cuda")
        _check(load().upc_cs_zm(self._ctx, zmin, zmax, nz, mmin, mmax, nm, flag, _dptr(out),
                                _stream(stream)), "upc_cs_zm")
        return out

    def cs_m(self, flag=0, stream=None):
        out = self._empty(self.config.nm)
        _check(load().upc_cs_m(self._ctx, flag, _dptr(out), _stream(stream)), "upc_cs_m")
        return out

    def elem_cs(self, flag, m, z=None, stream=None):
        """Elementary sigma: flag 0/1/2 = sigma(M), S, PS; 3/4/5 = dsigma/dz(z, M), S, PS."""
        out = self._empty(m.numel())
        if flag >= 3 and (z is None or z.numel() != m.numel()):
            raise UpcError("elem_cs: z must have as many elements as m")
        _check(load().upc_elem_cs(self._ctx, flag, self._in(m, name="m"), self._in(z, m.numel(), "z"), m.numel(),
                                  _dptr(out),
                                  _stream(stream)), "upc_elem_cs")
        return out

    def nuc_cs_ym(self, lumi, stream=None, tot_cs=0.0):
        """calcNucCrossSectionYM: returns (cs [ny][nm], ratio or None, totCS [mb])."""
        import torch
        nm, ny = self.config.nm, self.config.ny
        dev = f"cuda:{self.device}"
        cs = torch.empty((ny, nm), dtype=torch.float64, device=dev)
        tot = C.c_double(tot_cs)
        if self.config.use_polarized:
            ls, lp = lumi
            ratio = torch.empty((ny, nm), dtype=torch.float64, device=dev)
            _check(load().upc_nuc_cs_ym(self._ctx, None, self._in(ls, nm * ny, "Ls"), self._in(lp, nm * ny, "Lp"),
                                        _dptr(cs), _dptr(ratio), C.byref(tot), _stream(stream)), "upc_nuc_cs_ym")
            return cs, ratio, tot.value
        _check(load().upc_nuc_cs_ym(self._ctx, self._in(lumi, nm * ny, "lumi"), None, None, _dptr(cs), None,
                                    C.byref(tot), _stream(stream)), "upc_nuc_cs_ym")
        return cs, None, tot.value

    @staticmethod
    def pdf_init(bins, stream=None):
        """gsl_histogram_pdf_init per row of a 2-D tensor (rows, n) -> (rows, n+1)."""
        import torch
        b2 = bins.reshape(-1, bins.shape[-1]) if bins.dim() > 1 else bins.reshape(1, -1)
        b2 = b2.contiguous()
        rows, n = b2.shape
        out = torch.empty((rows, n + 1), dtype=torch.float64, device=b2.device)
        _check(load().upc_pdf_init(_dptr(b2, name="bins"), rows, n, _dptr(out), _stream(stream)),
               "upc_pdf_init")
        return out

    def _tables(self, sum_ym, sum_z, sum_zs, sum_zp, ratio):
        """Pointers to the sampler tables, checked against the grid sizes."""
        nm, ny, nz = self.config.nm, self.config.ny, self.config.nz
        if self.config.use_polarized:
            if sum_zs is None or sum_zp is None or ratio is None:
                raise UpcError("polarised sampling needs sum_zs, sum_zp and ratio")
        elif sum_z is None:
            raise UpcError("sampling needs sum_z")
        return (self._in(sum_ym, ny * nm + 1, "sum_ym"), self._in(sum_z, nm * (nz + 1), "sum_z"),
                self._in(sum_zs, nm * (nz + 1), "sum_zs"), self._in(sum_zp, nm * (nz + 1), "sum_zp"),
                self._in(ratio, ny * nm, "ratio"))

    def photon_pt_tables(self, gtot, stream=None):
        _check(load().upc_photon_pt_tables(self._ctx, gtot, _stream(stream)), "upc_photon_pt_tables")

    def photon_pt(self, E, seed=1, stream=None):
        out = self._empty(E.numel())
        _check(load().upc_photon_pt(self._ctx, self._in(E, name="E"), E.numel(), seed, _dptr(out), _stream(stream)),
               "upc_photon_pt")
        return out

    def generate_pairs(self, sum_ym, sum_z=None, sum_zs=None, sum_zp=None, ratio=None, n=1000, seed=12345,
                       offset=0, opts=None, stream=None):
        """Dilepton events: (p [n,2,4] (px,py,pz,E), pdg [n,2], pass [n])."""
        import torch
        o = opts if opts is not None else EventOpts(0, 0, 0.0, 0, 0.0, 0.0)
        dev = f"cuda:{self.device}"
        tables = self._tables(sum_ym, sum_z, sum_zs, sum_zp, ratio)
        p = torch.empty((n, 2, 4), dtype=torch.float64, device=dev)
        pdg = torch.empty((n, 2), dtype=torch.int32, device=dev)
        ok = torch.empty(n, dtype=torch.int32, device=dev)
        _check(load().upc_generate_pairs(self._ctx, *tables, C.byref(o), seed, offset, n, _dptr(p),
                                         _dptr(pdg, torch.int32), _dptr(ok, torch.int32), _stream(stream)),
               "upc_generate_pairs")
        return p, pdg, ok

    def sample_events(self, sum_ym, sum_z=None, sum_zs=None, sum_zp=None, ratio=None, n=1000,
                      seed=12345, offset=0, stream=None):
        tables = self._tables(sum_ym, sum_z, sum_zs, sum_zp, ratio)
        y, m, z = self._empty(n), self._empty(n), self._empty(n)
        _check(load().upc_sample_events(self._ctx, *tables, seed, offset, n, _dptr(y), _dptr(m), _dptr(z),
                                        _stream(stream)), "upc_sample_events")
        return y, m, z


def bessel_j1(x, stream=None):
    """J1 on the device, as the flux integrand evaluates it (float64 tensor)."""
    import torch
    x = torch.as_tensor(x, dtype=torch.float64, device="cuda").contiguous()
    out = torch.empty_like(x)
    _check(load().upc_bessel_j1(_dptr(x), x.numel(), _dptr(out), _stream(stream)), "upc_bessel_j1")
    return out


def wood_saxon_radius(A):
    """R = 1.12 A^(1/3) - 0.86 A^(-1/3) fm."""
    return 1.12 * A ** (1 / 3) - 0.86 * A ** (-1 / 3)


def mixed_systems_jobs(seed=2024, n=64, polarized=1, flux_point=0):
    """The 64-job mixed batch: Pb-Pb 5.02 TeV, Xe-Xe 5.44 TeV, O-O 7 TeV and
    Au-Au 200 GeV x breakup modes 1-4 (each combination n / 16 times), M bins
    drawn from {101, 251, 501, 1001, 2001} and Y bins from {41, 81, 121} with
    `seed`; R from wood_saxon_radius, a = 0.54 fm except Pb (0.447)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    systems = [(82, 208, 0.447, 5020.0), (54, 129, 0.54, 5440.0), (8, 16, 0.54, 7000.0), (79, 197, 0.54, 200.0)]
    jobs = []
    for i in range(n):
        Z, A, a, sqrts = systems[i % 4]
        mode = 1 + (i // 4) % 4
        jobs.append(Config.make(Z=Z, A=A, R=wood_saxon_radius(A), a=a, sqrts=sqrts, flux_point=flux_point,
                                breakup_mode=mode, use_polarized=polarized,
                                nm=int(rng.choice([101, 251, 501, 1001, 2001])),
                                ny=int(rng.choice([41, 81, 121]))))
    return jobs


class Batch:
    """Many configurations whose luminosity grids run as one launch sequence
    (one persistent QAGS launch, one contraction launch), optionally replayed
    from a captured hipGraph."""

    def __init__(self, configs, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise UpcError("no HIP device visible: upcgen_amd runs on the GPU only")
        self.configs = list(configs)
        arr = (Config * len(self.configs))(*self.configs)
        self._b = C.c_void_p()
        _check(load().upc_batch_create(arr, len(self.configs), _stream(stream), C.byref(self._b)),
               "upc_batch_create")
        n = len(self.configs)
        dev = "cuda"
        self.outputs = []
        for c in self.configs:
            if c.use_polarized:
                self.outputs.append((torch.empty((c.nm, c.ny), dtype=torch.float64, device=dev),
                                     torch.empty((c.nm, c.ny), dtype=torch.float64, device=dev)))
            else:
                self.outputs.append(torch.empty((c.nm, c.ny), dtype=torch.float64, device=dev))
        L = (_vp * n)(*[None if c.use_polarized else o.data_ptr() for c, o in zip(self.configs, self.outputs)])
        Ls = (_vp * n)(*[o[0].data_ptr() if c.use_polarized else None for c, o in zip(self.configs, self.outputs)])
        Lp = (_vp * n)(*[o[1].data_ptr() if c.use_polarized else None for c, o in zip(self.configs, self.outputs)])
        _check(load().upc_batch_set_outputs(self._b, L, Ls, Lp), "upc_batch_set_outputs")

    def close(self):
        if self._b:
            load().upc_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, graph=True, stream=None):
        """All grids into self.outputs ([nm][ny] x dm x dy; (Ls, Lp) for polarised jobs)."""
        _check(load().upc_batch_run(self._b, int(graph), _stream(stream)), "upc_batch_run")
        return self.outputs

    def autotune(self, stream=None) -> Tune:
        t = Tune()
        _check(load().upc_batch_autotune(self._b, _stream(stream), C.byref(t)), "upc_batch_autotune")
        return t

    def set_tune(self, qags_blocks_per_cu, lumi_waves):
        _check(load().upc_batch_set_tune(self._b, qags_blocks_per_cu, lumi_waves), "upc_batch_set_tune")

    def set_form(self, form):
        """FORM_FP64 (default), FORM_MFMA or FORM_MFMA_FP32 for every job's contraction."""
        _check(load().upc_batch_set_form(self._b, int(form)), "upc_batch_set_form")

    def set_mfma_tiles(self, tile_group):
        """MFMA column tiles per A-fragment load (1, 2 or 4; same results)."""
        _check(load().upc_batch_set_mfma_tiles(self._b, int(tile_group)), "upc_batch_set_mfma_tiles")

    def job_failures(self, stream=None):
        """Per job of the last run: flux integrals that ended with a QAGS error
        or a form-factor spline domain error (the reference aborts there)."""
        n = len(self.configs)
        out = (_ll * n)()
        _check(load().upc_batch_job_failures(self._b, _stream(stream), out), "upc_batch_job_failures")
        return list(out)

    def stats(self, stream=None) -> BatchStats:
        st = BatchStats()
        _check(load().upc_batch_info(self._b, _stream(stream), C.byref(st)), "upc_batch_info")
        return st
