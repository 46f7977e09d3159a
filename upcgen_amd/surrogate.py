"""Learned surrogate of log L(M, Y; system): data generation by the FP64 GPU
quadrature, training by the library's own HIP kernels, and serving as the
UPC_FORM_SURROGATE luminosity form (with quadrature fallback).

Training distribution (one system per context, points drawn per system):
Z ~ U[8, 92], A = round(2.5 Z), sqrt(s_NN) log-uniform in [0.2, 40] TeV,
breakup mode uniform in 1..4, form-factor flux, Woods-Saxon R = (1.12 A^(1/3)
- 0.86 A^(-1/3)) x U[0.95, 1.05] fm and a ~ U[0.42, 0.58] fm (so that the
Pb-Pb R = 6.68 fm, a = 0.447 fm system lies inside it); M log-uniform in
[1, 200] GeV, Y ~ U[-7, 7].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Config, Context, UpcError, _check, _dptr, _stream, load, wood_saxon_radius

IN = 32


def _dev():
    import torch
    if not torch.cuda.is_available():
        raise UpcError("no HIP device visible: upcgen_amd runs on the GPU only")
    return torch


def param_count() -> int:
    n = C.c_int()
    _check(load().upc_sur_param_count(C.byref(n)), "upc_sur_param_count")
    return n.value


def features(config: Config, M, Y, stream=None):
    """(X [n][32] float32, log L0 [n] float64) on the GPU for points of one system."""
    torch = _dev()
    M = torch.as_tensor(M, dtype=torch.float64, device="cuda").contiguous()
    Y = torch.as_tensor(Y, dtype=torch.float64, device="cuda").contiguous()
    n = M.numel()
    X = torch.empty((n, IN), dtype=torch.float32, device="cuda")
    L0 = torch.empty(n, dtype=torch.float64, device="cuda")
    _check(load().upc_sur_features(C.byref(config), _dptr(M, numel=n), _dptr(Y, numel=n), n,
                                   _dptr(X, torch.float32), _dptr(L0), _stream(stream)),
           "upc_sur_features")
    return X, L0


class Surrogate:
    """Two-network ensemble (upc_sur)."""

    def __init__(self, seed: int = 1):
        _dev()
        self.seed = seed
        self.split = True
        self._s = C.c_void_p()
        _check(load().upc_sur_create(seed, C.byref(self._s)), "upc_sur_create")
        self.nparams = param_count()

    def close(self):
        if self._s:
            load().upc_sur_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self, net: int) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(load().upc_sur_get_params(self._s, net, out.ctypes.data_as(C.c_void_p)), "upc_sur_get_params")
        return out

    def set_params(self, net: int, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise UpcError(f"expected {self.nparams} parameters")
        _check(load().upc_sur_set_params(self._s, net, p.ctypes.data_as(C.c_void_p)), "upc_sur_set_params")

    @property
    def norm(self):
        mu, sig = C.c_double(), C.c_double()
        _check(load().upc_sur_get_norm(self._s, C.byref(mu), C.byref(sig)), "upc_sur_get_norm")
        return mu.value, sig.value

    @norm.setter
    def norm(self, v):
        _check(load().upc_sur_set_norm(self._s, float(v[0]), float(v[1])), "upc_sur_set_norm")

    def set_precision(self, split: bool):
        """Split-bf16 products (default, about FP32 products) or plain bf16."""
        _check(load().upc_sur_set_precision(self._s, int(bool(split))), "upc_sur_set_precision")
        self.split = bool(split)

    @property
    def err_scale(self):
        v = C.c_double()
        _check(load().upc_sur_get_err_scale(self._s, C.byref(v)), "upc_sur_get_err_scale")
        return v.value

    @err_scale.setter
    def err_scale(self, v):
        _check(load().upc_sur_set_err_scale(self._s, float(v)), "upc_sur_set_err_scale")

    def calibrate(self, X, logL0, log_true, max_err=1e-3, coverage=0.99, min_served=0.01):
        """Error scale from held-out points: the smallest scale c such that,
        of the points the gate would serve (c * spread <= max_err), at least
        `coverage` are within max_err, counting only scales that serve at
        least a fraction min_served of the points (and at least 100 of them:
        a coverage measured on a handful of points says nothing).  When no
        scale achieves that, c is set so large that nothing is served."""
        pred, spread = self.log_lumi(X, logL0)
        err = (pred - log_true).abs().cpu().numpy()
        spr = np.maximum(spread.cpu().numpy(), 1e-300)
        order = np.argsort(spr)
        ok = np.cumsum(err[order] <= max_err) / np.arange(1, len(order) + 1)
        # serving the k smallest spreads is acceptable when the running coverage holds
        kmin = max(100, int(np.ceil(min_served * len(order))))
        good = np.nonzero(ok >= coverage)[0]
        good = good[good + 1 >= kmin]
        if good.size == 0:
            self.err_scale = 1e300
            return self.err_scale
        k = good.max()
        self.err_scale = float(max(1.0, max_err / spr[order[k]]))
        return self.err_scale

    def forward(self, net: int, X, stream=None):
        torch = _dev()
        B = X.shape[0]
        out = torch.empty(B, dtype=torch.float32, device="cuda")
        _check(load().upc_sur_forward(self._s, net, _dptr(X, torch.float32, B * IN, name="X"), B,
                                      _dptr(out, torch.float32), _stream(stream)), "upc_sur_forward")
        return out

    def grad(self, net: int, X, T, stream=None):
        torch = _dev()
        g = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
        B = X.shape[0]
        _check(load().upc_sur_grad(self._s, net, _dptr(X, torch.float32, B * IN, name="X"),
                                   _dptr(T, torch.float32, B, name="T"), B, _dptr(g, torch.float32),
                                   _stream(stream)),
               "upc_sur_grad")
        return g

    def train(self, X, T, steps=4000, batch=4096, lr=2e-3, graph=True, stream=None):
        """Adam on both networks; returns the per-step loss curves [2][steps]."""
        loss = np.zeros((2, steps), dtype=np.float32)
        import torch
        n = X.shape[0]
        _check(load().upc_sur_train(self._s, _dptr(X, torch.float32, n * IN, name="X"),
                                    _dptr(T, torch.float32, n, name="T"), n, steps, batch, lr, int(graph),
                                    _stream(stream), loss.ctypes.data_as(C.c_void_p)), "upc_sur_train")
        return loss

    def predict(self, X):
        """Normalised predictions of both networks for any number of rows."""
        torch = _dev()
        n = X.shape[0]
        n32 = (n + 31) // 32 * 32
        if n32 != n:
            X = torch.cat([X, torch.zeros((n32 - n, IN), dtype=X.dtype, device=X.device)])
        outs = []
        for net in (0, 1):
            parts = [self.forward(net, X[o:o + 65536].contiguous()) for o in range(0, n32, 65536)]
            outs.append(torch.cat(parts)[:n])
        return outs

    def log_lumi(self, X, logL0):
        """(log L prediction, ensemble spread |r_a - r_b|) in float64; the
        predicted error is err_scale times the spread."""
        pa, pb = self.predict(X)
        mu, sig = self.norm
        ra, rb = mu + sig * pa.double(), mu + sig * pb.double()
        return logL0 + 0.5 * (ra + rb), (ra - rb).abs()

    def copy(self, seed=None):
        """A new surrogate with the same weights, normalisation and error scale."""
        c = Surrogate(self.seed if seed is None else seed)
        c.set_precision(self.split)
        for net in (0, 1):
            c.set_params(net, self.get_params(net))
        c.norm = self.norm
        c.err_scale = self.err_scale
        return c

    def fit_system(self, config, n=32768, n_valid=4096, steps=15000, batch=4096, lr=5e-4, max_err=1e-3,
                   coverage=0.99, seed=7):
        """A copy of this surrogate fine-tuned on one system: n FP64
        quadrature labels of `config` at uniform (M, Y) over its grid range
        (inside the domain), `steps` Adam steps from the current weights, and
        the error scale calibrated on n_valid further points of the system.
        Returns (surrogate, seconds spent on labels, seconds of training)."""
        import time
        torch = _dev()
        cfg = Config.make(**{f: getattr(config, f) for f, _ in Config._fields_})
        mr, yr = (cfg.mmin, cfg.mmax), (cfg.ymin, cfg.ymax)
        t0 = time.perf_counter()
        X, r, _, _ = make_dataset(0, n, seed, systems=[cfg], mrange=mr, yrange=yr, log_m=False)
        Xv, _, lgv, L0v = make_dataset(0, n_valid, seed + 101, systems=[cfg], mrange=mr, yrange=yr, log_m=False)
        torch.cuda.synchronize()
        t_label = time.perf_counter() - t0
        s = self.copy()
        mu, sig = s.norm
        t0 = time.perf_counter()
        s.train(X.contiguous(), ((r - mu) / sig).float().contiguous(), steps=steps, batch=batch, lr=lr)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        s.calibrate(Xv.contiguous(), L0v, lgv, max_err=max_err, coverage=coverage)
        return s, t_label, t_train

    def attach(self, ctx: Context, max_err=1e-3, anchors=(16, 4)):
        """Serve ctx's UPC_FORM_SURROGATE grids.  anchors: the M and Y
        spacing of the FP64 anchor lattice that judges the network (a bin is
        served when the network is within max_err / 2 at the four anchors
        around it), or None for the ensemble-spread gate."""
        _check(load().upc_ctx_set_surrogate(ctx._ctx, self._s, max_err), "upc_ctx_set_surrogate")
        em, ey = anchors if anchors else (0, 0)
        _check(load().upc_ctx_set_surrogate_anchors(ctx._ctx, int(em), int(ey)), "upc_ctx_set_surrogate_anchors")
        ctx._surrogate = self  # keep alive while attached

    def save(self, path):
        mu, sig = self.norm
        np.savez(path, net0=self.get_params(0), net1=self.get_params(1), mu=mu, sigma=sig,
                 err_scale=self.err_scale)

    @classmethod
    def load(cls, path, seed=1):
        d = np.load(path)
        s = cls(seed)
        s.set_params(0, d["net0"])
        s.set_params(1, d["net1"])
        s.norm = (float(d["mu"]), float(d["sigma"]))
        s.err_scale = float(d["err_scale"])
        return s


def sample_systems(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Z = int(rng.integers(8, 93))
        A = int(round(2.5 * Z))
        sqrts = float(np.exp(rng.uniform(np.log(200.0), np.log(40000.0))))
        out.append(Config.make(Z=Z, A=A, R=wood_saxon_radius(A) * rng.uniform(0.95, 1.05),
                               a=rng.uniform(0.42, 0.58), sqrts=sqrts, flux_point=0,
                               breakup_mode=int(rng.integers(1, 5)), use_polarized=0, nm=2, ny=2))
    return out


XI_MAX = 4.0  # domain of the network: both photons with xi = k R / (gamma hc) <= XI_MAX
M_PROTON = 0.9382720813


def xi_max(cfg, M, Y):
    """max(xi1, xi2), xi = k R / (gamma hc), k = M/2 e^{+-Y}: the domain test."""
    gamma = cfg.sqrts / (2 * M_PROTON)
    return np.asarray(M) / 2 * np.exp(np.abs(np.asarray(Y))) * cfg.R / (gamma * 0.1973269718)


def sample_points(cfg, n, rng, mrange=(1.0, 200.0), yrange=(-7.0, 7.0), log_m=True):
    """n points (M, Y) of the training distribution (M log-uniform or uniform
    in mrange, Y uniform in yrange) inside the network's domain: draws
    outside it are rejected and drawn again.  Fewer than n come back only
    when the domain holds almost none of the range (at most 64 rounds)."""
    Ms, Ys, got = [], [], 0
    for _ in range(64):
        k = max(2 * (n - got), 1024)
        M = np.exp(rng.uniform(np.log(mrange[0]), np.log(mrange[1]), k)) if log_m else rng.uniform(*mrange, k)
        Y = rng.uniform(yrange[0], yrange[1], k)
        keep = xi_max(cfg, M, Y) <= XI_MAX
        Ms.append(M[keep][:n - got])
        Ys.append(Y[keep][:n - got])
        got += len(Ms[-1])
        if got >= n:
            break
    return np.concatenate(Ms), np.concatenate(Ys)


def make_dataset(n_systems, per_system, seed, systems=None, mrange=(1.0, 200.0), yrange=(-7.0, 7.0),
                 log_m=True):
    """Features, targets r = log L - log L0 and the labels, from the FP64 GPU
    quadrature (upc_lumi_points) at per_system random (M, Y) inside the
    domain (sample_points) of each of n_systems random systems (or of
    `systems`).  Points where L is zero or not finite are dropped."""
    torch = _dev()
    rng = np.random.default_rng(seed + 1)
    systems = systems if systems is not None else sample_systems(n_systems, seed)
    Xs, Rs, Ls, L0s = [], [], [], []
    for cfg in systems:
        M, Y = sample_points(cfg, per_system, rng, mrange, yrange, log_m)
        ctx = Context(cfg)
        L = ctx.lumi_points(torch.as_tensor(M, device="cuda"), torch.as_tensor(Y, device="cuda"))
        ctx.close()
        X, L0 = features(cfg, M, Y)
        lg = torch.log(L)
        ok = torch.isfinite(lg) & torch.isfinite(L0) & (L > 0)
        Xs.append(X[ok])
        Rs.append((lg - L0)[ok])
        Ls.append(lg[ok])
        L0s.append(L0[ok])
    return torch.cat(Xs), torch.cat(Rs), torch.cat(Ls), torch.cat(L0s)


def train_default(n_systems=256, per_system=4096, steps=30000, batch=8192, lr=2e-3, seed=11, n_valid=32):
    """The default training run (n_systems x per_system points inside the
    domain, 2^20 by default) with the error scale calibrated on n_valid
    further systems.  Returns (surrogate, loss curves, training rows)."""
    X, r, _, _ = make_dataset(n_systems, per_system, seed)
    mu, sig = float(r.mean()), float(r.std())
    T = ((r - mu) / sig).float().contiguous()
    s = Surrogate(seed)
    s.norm = (mu, sig)
    loss = s.train(X.contiguous(), T, steps=steps, batch=batch, lr=lr)
    Xv, _, lgv, L0v = make_dataset(0, 2048, seed + 5, systems=sample_systems(n_valid, seed + 7))
    s.calibrate(Xv.contiguous(), L0v, lgv)
    return s, loss, X.shape[0]
