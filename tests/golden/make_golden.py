"""Writes the golden fixtures under tests/golden/.

  bessel_kat.npz   K0, K1, J1 from mpmath (30 digits) on log/linear grids
  formfac_kat.npz  Fourier transform of the Woods-Saxon charge density
                   rho(r) = rho0 / (1 + exp((r - R)/a)), by mpmath quadrature,
                   for the Pb parameters (R = 6.68, a = 0.447 fm)
  oracle_grids.npz regression grids from the CPU oracle: the 101x41 point-flux
                   grid (Pb, mode 1) and 7x5 form-factor grids (modes 2, 3 pol)
  oracle_xnxn_1001x121.npz
                   the oracle's full 1001x121 Pb-Pb 5.02 TeV XNXN form-factor
                   grid (FLUX_POINT 0, BREAKUP_MODE 2) with the QAGS integrand
                   evaluations of every bin's k1 row and of the iy = 0 k2 rows
                   (the rows a mirror-deduplicated grid computes), the total
                   over all 2 x 121,121 rows, the ier != 0 count, and the
                   number of bins whose k2 row count differs from its mirror
                   bin's k1 row count
  oracle_config3.npz
                   the 64-job mixed batch: eight bins of every job from the
                   oracle, and full 9x7 polarised form-factor grids of O-O
                   7 TeV and Au-Au 200 GeV in breakup modes 1-4

Run from the repository root:  python tests/golden/make_golden.py
Nothing here imports or runs the upcgen reference.
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def bessel():
    mp.mp.dps = 30
    xk = np.concatenate([np.logspace(-10, 0, 60), np.linspace(1.01, 10, 60), np.logspace(1.01, np.log10(700), 60)])
    k0 = np.array([float(mp.besselk(0, x)) for x in xk])
    k1 = np.array([float(mp.besselk(1, x)) for x in xk])
    xj = np.concatenate([np.linspace(0.0, 20, 201), np.linspace(20.5, 700, 400)])
    j1 = np.array([float(mp.besselj(1, x)) for x in xj])
    np.savez(os.path.join(HERE, "bessel_kat.npz"), xk=xk, k0=k0, k1=k1, xj=xj, j1=j1)


def formfac():
    mp.mp.dps = 30
    R, a, hc = mp.mpf("6.68"), mp.mpf("0.447"), mp.mpf("0.1973269718")
    rho0 = mp.mpf(1)  # F is linear in rho0: tests scale by the oracle's rho0
    q2s = [1e-6, 1e-4, 1e-3, 5e-3, 0.01, 0.03, 0.1, 0.3, 1.0]
    vals = []
    for q2 in q2s:
        q = mp.sqrt(q2) / hc
        rho = lambda r: rho0 / (1 + mp.exp((r - R) / a))
        f = mp.quad(lambda r: 4 * mp.pi * r * r * rho(r) * mp.sin(q * r) / (q * r), mp.linspace(0, 60, 61))
        vals.append(float(f))
    np.savez(os.path.join(HERE, "formfac_kat.npz"), q2=np.array(q2s), f_over_rho0=np.array(vals))


def oracle_grids():
    import oracle as O
    o1 = O.Oracle(flux_point=1, breakup_mode=1, nm=101, ny=41, nz=50, proc_id=13, num_threads=8)
    g1, bad1 = o1.lumi_grid()
    assert bad1 == 0
    o2 = O.Oracle(flux_point=0, breakup_mode=2, nm=7, ny=5, num_threads=8)
    g2, bad2 = o2.lumi_grid()
    assert bad2 == 0
    o3 = O.Oracle(flux_point=0, breakup_mode=3, use_polarized=1, nm=7, ny=5, num_threads=8)
    (g3s, g3p), bad3 = o3.lumi_grid()
    assert bad3 == 0
    np.savez(os.path.join(HERE, "oracle_grids.npz"), cfg1_point_mode1=g1, ff_mode2_7x5=g2,
             ff_mode3_pol_s_7x5=g3s, ff_mode3_pol_p_7x5=g3p)


def oracle_full_grid(num_threads=None):
    """About 4 minutes on 8 cores."""
    import oracle as O
    o = O.Oracle(flux_point=0, breakup_mode=2, nm=1001, ny=121, nz=100, proc_id=13,
                 num_threads=num_threads or os.cpu_count() or 1)
    g, bad, cnt = o.lumi_grid(counts=True)
    ev1, ev2, bad1, bad2 = (cnt[..., i] for i in range(4))
    ny = g.shape[1]
    mirror_mismatch = int(np.count_nonzero(ev2[:, 1:] != ev1[:, ny - 1:0:-1]) +
                          np.count_nonzero(bad2[:, 1:] != bad1[:, ny - 1:0:-1]))
    np.savez_compressed(os.path.join(HERE, "oracle_xnxn_1001x121.npz"), L=g, qags_bad=np.int64(bad),
                        ev1=ev1.astype(np.int32), ev2_col0=ev2[:, 0].astype(np.int32),
                        bad1=bad1.astype(np.int16), bad2_col0=bad2[:, 0].astype(np.int16),
                        evals_all_rows=np.int64(ev1.sum() + ev2.sum()),
                        evals_dedup_rows=np.int64(ev1.sum() + ev2[:, 0].sum()),
                        bad_dedup_rows=np.int64(bad1.sum() + bad2[:, 0].sum()),
                        mirror_mismatch=np.int64(mirror_mismatch))


def oracle_config3(num_threads=None):
    """The 64-job mixed batch (upcgen_amd.mixed_systems_jobs(seed=2024):
    Pb/Xe/O/Au x breakup modes 1-4, polarised, form-factor flux): eight bins
    of every job (its four corners and four drawn with a per-job seed) from
    the oracle's calcTwoPhotonLumiPol, x dm x dy; plus the full 9 x 7
    polarised form-factor grids of O-O 7 TeV and Au-Au 200 GeV in modes 1-4.
    About 2 minutes on 8 cores."""
    import oracle as O
    import upcgen_amd as U
    nt = num_threads or os.cpu_count() or 1
    jobs = U.mixed_systems_jobs(seed=2024)
    ctxs = {}
    rows = []
    for j, c in enumerate(jobs):
        key = (c.Z, c.breakup_mode)
        if key not in ctxs:
            ctxs[key] = O.Oracle(**{k: getattr(c, k) for k in ("Z", "A", "R", "a", "sqrts", "flux_point",
                                                                  "breakup_mode", "use_polarized")},
                                 num_threads=nt)
        o = ctxs[key]
        rng = np.random.default_rng(1000 + j)
        ims = np.concatenate([[0, 0, c.nm - 1, c.nm - 1], rng.integers(0, c.nm, 4)])
        iys = np.concatenate([[0, c.ny - 1, 0, c.ny - 1], rng.integers(0, c.ny, 4)])
        dm, dy = (c.mmax - c.mmin) / c.nm, (c.ymax - c.ymin) / c.ny
        for im, iy in zip(ims, iys):
            ls, lp = o.lumi_pol(c.mmin + dm * im, c.ymin + dy * iy)
            rows.append((j, im, iy, ls * dm * dy, lp * dm * dy))
    rows = np.array(rows)
    grids = {}
    for name, sysp in (("OO", dict(Z=8, A=16, R=U.wood_saxon_radius(16), a=0.54, sqrts=7000.0)),
                       ("AuAu", dict(Z=79, A=197, R=U.wood_saxon_radius(197), a=0.54, sqrts=200.0))):
        for mode in (1, 2, 3, 4):
            o = O.Oracle(flux_point=0, breakup_mode=mode, use_polarized=1, nm=9, ny=7, num_threads=nt, **sysp)
            (s, p), bad, cnt = o.lumi_grid(counts=True)
            grids[f"{name}_mode{mode}_s"] = s
            grids[f"{name}_mode{mode}_p"] = p
            # failed integrals of the rows a mirror-deduplicated grid computes:
            # every bin's k1 row and the k2 rows of the iy = 0 column
            grids[f"{name}_mode{mode}_bad"] = np.int64(cnt[..., 2].sum() + cnt[:, 0, 3].sum())
            grids[f"{name}_mode{mode}_evals"] = np.int64(cnt[..., 0].sum() + cnt[:, 0, 1].sum())
    np.savez_compressed(os.path.join(HERE, "oracle_config3.npz"), job=rows[:, 0].astype(np.int32),
                        im=rows[:, 1].astype(np.int32), iy=rows[:, 2].astype(np.int32), Ls=rows[:, 3],
                        Lp=rows[:, 4], **grids)


if __name__ == "__main__":
    what = sys.argv[1:] or ["bessel", "formfac", "oracle_grids", "oracle_full_grid", "oracle_config3"]
    for name in what:
        globals()[name]()
    print("golden fixtures written to", HERE)
