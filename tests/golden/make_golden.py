"""Writes the golden fixtures under tests/golden/.

  bessel_kat.npz   K0, K1, J1 from mpmath (30 digits) on log/linear grids
  formfac_kat.npz  Fourier transform of the Woods-Saxon charge density
                   rho(r) = rho0 / (1 + exp((r - R)/a)), by mpmath quadrature,
                   for the Pb parameters (R = 6.68, a = 0.447 fm)
  oracle_grids.npz regression grids from the CPU oracle: the 101x41 point-flux
                   grid (Pb, mode 1) and 7x5 form-factor grids (modes 2, 3 pol)

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


if __name__ == "__main__":
    bessel()
    formfac()
    oracle_grids()
    print("golden fixtures written to", HERE)
