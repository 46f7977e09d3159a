"""The drop-in facade on the GPU: the upcgen command line end to end
(parameters.in -> UpcGenerator::init -> luminosity cache + lock -> nuclear
cross section -> events in HepMC), its cache/lock protocol, and the event
kernels' distributions against the oracle."""
import math
import os
import re
import subprocess
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "upcgen_amd", "upcgen")

SMALL = dict(NUCLEUS_Z=82, NUCLEUS_A=208, WS_R=6.68, WS_A=0.447, SQRTS=5020, PROC_ID=13, LEP_A=0,
             NEVENTS=3000, DO_PT_CUT=1, PT_MIN=0.5, DO_ETA_CUT=1, ETA_MIN=-2.0, ETA_MAX=2.0,
             ZMIN=-1, ZMAX=1, MMIN=3.56, MMAX=50, YMIN=-6, YMAX=6, BINS_Z=20, BINS_M=21, BINS_Y=9,
             FLUX_POINT=1, BREAKUP_MODE=1, NON_ZERO_GAM_PT=1, USE_POLARIZED_CS=0, SEED=777,
             USE_ROOT_OUTPUT=0, USE_HEPMC_OUTPUT=1)


def write_params(path, **over):
    p = dict(SMALL)
    p.update(over)
    path.write_text("".join(f"{k} {v}\n" for k, v in p.items()))
    return p


def run(cwd, *extra, timeout=240):
    r = subprocess.run([EXE, "-parfile", "p.in"] + list(extra), cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def logged(out, key):
    m = re.search(key + r" ([0-9.eE+-]+)", out)
    return float(m.group(1))


def oracle_tot(oracle_ctx, **kw):
    o = oracle_ctx(**kw)
    g, nbad = o.lumi_grid()
    assert nbad == 0
    _, _, tot = o.nuc_cs_ym(lumi=g)
    return tot


def parse_hepmc(path):
    events, cur = [], None
    lines = path.read_text().splitlines()
    assert lines[0].startswith("HepMC::Version") and lines[1] == "HepMC::Asciiv3-START_EVENT_LISTING"
    assert lines[-1] == "HepMC::Asciiv3-END_EVENT_LISTING"
    for ln in lines[2:-1]:
        t = ln.split()
        if t[0] == "E":
            cur = []
            events.append(cur)
        elif t[0] == "P":
            cur.append((int(t[3]),) + tuple(float(x) for x in t[4:9]))
    return events


def test_cli_end_to_end_matches_oracle(tmp_path, oracle_ctx):
    write_params(tmp_path / "p.in")
    rc, out = run(tmp_path)
    assert rc == 0, out[-3000:]
    assert (tmp_path / "twoPhotonLumi.upcbin").exists()
    assert not (tmp_path / ".lumiIsCalculated").exists()
    tot = logged(out, "totCS_mb")
    ref = oracle_tot(oracle_ctx, flux_point=1, breakup_mode=1, nm=21, ny=9, nz=20, proc_id=13)
    assert tot == pytest.approx(ref, rel=1e-9)
    fid = logged(out, "fidCS_mb")
    assert 0 < fid < tot
    ev = parse_hepmc(tmp_path / "events.hepmc")
    assert len(ev) == 3000
    for parts in ev[:500]:
        assert len(parts) == 2
        (pdg1, *p1), (pdg2, *p2) = parts
        assert {pdg1, pdg2} == {13, -13}
        for px, py, pz, e, m in (p1, p2):
            assert m == pytest.approx(0.1056583745, rel=1e-5)
            pt = math.hypot(px, py)
            assert pt >= 0.5 and abs(math.asinh(pz / pt)) <= 2.0
        tot4 = [a + b for a, b in zip(p1[:4], p2[:4])]
        mpair = math.sqrt(max(tot4[3] ** 2 - tot4[0] ** 2 - tot4[1] ** 2 - tot4[2] ** 2, 0))
        assert 3.56 - 1e-9 <= mpair <= 50 + 1e-9
        assert math.hypot(tot4[0], tot4[1]) < 2 * 6 * 0.1973269718 / 6.68 + 1e-9


def test_cache_hit_and_key_mismatch(tmp_path):
    write_params(tmp_path / "p.in", NEVENTS=10, USE_HEPMC_OUTPUT=0)
    rc, out = run(tmp_path)
    assert rc == 0 and "Starting all over" in out
    tot1 = logged(out, "totCS_mb")
    rc, out = run(tmp_path)
    assert rc == 0 and "Found pre-calculated unpolarized 2D luminosity" in out
    assert logged(out, "totCS_mb") == tot1
    write_params(tmp_path / "p.in", NEVENTS=10, USE_HEPMC_OUTPUT=0, BINS_Y=11)
    rc, out = run(tmp_path)
    assert rc == 0 and "different parameters" in out and "Starting all over" in out


def test_lock_file_wait_and_stale_takeover(tmp_path):
    write_params(tmp_path / "p.in", NEVENTS=10, USE_HEPMC_OUTPUT=0)
    lock = tmp_path / ".lumiIsCalculated"
    lock.write_text("")
    p = subprocess.Popen([EXE, "-parfile", "p.in"], cwd=tmp_path, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True)
    time.sleep(3)
    assert p.poll() is None, "the generator must wait while another holds the lock"
    lock.unlink()
    out, _ = p.communicate(timeout=240)
    assert p.returncode == 0 and "totCS_mb" in out
    # a lock older than the timeout is taken over
    lock.write_text("")
    old = time.time() - 7200
    os.utime(lock, (old, old))
    rc, out = run(tmp_path, "-locktimeout", "60")
    assert rc == 0 and "taking it over" in out
    assert not lock.exists()


def test_polarised_generator(tmp_path, oracle_ctx):
    write_params(tmp_path / "p.in", USE_POLARIZED_CS=1, PROC_ID=15, FLUX_POINT=0, BREAKUP_MODE=2, BINS_M=7,
                 BINS_Y=5, BINS_Z=10, NEVENTS=500, DO_PT_CUT=0, DO_ETA_CUT=0)
    rc, out = run(tmp_path)
    assert rc == 0, out[-3000:]
    assert (tmp_path / "twoPhotonLumiPol.upcbin").exists()
    o = oracle_ctx(flux_point=0, breakup_mode=2, use_polarized=1, nm=7, ny=5, nz=10, proc_id=15)
    (s, p), nbad = o.lumi_grid()
    _, _, ref = o.nuc_cs_ym(lumi_s=s, lumi_p=p)
    assert logged(out, "totCS_mb") == pytest.approx(ref, rel=1e-9)
    ev = parse_hepmc(tmp_path / "events.hepmc")
    assert len(ev) == 500 and all({a[0] for a in e} == {15, -15} for e in ev)


def test_photon_pt_distribution(gpu_ctx, oracle_ctx):
    """getPhotonPt draws follow the reference's 5000-bin histogram (chi-square
    on 50 merged bins) at photon energies across the grid."""
    import torch
    from scipy.stats import chi2
    g = gpu_ctx(flux_point=0, breakup_mode=1, nm=3, ny=3)
    o = oracle_ctx(flux_point=0, breakup_mode=1, nm=3, ny=3)
    gtot = 5020 / (2 * 0.9382720813)
    g.photon_pt_tables(gtot)
    n = 400_000
    ptmax = 6 * 0.1973269718 / 6.68
    for E in (0.005, 1.0, 40.0, 300.0, 2500.0):
        pts = g.photon_pt(torch.full((n,), E, dtype=torch.float64, device="cuda"), seed=99).cpu().numpy()
        assert pts.min() >= 0 and pts.max() <= ptmax
        w = o.photon_pt_hist(E, gtot)
        p = w.reshape(50, 100).sum(axis=1) / w.sum()
        h, _ = np.histogram(pts, bins=50, range=(0, ptmax))
        keep = p * n > 20
        stat = np.sum((h[keep] - n * p[keep]) ** 2 / (n * p[keep]))
        assert stat < chi2.ppf(0.9999, keep.sum() - 1), (E, stat)


def test_pair_kinematics(gpu_ctx):
    import torch
    import upcgen_amd as U
    g = gpu_ctx(flux_point=1, breakup_mode=1, nm=21, ny=9, nz=20, proc_id=15)
    L = g.lumi_grid()
    cs, _, _ = g.nuc_cs_ym(L)
    sum_ym = g.pdf_init(cs.reshape(1, -1))[0].contiguous()
    sum_z = g.pdf_init(g.cs_zm(0)).contiguous()
    opts = U.EventOpts(0, 1, 1.0, 1, -1.5, 1.5)
    p, pdg, ok = g.generate_pairs(sum_ym, sum_z=sum_z, n=200_000, seed=5, opts=opts)
    p, pdg, ok = p.cpu().numpy(), pdg.cpu().numpy(), ok.cpu().numpy()
    assert set(np.unique(ok)) <= {1, 2}
    m2 = p[..., 3] ** 2 - (p[..., :3] ** 2).sum(-1)
    assert np.allclose(np.sqrt(m2), 1.77686, rtol=1e-6)
    assert np.all(pdg[:, 0] == -pdg[:, 1]) and set(np.abs(pdg).ravel()) == {15}
    pair = p.sum(axis=1)
    # without photon pt the pair moves along the beam: pt = 0, M and Y from the tables
    assert np.max(np.abs(pair[:, :2])) < 1e-9
    mp = np.sqrt(pair[:, 3] ** 2 - pair[:, 2] ** 2)
    assert mp.min() >= 3.56 - 1e-9 and mp.max() <= 50 + 1e-9
    pt = np.hypot(p[..., 0], p[..., 1])
    eta = np.arcsinh(p[..., 2] / pt)
    passed = (pt >= 1.0).all(axis=1) & (np.abs(eta) <= 1.5).all(axis=1)
    assert np.array_equal(passed, ok == 1)


def test_cli_mfma_form(tmp_path):
    """-form mfma: the same run with the MFMA luminosity; total cross section
    within the MFMA form's 1e-5 of the FP64 run."""
    write_params(tmp_path / "p.in", NEVENTS=10, USE_HEPMC_OUTPUT=0, FLUX_POINT=0, BREAKUP_MODE=2)
    rc, out = run(tmp_path)
    assert rc == 0, out[-2000:]
    tot64 = logged(out, "totCS_mb")
    # the cache is keyed by the form: the MFMA run does not reuse the FP64 grid
    rc, out = run(tmp_path, "-form", "mfma")
    assert rc == 0, out[-2000:]
    assert "different parameters" in out and "Starting all over" in out
    tot = logged(out, "totCS_mb")
    assert tot == pytest.approx(tot64, rel=1e-5) and tot != tot64


def test_facade_members_beyond_the_cli(oracle_ctx):
    """calcPhotonFlux, calcNucCrossSectionY (with a test elementary sigma(y)),
    getPhotonPt and getPairMomentum of the C++ facade (driven by
    tests/facade_members.cpp) against the oracle: the vector-meson photon flux
    and sigma(y) to 1e-9, the photon-pt draws' mean against the reference
    histogram's (5 standard errors), and the pair kinematics exactly."""
    import json
    exe = os.path.join(ROOT, "tests", "facade_members")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    d = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    o = oracle_ctx(flux_point=0, breakup_mode=2, nm=3, ny=9, ymin=-4.0, ymax=4.0)
    for y, v in d["photon_flux"]:
        ref = o.photon_flux(3.0969, y)
        assert v == pytest.approx(ref, rel=1e-9), y
    ys = -4.0 + 8.0 / 9 * np.arange(9)
    sig = lambda y: 1.0 + 0.1 * y + 0.01 * y * y
    cs, ratio, tot = o.nuc_cs_y(3.0969, sig(ys), sig(-ys))
    assert np.allclose(d["cs_y"], cs, rtol=1e-9, atol=0) and np.allclose(d["ratio_y"], ratio, rtol=1e-9, atol=0)
    assert d["tot_y"] == pytest.approx(tot, rel=1e-9)
    w = o.photon_pt_hist(2.0, d["gtot"])
    ptmax = 6 * 0.1973269718 / 6.68
    centers = (np.arange(5000) + 0.5) * ptmax / 5000
    mean = np.sum(w * centers) / w.sum()
    rms = np.sqrt(np.sum(w * centers ** 2) / w.sum())
    print(f"photon pt at 2 GeV: draws mean {d['pt_mean']:.5f} rms {d['pt_rms']:.5f}; "
          f"histogram {mean:.5f} {rms:.5f}")
    assert abs(d["pt_mean"] - mean) < 5 * rms / np.sqrt(4000)
    assert d["pair_max_dm"] < 1e-9 and d["pair_max_dy"] < 1e-9
    assert d["pair_mean_pt"] > 0 and d["pair_pt_zero"] == 0.0
