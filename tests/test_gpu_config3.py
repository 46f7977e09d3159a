"""Config 3 at its stated workload: the 64-job mixed batch
upcgen_amd.mixed_systems_jobs(seed=2024) -- Pb-Pb 5.02 TeV, Xe-Xe 5.44 TeV,
O-O 7 TeV, Au-Au 200 GeV x breakup modes 1-4, USE_POLARIZED_CS 1,
form-factor flux, M bins in {101, ..., 2001}, Y bins in {41, 81, 121};
3,844,184 (M, Y) points -- in one graph-replayed launch sequence.

Bars:
  - every job's (L_s, L_p) equals the same configuration run alone through
    upc_lumi_grid to 1e-12 relative (NaN in the same bins);
  - eight bins of every job (its corners and four drawn ones) equal the
    oracle's calcTwoPhotonLumiPol (tests/golden/oracle_config3.npz) to 1e-9;
  - the per-job count of integrals that end with a QAGS error or a
    form-factor domain error equals the single-context count, and is printed
    with the job's system and mode (the reference process stops at the first);
  - O-O 7 TeV and Au-Au 200 GeV 9 x 7 polarised form-factor grids in modes
    1-4 equal the oracle's full grids to 1e-9 (src/UpcCrossSection.cpp:194-218,
    274-335, 853-969).
"""
import os

import numpy as np
import pytest

import upcgen_amd as U

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "oracle_config3.npz")
SYSNAME = {82: "Pb-Pb 5.02 TeV", 54: "Xe-Xe 5.44 TeV", 8: "O-O 7 TeV", 79: "Au-Au 200 GeV"}


def _rel(a, b):
    nan = np.isnan(b)
    assert np.array_equal(nan, np.isnan(a))
    return float(np.max(np.abs(a[~nan] - b[~nan]) / np.abs(b[~nan]), initial=0.0))


@pytest.fixture(scope="module")
def batch_run(gpu):
    import torch
    cfgs = U.mixed_systems_jobs(seed=2024)
    b = U.Batch(cfgs)
    outs = b.run(graph=True)
    torch.cuda.synchronize()
    grids = [(s.cpu().numpy(), p.cpu().numpy()) for s, p in outs]
    st = b.stats()
    failed = b.job_failures()
    b.close()
    return cfgs, grids, st, failed


def test_batch_workload_size(batch_run):
    cfgs, grids, st, failed = batch_run
    assert len(cfgs) == 64 and st.points == 3_844_184
    print(f"batch: {st.points} points, {st.rows} flux rows, {st.qags_calls} integrals, "
          f"{st.qags_evals} evaluations, {st.qags_bad} failed")
    assert sum(failed) == st.qags_bad


def test_every_job_equals_its_single_context(batch_run):
    cfgs, grids, st, failed = batch_run
    worst = 0.0
    for j, c in enumerate(cfgs):
        ctx = U.Context(c)
        ls, lp = ctx.lumi_grid()
        ref = (ls.cpu().numpy(), lp.cpu().numpy())
        single_failed = ctx.info.qags_bad
        ctx.close()
        e = max(_rel(grids[j][0], ref[0]), _rel(grids[j][1], ref[1]))
        worst = max(worst, e)
        if failed[j] or single_failed:
            print(f"job {j:2d} {SYSNAME[c.Z]} mode {c.breakup_mode} {c.nm}x{c.ny}: {failed[j]} failed integrals, "
                  f"{int(np.isnan(grids[j][0]).sum())} NaN bins")
        assert failed[j] == single_failed, j
        assert e <= 1e-12, (j, e)
    print(f"worst batch vs single-context difference {worst:.2e}")


def test_sampled_bins_of_every_job_against_oracle(batch_run):
    cfgs, grids, st, failed = batch_run
    g = np.load(GOLD)
    got_s = np.array([grids[j][0][i, k] for j, i, k in zip(g["job"], g["im"], g["iy"])])
    got_p = np.array([grids[j][1][i, k] for j, i, k in zip(g["job"], g["im"], g["iy"])])
    assert set(g["job"].tolist()) == set(range(64))
    es, ep = _rel(got_s, g["Ls"]), _rel(got_p, g["Lp"])
    print(f"{len(got_s)} oracle bins: max rel err s {es:.2e}, p {ep:.2e}")
    assert es <= 1e-9 and ep <= 1e-9


@pytest.mark.parametrize("name", ["OO", "AuAu"])
@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_light_and_low_energy_systems_full_grids(gpu_ctx, name, mode):
    g = np.load(GOLD)
    sysp = dict(Z=8, A=16, R=U.wood_saxon_radius(16), a=0.54, sqrts=7000.0) if name == "OO" else \
        dict(Z=79, A=197, R=U.wood_saxon_radius(197), a=0.54, sqrts=200.0)
    ctx = gpu_ctx(flux_point=0, breakup_mode=mode, use_polarized=1, nm=9, ny=7, **sysp)
    ls, lp = ctx.lumi_grid()
    es = _rel(ls.cpu().numpy(), g[f"{name}_mode{mode}_s"])
    ep = _rel(lp.cpu().numpy(), g[f"{name}_mode{mode}_p"])
    inf = ctx.info
    print(f"{name} mode {mode}: max rel err s {es:.2e}, p {ep:.2e}; failed {inf.qags_bad} "
          f"(oracle {int(g[f'{name}_mode{mode}_bad'])}), evaluations {inf.qags_evals} "
          f"(oracle {int(g[f'{name}_mode{mode}_evals'])})")
    assert inf.qags_bad == int(g[f"{name}_mode{mode}_bad"])
    assert inf.qags_evals == int(g[f"{name}_mode{mode}_evals"])
    assert es <= 1e-9 and ep <= 1e-9
