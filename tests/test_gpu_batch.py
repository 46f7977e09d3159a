"""Batched luminosity grids (one persistent QAGS launch and one contraction
launch for many configurations, replayed from a hipGraph) against the
single-configuration path and the CPU oracle.

Bar: every job's grid equals the same configuration run alone to 1e-12
relative (the same arithmetic; the QAGS kernel instantiated for the batch
task source may contract a few FMAs differently, measured <= 5e-15, and
point-flux jobs are bit-identical); one job is also checked against the
oracle at the FP64 form's 1e-9.  Graph replays are bit-identical to the
eager batch run.
"""
import numpy as np
import pytest

import upcgen_amd as U

pytestmark = pytest.mark.gpu

JOBS = [
    dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0, flux_point=0, breakup_mode=2, use_polarized=1, nm=7, ny=5),
    dict(Z=54, A=129, R=5.4889, a=0.54, sqrts=5440.0, flux_point=0, breakup_mode=3, use_polarized=0, nm=6, ny=5),
    dict(Z=8, A=16, R=2.4809, a=0.54, sqrts=7000.0, flux_point=1, breakup_mode=4, use_polarized=1, nm=5, ny=7),
    dict(Z=79, A=197, R=6.3691, a=0.54, sqrts=200.0, flux_point=1, breakup_mode=1, use_polarized=1, nm=6, ny=5,
         mmin=1.0, mmax=10.0, ymin=-3.0, ymax=3.0),
    dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0, flux_point=0, breakup_mode=2, use_polarized=0, nm=6, ny=4,
         ymin=-5.0, ymax=3.0),
    dict(Z=82, A=208, R=6.68, a=0.447, sqrts=5020.0, flux_point=0, breakup_mode=1, use_polarized=0, nm=9, ny=7),
]


def _np(t):
    return t.detach().cpu().numpy()


def _grids(outs, cfgs):
    res = []
    for o, c in zip(outs, cfgs):
        res.append((_np(o[0]), _np(o[1])) if c.use_polarized else _np(o))
    return res


def _close(a, b, tol=1e-12):
    if isinstance(a, tuple):
        return all(_close(x, y, tol) for x, y in zip(a, b))
    return np.max(np.abs(a - b) / np.abs(b)) <= tol


def _same(a, b):
    if isinstance(a, tuple):
        return all(np.array_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b)


@pytest.fixture(scope="module")
def batch_and_refs(gpu):
    cfgs = [U.Config.make(**kw) for kw in JOBS]
    refs, calls = [], 0
    for c in cfgs:
        ctx = U.Context(c)
        out = ctx.lumi_grid()
        refs.append((_np(out[0]), _np(out[1])) if c.use_polarized else _np(out))
        calls += ctx.info.qags_calls
        ctx.close()
    return U.Batch(cfgs), cfgs, refs, calls


def test_batch_matches_single_contexts(batch_and_refs):
    import torch
    b, cfgs, refs, calls = batch_and_refs
    got = _grids(b.run(graph=False), cfgs)
    torch.cuda.synchronize()
    for j, (g, r, c) in enumerate(zip(got, refs, cfgs)):
        assert _close(g, r), f"job {j}"
        if c.flux_point:
            assert _same(g, r), f"job {j}"
    st = b.stats()
    assert st.njobs == len(cfgs) and st.qags_calls == calls and st.qags_bad == 0
    assert st.points == sum(c.nm * c.ny for c in cfgs)


def test_batch_graph_replays(batch_and_refs):
    import torch
    b, cfgs, _, _ = batch_and_refs
    refs = _grids(b.run(graph=False), cfgs)
    torch.cuda.synchronize()
    for rep in range(3):
        for o in b.outputs:  # poison the outputs between replays
            for t in (o if isinstance(o, tuple) else (o,)):
                t.fill_(-1.0)
        x = torch.randn(1 << 16, device="cuda")  # unrelated eager work between replays
        (x * 2).sum().item()
        got = _grids(b.run(graph=True), cfgs)
        torch.cuda.synchronize()
        for j, (g, r) in enumerate(zip(got, refs)):
            assert _same(g, r), f"replay {rep} job {j}"


def test_batch_autotune(batch_and_refs):
    import torch
    b, cfgs, _, _ = batch_and_refs
    refs = _grids(b.run(graph=False), cfgs)
    torch.cuda.synchronize()
    t = b.autotune()
    assert t.qags_blocks_per_cu in list(t.qags_candidates)
    assert t.lumi_waves in list(t.lumi_candidates)
    assert all(ms > 0 for ms in t.qags_ms) and all(ms > 0 for ms in t.lumi_ms)
    for qb in (1, 3):
        for lw in (2, 8):
            b.set_tune(qb, lw)
            got = _grids(b.run(graph=True), cfgs)
            torch.cuda.synchronize()
            assert all(_same(g, r) for g, r in zip(got, refs)), (qb, lw)
            st = b.stats()
            assert (st.qags_blocks_per_cu, st.lumi_waves) == (qb, lw)
    with pytest.raises(U.UpcError):
        b.set_tune(5, 4)


@pytest.mark.parametrize("form", [U.FORM_MFMA, U.FORM_MFMA_FP32])
def test_batch_mfma_forms(batch_and_refs, form):
    """The batch's MFMA forms: every job's grid is the one its own context
    computes in that form (1e-12; point-flux jobs bit-identical), each tile
    shape gives the same bits, graph replays equal the eager run, and the
    grids stay within the MFMA form's accuracy of the FP64 grids (1e-5,
    3e-5 for the 1-10 GeV Au-Au job as in the single-context MFMA tests)."""
    import torch
    b, cfgs, refs, _ = batch_and_refs
    singles = []
    for c in cfgs:
        ctx = U.Context(c)
        out = ctx.lumi_grid(form=form)
        singles.append((_np(out[0]), _np(out[1])) if c.use_polarized else _np(out))
        ctx.close()
    try:
        b.set_form(form)
        first = None
        for tg in (1, 2, 4):
            b.set_mfma_tiles(tg)
            eager = _grids(b.run(graph=False), cfgs)
            torch.cuda.synchronize()
            for o in b.outputs:
                for t in (o if isinstance(o, tuple) else (o,)):
                    t.fill_(-1.0)
            replay = _grids(b.run(graph=True), cfgs)
            torch.cuda.synchronize()
            st = b.stats()
            assert (st.form, st.mfma_tile_group) == (form, tg)
            for j, (e, r, s1, ref, c) in enumerate(zip(eager, replay, singles, refs, cfgs)):
                assert _same(e, r), (tg, j)
                assert _close(e, s1), (tg, j)
                if c.flux_point:
                    assert _same(e, s1), (tg, j)
                assert _close(e, ref, 3e-5 if j == 3 else 1e-5), (tg, j)
            if first is None:
                first = eager
            else:
                assert all(_same(g, r) for g, r in zip(eager, first)), tg
        t = b.autotune()
        assert t.form == form and t.mfma_tile_group in list(t.mfma_candidates)
        assert sorted(t.mfma_candidates) == [1, 2, 4] and all(ms > 0 for ms in t.mfma_ms)
        assert b.stats().mfma_tile_group == t.mfma_tile_group
        got = _grids(b.run(graph=True), cfgs)
        torch.cuda.synchronize()
        assert all(_same(g, r) for g, r in zip(got, first))
        with pytest.raises(U.UpcError):
            b.set_mfma_tiles(3)
        with pytest.raises(U.UpcError):
            b.set_form(U.FORM_SURROGATE)
    finally:
        b.set_form(U.FORM_FP64)
    got = _grids(b.run(graph=True), cfgs)
    torch.cuda.synchronize()
    assert all(_close(g, r) for g, r in zip(got, refs))


def test_batch_job_against_oracle(batch_and_refs, oracle_ctx):
    b, cfgs, refs, _ = batch_and_refs
    kw = dict(JOBS[1])
    ref, nbad = oracle_ctx(**kw).lumi_grid()
    assert nbad == 0
    got = _grids(b.run(graph=True), cfgs)[1]
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-9


def test_batch_mfma_cancellation_guard(gpu):
    """A batch with an Au-Au 200 GeV form-factor job over the full default M
    range, where the MFMA product cancels most of the far-field term at high
    M: the guard recomputes those work items with the FP64 contraction
    (counted in the batch stats) and every bin of both jobs stays within
    1e-5 of the FP64 batch."""
    import torch
    cfgs = [U.Config.make(Z=79, A=197, R=6.3691, a=0.54, sqrts=200.0, flux_point=0, breakup_mode=2,
                          use_polarized=1, nm=101, ny=41),
            U.Config.make(flux_point=0, breakup_mode=2, use_polarized=0, nm=51, ny=21)]
    b = U.Batch(cfgs)
    try:
        ref = _grids(b.run(graph=False), cfgs)
        torch.cuda.synchronize()
        b.set_form(U.FORM_MFMA)
        got = _grids(b.run(graph=True), cfgs)
        torch.cuda.synchronize()
        st = b.stats()
        print(f"work items recomputed in FP64: {st.mfma_fp64_items}")
        assert st.mfma_fp64_items > 0
        for g, r in zip(got, ref):
            for x, y in (zip(g, r) if isinstance(g, tuple) else ((g, r),)):
                fin = np.isfinite(y)
                assert np.array_equal(np.isfinite(x), fin)
                assert np.max(np.abs(x[fin] - y[fin]) / np.abs(y[fin])) <= 1e-5
    finally:
        b.close()
