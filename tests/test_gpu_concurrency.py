"""Contexts are independent: two contexts driven at the same time on two
streams (table rebuilds and grids in flight together, both forms) give the
same bits and the same QAGS counts as each context run alone.  Each context
owns its work buffers, its QAGS overflow workspace and the side streams of
its table build."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW_A = dict(flux_point=0, breakup_mode=2, nm=41, ny=21)
KW_B = dict(Z=54, A=129, R=5.4889, a=0.54, sqrts=5440.0, flux_point=0, breakup_mode=3, use_polarized=1,
            nm=31, ny=16)


def _bits(x):
    xs = x if isinstance(x, tuple) else (x,)
    return [t.cpu().numpy().view(np.uint64).copy() for t in xs]


def test_two_contexts_on_two_streams(gpu):
    import torch
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    with gpu.Context(stream=sa, **KW_A) as a, gpu.Context(stream=sb, **KW_B) as b:
        alone = {}
        for name, ctx, s in (("a", a, sa), ("b", b, sb)):
            for form in (gpu.FORM_FP64, gpu.FORM_MFMA):
                with torch.cuda.stream(s):
                    out = ctx.lumi_grid(form, stream=s)
                torch.cuda.synchronize()
                alone[name, form] = (_bits(out), ctx.info.qags_evals)
        for rnd in range(3):
            for form in (gpu.FORM_FP64, gpu.FORM_MFMA):
                with torch.cuda.stream(sa):
                    a.rebuild_tables(stream=sa)
                    out_a = a.lumi_grid(form, stream=sa)
                with torch.cuda.stream(sb):
                    b.rebuild_tables(stream=sb)
                    out_b = b.lumi_grid(form, stream=sb)
                torch.cuda.synchronize()
                for name, ctx, out in (("a", a, out_a), ("b", b, out_b)):
                    bits, evals = alone[name, form]
                    got = _bits(out)
                    assert all(np.array_equal(g, w) for g, w in zip(got, bits)), (rnd, name, form)
                    assert ctx.info.qags_evals == evals, (rnd, name, form)
