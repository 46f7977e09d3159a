"""Host-side logic of the mixed-systems batch (no GPU needed)."""
import pytest

import upcgen_amd as U


def test_mixed_systems_jobs_spec():
    jobs = U.mixed_systems_jobs(seed=7)
    assert len(jobs) == 64
    combos = {(c.Z, c.breakup_mode) for c in jobs}
    assert len(combos) == 16
    assert all(c.use_polarized == 1 for c in jobs)
    assert {c.nm for c in jobs} <= {101, 251, 501, 1001, 2001} and {c.ny for c in jobs} <= {41, 81, 121}
    pb = [c for c in jobs if c.Z == 82][0]
    assert pb.a == 0.447 and pb.R == pytest.approx(6.4908, abs=1e-4)
