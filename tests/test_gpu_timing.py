"""Kernel timers: the sweep's event pairs selected by name and sampled by a
stride (the throughput windows of bench.py time every fourth sweep)."""
from math import ceil, gcd

import pytest

pytestmark = pytest.mark.gpu


def _sweep_counts(hmec, stride, steps=3):
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    ctx = hmec.Context(p, soa, precision=hmec.PREC_FP32)
    ctx.timing(True)
    ctx.timing_select("sweep_fwd,sweep_bwd")
    ctx.timing_sample(stride)
    ctx.timing_reset()
    for _ in range(steps):
        ctx.step()
    out = {}
    for name in ("sweep_fwd", "sweep_bwd"):
        n, ms, by = ctx.kernel_time(name)
        out[name] = (n, ms, by)
    ctx.close()
    return out


def test_timing_sample(hmec, gpu):
    """Two contexts on the same inputs run the same preconditioner
    applications; with stride 4 the timers cover the first, fifth, ... of
    them: launches counted = ceil(applications / 4) x launches per
    application, with the same per-launch averages in range."""
    all_ = _sweep_counts(hmec, 1)
    some = _sweep_counts(hmec, 4)
    for name in ("sweep_fwd", "sweep_bwd"):
        n1, ms1, _ = all_[name]
        n4, ms4, _ = some[name]
        assert n1 > 0 and 0 < n4 < n1, (name, n1, n4)
        assert ms1 > 0 and ms4 > 0, (name, ms1, ms4)
        g = gcd(n1, n4)
        ok = [L for L in range(1, g + 1)
              if g % L == 0 and n4 // L == ceil((n1 // L) / 4)]
        assert ok, (name, n1, n4)


def test_timing_sample_rejects_zero(hmec, gpu):
    p = hmec.params(1)
    soa = hmec.generate_markers(p, cfg=1, seed=1)
    ctx = hmec.Context(p, soa, precision=hmec.PREC_FP32)
    with pytest.raises(hmec.HmecError):
        ctx.timing_sample(0)
    ctx.close()
