"""The oracle's liquid-volume accounting closes — CPU only.

Over a few dam-break steps the change of sum(F vol) equals the sum of the
accounted terms (stage-1 clip, stage-4 clamp01, sweep dilation,
post_process_f's clamp01) up to the rounding of the sums; with the
instrumentation on, the step's bits are unchanged.  The study itself is
tests/conservation_study.py.
"""
import numpy as np
import pytest

from conservation_study import account
from oracle import Oracle
from vof import default_params


def test_volume_accounting_closes_fp64():
    """In FP64 the terms add up to the change to round-off of the sums."""
    r = account(24, 30, 8, 10, init="box", dt=2e-4)
    assert abs(r["rest"]) <= 1e-8 * abs(r["change"]), r
    # a dam break at this time step is dominated by the dilation of the
    # not-quite-divergence-free velocity (10 Jacobi sweeps), to first order
    # dt vol sum F div u
    assert abs(r["dil"] - r["div"]) < 0.1 * abs(r["dil"])


def test_volume_accounting_fp32_rounding_is_a_term_of_its_own():
    """In FP32 the sweeps' own arithmetic rounds each cell's new value, so the
    telescoping sums hold only to that rounding: 'rest' is nonzero but of
    the order of the stage-4 rounding term, and far below the dilation."""
    r = account(24, 30, 4, 10, init="box", dt=2e-4)
    assert r["rest"] != 0.0
    assert abs(r["rest"]) < 5 * abs(r["round4"]) + 1e-9
    assert abs(r["rest"]) < 0.1 * abs(r["dil"])


def test_accounting_does_not_change_the_step():
    kw = dict(n=20, precision=4, g=(0.0, 0.0, -9.81), dt=2e-4)
    a, b = Oracle(default_params(**kw)), Oracle(default_params(**kw))
    for o in (a, b):
        o.init_box(0.04, 0.04, 0.06)
    b.stats_enable(True)
    a.step(1, 6)
    b.step(1, 6)
    for f in ("F", "U", "V", "W", "P"):
        assert np.array_equal(a.get(f), b.get(f)), f
