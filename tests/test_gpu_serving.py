"""Serving (continuous batching): every request's result is the same alone and in the mix, the
HBM budget is respected by admission, and the moments stay consistent."""
import numpy as np
import pytest

from dtx import serving

pytestmark = pytest.mark.gpu


def small_mix():
    # shapes drawn like the serving workload but small enough for a unit test
    specs = serving.request_mix(count=8, rate_per_s=1e6, steps_range=(20, 60))
    for s in specs:
        s.p_num = min(s.p_num, 16)
        s.x_num = min(s.x_num, 32)
    return specs


@pytest.mark.parametrize("precision", ["fp32", "fp64", "bf16x3"])
def test_alone_equals_mixed_bitwise(precision):
    specs = small_mix()
    mixed, _, f_mixed = serving.run(specs, precision=precision, realtime=False, fetch=True)
    for i, s in enumerate(specs):
        alone, _, f_alone = serving.run([s], precision=precision, realtime=False, fetch=True)
        np.testing.assert_array_equal(f_alone[0], f_mixed[i])
        assert alone[0].number_total == mixed[i].number_total
        assert mixed[i].steps_done == s.steps


def test_budget_limits_concurrency_but_all_complete():
    specs = small_mix()[:4]
    res, _, f_small_budget = serving.run(specs, precision="fp32", hbm_budget_gb=0.001,
                                         realtime=False, fetch=True)
    res2, _, f_big = serving.run(specs, precision="fp32", realtime=False, fetch=True)
    for a, b in zip(f_small_budget, f_big):
        np.testing.assert_array_equal(a, b)
    # with a tiny budget only one request runs at a time: each starts after the previous ends
    order = sorted(res, key=lambda r: r.start_ms)
    for prev, nxt in zip(order, order[1:]):
        assert nxt.start_ms >= prev.finish_ms - 1.0


def test_number_conserved_per_request():
    """Collisions conserve number per input pair, periodic advection and closed-boundary
    cooling move particles only: the total stays at its initial value (f0 sums to one per
    species and cell, times the 1% noise)."""
    specs = small_mix()[:4]
    res, _, states = serving.run(specs, precision="fp64", realtime=False, fetch=True)
    for s, r, f in zip(specs, res, states):
        assert abs(r.number_total - f.sum()) <= 1e-12 * abs(f.sum())
        assert abs(f.sum() / (s.species * s.x_num) - 1.0) < 0.05


def test_large_request_in_group():
    """A request large enough for the two-level fold of the step partials (more than 16 fold
    groups: 520 cells x 8 update segments), run alone and next to a small one: same state, and
    the device moments still sum the state."""
    big = serving.RequestSpec(p_num=32, u_num=8, h_num=1, x_num=520, species=2, steps=12,
                              seed=11)
    small = small_mix()[0]
    mixed, _, f_mixed = serving.run([big, small], precision="fp64", realtime=False, fetch=True)
    alone, _, f_alone = serving.run([big], precision="fp64", realtime=False, fetch=True)
    np.testing.assert_array_equal(f_alone[0], f_mixed[0])
    assert alone[0].number_total == mixed[0].number_total
    f = f_alone[0]
    assert abs(alone[0].number_total - f.sum()) <= 1e-12 * abs(f.sum())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_wide_items_in_group(precision):
    """128 cells: FP64 gain items take two k tiles x 128 cells; alone and mixed with a
    64-cell-tile request the state is the same."""
    req = serving.RequestSpec(p_num=16, u_num=8, h_num=1, x_num=128, species=2, steps=10,
                              seed=5)
    small = small_mix()[1]
    mixed, _, f_mixed = serving.run([req, small], precision=precision, realtime=False,
                                    fetch=True)
    alone, _, f_alone = serving.run([req], precision=precision, realtime=False, fetch=True)
    np.testing.assert_array_equal(f_alone[0], f_mixed[0])
    assert alone[0].number_total == mixed[0].number_total

