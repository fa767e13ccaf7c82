"""Serving (continuous batching): every request's result is the same alone, in the mix, and
stepped on its own through dtx_step (bitwise); a caller-built simulation served in a mix agrees
with the oracle; the FP64 moment-rate correction holds in grouped FP32 / split-BF16 steps; the
HBM budget is respected by admission; the real request shapes of the serving workload run."""
import numpy as np
import pytest

import oracle as O
from dtx import serving, workloads

pytestmark = pytest.mark.gpu


def small_mix(count=8, steps=(20, 60)):
    # shapes drawn like the serving workload but small enough for a unit test
    specs = serving.request_mix(count=count, rate_per_s=1e6, steps_range=steps)
    for s in specs:
        s.p_num = min(s.p_num, 16)
        s.x_num = min(s.x_num, 32)
    return specs


@pytest.mark.parametrize("precision", ["fp32", "fp64", "bf16x3"])
def test_alone_equals_mixed_bitwise(precision):
    specs = small_mix()
    mixed, _, f_mixed = serving.run(specs, precision=precision, realtime=False, fetch=True)
    for i, s in enumerate(specs[:4]):
        alone, _, f_alone = serving.run([s], precision=precision, realtime=False, fetch=True)
        np.testing.assert_array_equal(f_alone[0], f_mixed[i])
        assert alone[0].number_total == mixed[i].number_total
        assert mixed[i].steps_done == s.steps


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_served_equals_single_handle_steps(precision):
    """The served state equals the request's own simulation (dtx_request_create) stepped with
    dtx_step, bit for bit, with the moment-rate correction in the reduced-precision modes."""
    specs = small_mix()
    _, _, served = serving.run(specs, precision=precision, realtime=False, fetch=True)
    for i in (0, 3, 5):
        h = serving.RequestHandle(specs[i], precision=precision)
        h.step(specs[i].steps)
        np.testing.assert_array_equal(h.f, served[i])
        h.close()


@pytest.mark.parametrize("emission", [False, True], ids=["no_emission", "emission"])
def test_caller_built_simulation_matches_oracle(emission):
    """A simulation built by the caller (host flux CSC, host gain / loss channels, f0 with 1%
    noise, and per-cell emission blocks in the second case) is served next to synthetic
    requests; its final state agrees with the oracle at 1e-12 and equals the same simulation
    stepped alone bit for bit."""
    w = workloads.transport_1d(p_num=8, u_num=4, x_num=24, steps=40, emission=emission)
    m = workloads.make_struct(w, precision="fp64", use_monolithic=False)
    alone = workloads.make_struct(w, precision="fp64", use_monolithic=False)
    t = 0.0
    for _ in range(w.steps):
        alone(t, 1e300, w.dt, 0)
        t += w.dt
    f_alone = alone.f
    alone.close()
    srv = serving.Server("fp64")
    specs = small_mix(count=4)
    ids = [srv.submit(s) for s in specs[:2]]
    mine = srv.submit_handle(m, w.steps, w.dt)
    ids += [srv.submit(s) for s in specs[2:]]
    srv.drain()
    res = srv.poll(mine)
    assert res.status == 2 and res.steps_done == w.steps
    f_served = srv.fetch(mine, w.ps.n)
    assert all(srv.poll(i).status == 2 for i in ids)
    srv.close()
    np.testing.assert_array_equal(f_served, f_alone)
    prob = O.problem_from_workload(w, use_monolithic=False, batched=True)
    st = O.Stepper(prob, w.f0)
    t = 0.0
    for _ in range(w.steps):
        st(t, 1e300, w.dt, 0)
        t += w.dt
    fo = st.f
    assert np.max(np.abs(f_served - fo)) <= 1e-12 * np.max(np.abs(fo))
    assert abs(res.number_total - f_served.sum()) <= 1e-12 * f_served.sum()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_grouped_correction_holds_number_and_energy(precision):
    """Grouped FP32 / split-BF16 steps with the FP64 moment-rate correction: number and energy
    of every request follow the FP64 run to < 1e-9 over hundreds of steps, where the
    uncorrected grouped steps drift measurably."""
    specs = small_mix(count=6, steps=(300, 400))
    r64, _, _ = serving.run(specs, precision="fp64", realtime=False)
    rc, _, _ = serving.run(specs, precision=precision, realtime=False)
    ru, _, _ = serving.run(specs, precision=precision, realtime=False, moment_correction=False)
    worst_c = worst_u = 0.0
    for a, b, u in zip(r64, rc, ru):
        for x64, xc, xu in ((a.number_total, b.number_total, u.number_total),
                            (a.energy_total, b.energy_total, u.energy_total)):
            worst_c = max(worst_c, abs(xc - x64) / abs(x64))
            worst_u = max(worst_u, abs(xu - x64) / abs(x64))
    assert worst_c < 1e-9, worst_c
    assert worst_u > worst_c


def test_budget_limits_concurrency_but_all_complete():
    specs = small_mix()[:4]
    res, _, f_small_budget = serving.run(specs, precision="fp32", hbm_budget_gb=0.001,
                                         realtime=False, fetch=True)
    res2, _, f_big = serving.run(specs, precision="fp32", realtime=False, fetch=True)
    for a, b in zip(f_small_budget, f_big):
        np.testing.assert_array_equal(a, b)
    # with a tiny budget only one request is admitted at a time: each starts after the
    # previous one has finished and released its memory
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


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_wide_items_in_group(precision):
    """128 cells: FP64 and split-BF16 gain items take k tiles x 128 cells; alone and mixed
    with a 64-cell-tile request the state is the same, and equal to its own handle's steps."""
    req = serving.RequestSpec(p_num=16, u_num=8, h_num=1, x_num=128, species=2, steps=10,
                              seed=5)
    small = small_mix()[1]
    mixed, _, f_mixed = serving.run([req, small], precision=precision, realtime=False,
                                    fetch=True)
    alone, _, f_alone = serving.run([req], precision=precision, realtime=False, fetch=True)
    np.testing.assert_array_equal(f_alone[0], f_mixed[0])
    assert alone[0].number_total == mixed[0].number_total
    h = serving.RequestHandle(req, precision=precision)
    h.step(req.steps)
    np.testing.assert_array_equal(h.f, f_mixed[0])
    h.close()


def test_serving_workload_real_shapes():
    """The serving workload's own request shapes (p in {16, 32, 64}, u in {4, 8}, x in
    {1, 32, 128}, 1-2 species, 200-2000 steps), 24 requests at once in FP32 with the
    correction: all complete with their full step count; the largest request and one of each
    cell count equal their own handles stepped alone."""
    specs = serving.request_mix(count=24, rate_per_s=1e9)
    assert {s.p_num for s in specs} == {16, 32, 64} and {s.x_num for s in specs} == {1, 32, 128}
    res, wall, states = serving.run(specs, precision="fp32", realtime=False, fetch=True)
    for s, r in zip(specs, res):
        assert r.status == 2 and r.steps_done == s.steps
    picks = {max(range(len(specs)), key=lambda i: specs[i].n)}
    for x in (1, 32, 128):
        picks.add(next(i for i, s in enumerate(specs) if s.x_num == x))
    for i in sorted(picks):
        h = serving.RequestHandle(specs[i], precision="fp32")
        h.step(specs[i].steps)
        np.testing.assert_array_equal(h.f, states[i])
        h.close()
    print(f"24 real-shape requests in {wall:.2f} s")
