"""The benchmark configurations at their real sizes on the GPU.

* homogeneous_box (2 x 256 states, one cell, monolithic 512^2 x 512 M_Bin): the full 1000-step
  Solve in FP64 against the oracle at 1e-12, every save and every step's dt / save / Cr.
* transport_1d (2 x 256 states x 128 cells): the full 2000-step FP64 Solve on the GPU; the
  oracle checks the first 30 steps from f0 and the last 30 steps from the GPU's own state at
  step 1970 (same dt sequence), both at 1e-12; total number per species is conserved.  With
  dense per-cell emission blocks as well: 10 steps against the oracle at 1e-12; and 10
  ForwardSymplecticEuler steps with their Verbose-3 diagnostics.
* precision_box (3 x 1024 states) and transport_2d (2 x 512 states x 32 x 32 cells): FP32 and
  split-BF16 against the FP64 GPU path over the full 1000 / 500 steps, at the accuracy bars of
  the reduced-precision modes (< 1e-6 relative in f, < 1e-9 number and energy drift); FP64
  conserves number per species.

The oracle is too slow for the precision_box and transport_2d operators (8 and 0.5 G entries);
their FP64 path is the one pinned to the oracle on the smaller configurations above.
"""
import numpy as np
import pytest

import oracle as O
from dtx import workloads
from dtx.api import Solve

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
# Cr = -min(df / f) is a difference of gain and loss rates, each of order dt x (loss rate) = 0.1
# by the workloads' choice of dt; near equilibrium it is far smaller than its terms, so it is
# compared at 1e-12 of that scale (the state itself at 1e-12 of its own norm).
CR_SCALE = 0.1
F_BAR = 1e-6       # relative error in f of the reduced-precision modes
DRIFT_BAR = 1e-9   # relative number / energy drift of the reduced-precision modes


def rel_linf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rel_elementwise(a, b, floor=1e-8):
    """Largest relative error over the states above floor * max|b| (the error bar of the
    reduced-precision modes: a relative bar on states that carry the solution)."""
    a, b = np.asarray(a), np.asarray(b)
    mask = np.abs(b) > floor * np.max(np.abs(b))
    return float(np.max(np.abs(a[mask] - b[mask]) / np.abs(b[mask])))


def cr_close(a, b):
    return abs(a - b) <= TOL64 * max(abs(b), CR_SCALE)


def t_save_marks(w, marks):
    """Save times after the given step counts (exact multiples of the workload's dt)."""
    out, t, k = [0.0], 0.0, 0
    for m in marks[1:]:
        while k < m:
            t += w.dt
            k += 1
        out.append(t)
    return np.array(out)


def species_totals(ps, f, weights=None):
    """Number (or weighted sum) per species, summed over cells."""
    N, C = ps.n_momentum, ps.n_space
    fc = np.asarray(f).reshape(C, N)
    if weights is not None:
        fc = fc * weights[None, :]
    return np.array([fc[:, o:o + s.size].sum() for o, s in zip(ps.species_offset, ps.species)])


def test_homogeneous_box_full_solve_vs_oracle():
    w = workloads.homogeneous_box()
    assert (w.ps.n_momentum, w.ps.n_space, w.steps) == (512, 1, 1000)
    ts = w.t_save(5)
    m = workloads.make_struct(w, precision="fp64")  # the monolithic M_Bin through the C ABI
    out = Solve(m, w.dt, ts, Verbose=1)
    gstats = m.recent_stats(w.steps)
    fo, to, ostats, nsteps = O.problem_from_workload(w).solve(w.f0, w.dt, ts, verbose=1)
    assert nsteps == w.steps == len(gstats)
    np.testing.assert_array_equal(out.t, to)
    for i in range(len(ts)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, i
    for k, (g, o) in enumerate(zip(gstats, ostats)):
        assert g.dt == o.dt and g.save == o.save and g.finite == o.finite == 1, k
        assert cr_close(g.cr, o.cr), (k, g.cr, o.cr)
    num, en = m.moments()
    for s, (off, sp) in enumerate(zip(w.ps.species_offset, w.ps.species)):
        n0, e0 = O.moments(fo[-1][off:off + sp.size], sp)
        assert abs(num[0, s] - n0) <= TOL64 * abs(n0)
        assert abs(en[0, s] - e0) <= TOL64 * abs(e0)
    # collisions a + b -> a + b conserve the number of every species
    t0, t1 = species_totals(w.ps, w.f0), species_totals(w.ps, out.f[-1])
    assert np.all(np.abs(t1 - t0) <= 1e-12 * t0), (t0, t1)
    m.close()


def test_transport_1d_full_solve_windows_vs_oracle():
    w = workloads.transport_1d()
    assert (w.ps.n_momentum, w.ps.n_space, w.steps) == (512, 128, 2000)
    marks = [0, 30, 1000, 1970, 2000]
    ts = t_save_marks(w, marks)
    m = workloads.make_struct(w, precision="fp64", use_monolithic=False)
    out = Solve(m, w.dt, ts, Verbose=1)
    gstats = m.recent_stats(w.steps)
    assert len(gstats) == w.steps
    prob = O.problem_from_workload(w, use_monolithic=False, batched=True)
    # window 1: the oracle from f0
    fo, _, ostats, n = prob.solve(w.f0, w.dt, ts[:2], verbose=1)
    assert n == 30
    assert rel_linf(out.f[1], fo[1]) < TOL64
    for g, o in zip(gstats[:30], ostats):
        assert g.dt == o.dt and g.finite == o.finite == 1
        assert cr_close(g.cr, o.cr), (g.cr, o.cr)
    # window 2: the oracle from the GPU's state after 1970 steps, with the GPU's dt sequence
    st = O.Stepper(prob, out.f[3])
    for k in range(1970, 2000):
        g = gstats[k]
        dt, save, so = st(g.t_start, ts[4], g.dt, 1)
        assert dt == g.dt and save == bool(g.save) and so.finite == g.finite == 1, k
        assert cr_close(g.cr, so.cr), (k, g.cr, so.cr)
    assert rel_linf(out.f[4], st.f) < TOL64
    # periodic advection, closed-boundary cooling and collisions conserve each species' number
    t0 = species_totals(w.ps, w.f0)
    for f in out.f[1:]:
        t1 = species_totals(w.ps, f)
        assert np.all(np.abs(t1 - t0) <= 1e-12 * t0), (t0, t1)
    m.close()


def test_transport_1d_emission_vs_oracle():
    """transport_1d at full size with dense per-cell emission blocks (128 x 512^2, 269 MB): the
    emission GEMV runs with one column range per row tile; 10 FP64 steps against the oracle
    (block-CSC emission) at 1e-12, with every step's Cr."""
    w = workloads.transport_1d(emission=True, steps=10)
    ts = t_save_marks(w, [0, w.steps])
    m = workloads.make_struct(w, precision="fp64", use_monolithic=False)
    out = Solve(m, w.dt, ts, Verbose=1)
    gstats = m.recent_stats(w.steps)
    m.close()
    prob = O.problem_from_workload(w, use_monolithic=False, batched=True, sparse_emission=True)
    fo, _, ostats, n = prob.solve(w.f0, w.dt, ts, verbose=1)
    assert n == w.steps
    assert np.all(np.isfinite(out.f[1]))
    assert rel_linf(out.f[1], fo[1]) < TOL64
    for g, o in zip(gstats, ostats):
        assert g.dt == o.dt and g.finite == o.finite == 1
        assert cr_close(g.cr, o.cr), (g.cr, o.cr)


def test_transport_1d_symplectic_vs_oracle():
    """ForwardSymplecticEuler at transport_1d's full size (P flux with the momentum half, X flux
    with the space half): 10 FP64 steps at Verbose 3 against the oracle, the state at 1e-12 and
    every step's Cr, Cr_Bin, Cr_Momentum and Cr_Space."""
    w = workloads.transport_1d(steps=10)
    ts = t_save_marks(w, [0, w.steps])
    m = workloads.make_struct(w, precision="fp64", method="symplectic", use_monolithic=False)
    out = Solve(m, w.dt, ts, Verbose=3)
    gstats = m.recent_stats(w.steps)
    m.close()
    prob = O.problem_from_workload(w, use_monolithic=False, batched=True)
    fo, _, ostats, n = prob.solve(w.f0, w.dt, ts, verbose=3, method=1)
    assert n == w.steps == len(gstats)
    assert rel_linf(out.f[1], fo[1]) < TOL64
    for g, o in zip(gstats, ostats):
        assert g.dt == o.dt and g.finite == o.finite
        for name in ("cr", "cr_bin", "cr_momentum", "cr_space"):
            assert cr_close(getattr(g, name), getattr(o, name)), (name, getattr(g, name),
                                                                 getattr(o, name))


def _run_modes(w, steps):
    """Each precision over `steps` steps from f0 (dtx_solve, one save at the end)."""
    ts = t_save_marks(w, [0, steps])
    weights = w.ps.energy_weights()
    res = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        m = workloads.make_struct(w, precision=prec)
        out, n = m.solve_native(w.dt, ts, Verbose=0)
        assert n == steps
        f = out.f[-1]
        res[prec] = dict(f=f, number=species_totals(w.ps, f),
                         energy=species_totals(w.ps, f, weights))
        m.close()
    return res


def _check_modes(w, res):
    base = res["fp64"]
    t0 = species_totals(w.ps, w.f0)
    assert np.all(np.abs(base["number"] - t0) <= 1e-12 * t0), (base["number"], t0)
    for prec in ("fp32", "bf16x3"):
        r = res[prec]
        err = rel_elementwise(r["f"], base["f"])
        dn = np.max(np.abs(r["number"] - base["number"]) / base["number"])
        de = np.max(np.abs(r["energy"] - base["energy"]) / base["energy"])
        assert err < F_BAR, (prec, err)
        assert dn < DRIFT_BAR, (prec, dn)
        assert de < DRIFT_BAR, (prec, de)
        print(f"{w.name} {prec}: rel f {err:.2e}, number drift {dn:.2e}, energy drift {de:.2e}")


def test_precision_box_full_run():
    w = workloads.precision_box()
    assert (w.ps.n_momentum, w.ps.n_space, w.steps) == (3072, 1, 1000)
    _check_modes(w, _run_modes(w, w.steps))


def test_transport_2d_full_run():
    w = workloads.transport_2d()
    assert (w.ps.n_momentum, w.ps.n_space, w.steps) == (1024, 1024, 500)
    _check_modes(w, _run_modes(w, w.steps))
