"""GPU: reduced-precision contraction (FP32 MFMA, split-BF16 x3) against the FP64 GPU path,
which the parity tests pin to the FP64 oracle.

Bars (over 60 steps of small cases):
* relative L-inf error of f against FP64 (over entries above 1e-8 of the maximum):
  FP32 <= 1e-6, BF16x3 <= 1e-4;
* with the FP64 moment-rate correction, the number of every (cell, species) follows the FP64
  run to <= 1e-12 relative (the synthetic collisions conserve number exactly), and the energy
  error stays within twice the uncorrected one (the energy rate is exact for the current f, so
  what remains comes from the f error).
"""
import numpy as np
import pytest

from dtx import workloads

pytestmark = pytest.mark.gpu
STEPS = 60
F_TOL = {"fp32": 1e-6, "bf16x3": 1e-4}


def run(w, precision, correction=True, steps=STEPS, method="euler"):
    m = workloads.make_struct(w, precision=precision, use_monolithic=False,
                              moment_correction=correction, method=method)
    t = 0.0
    for _ in range(steps):
        dt, _ = m(t, 1e300, w.dt, 0)
        t += dt
    m.synchronize()
    f = m.f
    num, en = m.moments()
    m.close()
    return f, num, en


def rel_err_f(f, ref):
    mask = np.abs(ref) > 1e-8 * np.max(np.abs(ref))
    return float(np.max(np.abs(f[mask] - ref[mask]) / np.abs(ref[mask])))


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def homogeneous_channels():
    """Channel-form tiny case without flux: collisions alone, number conserved per cell."""
    return workloads.tiny(p_num=8, u_num=4, x_num=3, with_flux=False, monolithic=False)


def synthetic_small():
    """Device-generated channels (3 species), the precision configuration at small size."""
    return workloads.precision_box(p_num=8, u_num=4, h_num=2, num_species=3, steps=STEPS)


def many_cells():
    """80 cells (the split-BF16 path switches to its 128-cell tiles above 64 cells: one partial
    cell tile) and 160 states per species (three 64-row k tiles: one full pair, one single)."""
    return workloads.tiny(p_num=20, u_num=8, x_num=80, with_flux=False, monolithic=False)


def wide_slices():
    """520 cells: the split-BF16 path takes four k tiles per item from 512 cells on (five cell
    tiles, the last with 8 cells; three k tiles per species, so one slice of each item is
    absent)."""
    return workloads.tiny(p_num=20, u_num=8, x_num=520, with_flux=False, monolithic=False)


def cells_32():
    """24 cells: 32-cell tiles (one partial tile) and 160 states per species."""
    return workloads.tiny(p_num=20, u_num=8, x_num=24, with_flux=False, monolithic=False)


def single_cell():
    """One cell, 160 states per species: GEMV items over three k tiles."""
    return workloads.tiny(p_num=20, u_num=8, x_num=1, with_flux=False, monolithic=False)


CASES = {"host_channels": homogeneous_channels, "device_channels": synthetic_small,
         "many_cells": many_cells, "wide_slices": wide_slices, "single_cell": single_cell,
         "cells_32": cells_32}


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("case", list(CASES))
def test_moment_rate_correction(precision, case):
    w = CASES[case]()
    print(f"[case] {case} {precision} N={w.ps.n_momentum} C={w.ps.n_space}", flush=True)
    f64, n64, e64 = run(w, "fp64")
    fc, nc, ec = run(w, precision, correction=True)
    fu, nu, eu = run(w, precision, correction=False)
    assert rel_err_f(fc, f64) <= F_TOL[precision]
    assert rel_err_f(fu, f64) <= F_TOL[precision]
    assert rel_err(nc, n64) <= 1e-12, (rel_err(nc, n64), rel_err(nu, n64))
    assert rel_err(nu, n64) > rel_err(nc, n64)
    assert rel_err(ec, e64) <= 2.0 * rel_err(eu, e64) + 1e-12


def test_correction_off_in_fp64_is_a_no_op():
    w = homogeneous_channels()
    a = run(w, "fp64", correction=True, steps=10)
    b = run(w, "fp64", correction=False, steps=10)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_correction_must_precede_first_step():
    w = homogeneous_channels()
    m = workloads.make_struct(w, precision="fp32", use_monolithic=False)
    m(0.0, 1e300, w.dt, 0)
    rc = m.lib.dtx_set_moment_correction(m._h, 0)
    assert rc == -4  # DTX_ERR_STATE
    m.close()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_symplectic_reduced_precision(precision):
    """ForwardSymplecticEuler with the reduced-precision contraction follows its FP64 run to the
    same bars (the momentum and space kernels take the same corrected binary sums)."""
    w = workloads.tiny(p_num=8, u_num=4, x_num=3, monolithic=False, injection=True)
    f64, n64, e64 = run(w, "fp64", method="symplectic", steps=30)
    fc, nc, ec = run(w, precision, method="symplectic", steps=30)
    assert rel_err_f(fc, f64) <= F_TOL[precision]
    assert np.all(np.isfinite(nc)) and np.all(np.isfinite(ec))


def _steps_with_stats(w, precision, correction, steps, verbose=3):
    m = workloads.make_struct(w, precision=precision, use_monolithic=False,
                              moment_correction=correction)
    assert m.work().fused_correction == (1 if correction is True else 0)
    t, stats = 0.0, []
    for _ in range(steps):
        dt, _, st = m(t, 1e300, w.dt, verbose, stats=True)
        stats.append((st.finite, st.cr, st.cr_bin, st.cr_emi, st.cr_flux, st.sum_f))
        t += dt
    m.synchronize()
    f = m.f
    num, en = m.moments()
    m.close()
    return f, num, en, stats


FUSED_CASES = {
    # 2 species x 256 states, 128 cells: one pass of four segments per species
    "transport_1d": lambda: workloads.transport_1d(steps=6),
    # 320 states per species: a second pass with one live segment group of four
    "two_passes": lambda: workloads.transport_1d(p_num=40, x_num=64, steps=6),
    # 128 cells with emission blocks (the GEMV's df_Emi read by the update), a binary Domain
    # (the workgroup's all-zero test of the cell) and injection
    "emission_domain": lambda: workloads.tiny(p_num=16, u_num=8, x_num=128, monolithic=False,
                                              emission=True, injection=True,
                                              domain=list(range(1, 129, 3)), steps=6),
}


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("case", list(FUSED_CASES))
def test_fused_correction_equals_three_launches(precision, case):
    """Corrected Forward Euler steps by the one-launch form (slot sums, coefficients and update
    fused per (cell, species)) against the three-launch form: f, moments and the Verbose-3
    diagnostics bit for bit."""
    w = FUSED_CASES[case]()
    print(f"[case] {case} {precision} N={w.ps.n_momentum} C={w.ps.n_space}", flush=True)
    fa, na, ea, sa = _steps_with_stats(w, precision, True, w.steps)
    fb, nb, eb, sb = _steps_with_stats(w, precision, 2, w.steps)
    np.testing.assert_array_equal(fa.view(np.uint64), fb.view(np.uint64))
    np.testing.assert_array_equal(na.view(np.uint64), nb.view(np.uint64))
    np.testing.assert_array_equal(ea.view(np.uint64), eb.view(np.uint64))
    assert sa == sb
