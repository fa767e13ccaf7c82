"""The reference's elementwise edge cases through the HIP kernels (via the C ABI), compared bit
for bit, the sign of zero included:

* the hand-computed cases of tests/hand_cases.py (both methods, Verbose 1 and 3): the state after
  the step and every diagnostic the functor reports (finite flag, Cr, Cr_Bin, Cr_Emi, Cr_Flux /
  Cr_PFlux, Cr_Momentum, Cr_Space);
* NaN, +-Inf, negative values, -0.0 and values around Float64(1f-10) / Float64(1f-20) planted in
  a multi-cell workload with flux, collisions, emission and injection, against the oracle.

The one allowance is the sign of a NaN that an operation creates (0 * Inf, Inf - Inf): IEEE 754
leaves it to the hardware, so a NaN clamped by `f * (f >= thr)` may become +0.0 on one side and
-0.0 on the other (documented deviation).  Zeros computed in a cell that held a NaN or an
infinity, or in the cells next to it, may therefore differ in sign; everything else is exact.
"""
import math

import numpy as np
import pytest

import hand_cases as H
import oracle as O
from dtx import workloads
from dtx.api import (BinaryMatricesStruct, EmissionMatricesStruct, FluxMatricesStruct,
                     ForwardEulerStruct, ForwardSymplecticEulerStruct)

pytestmark = pytest.mark.gpu


def _hand_struct(case):
    ps = case.phase_space()
    n = case.n
    BinM = BinaryMatricesStruct(M_Bin=case.M_Bin, Binary_list=["hand"] if case.M_Bin is not None
                                else [])
    FluxM = FluxMatricesStruct(np.ones(n), case.F if case.method == 0 else case.X,
                               None if case.method == 0 else case.P, np.ones(1))
    cls = ForwardEulerStruct if case.method == 0 else ForwardSymplecticEulerStruct
    return cls(ps, np.asarray(case.f, dtype=np.float64), case.injection, BinM,
               EmissionMatricesStruct(), FluxM)


@pytest.mark.parametrize("verbose", [1, 3])
@pytest.mark.parametrize("case", H.CASES, ids=lambda c: c.name)
def test_hand_cases_on_gpu(case, verbose):
    m = _hand_struct(case)
    dt, save, st = m(0.0, 10.0, 1.0, verbose, stats=True)
    assert dt == 1.0 and not save
    H.check_state(case, m.f)
    H.check_stats(case, st, verbose)
    m.close()


# The console lines of one step (t_start = 0, t_stop = 10, dt = 1), written out from the hand
# cases' diagnostics and the reference's println text (src/SteppingMethods.jl:49,87-93 and
# :222-228; Float64 values as Julia prints them, Cr of the symplectic lines to 3 digits).
NONFINITE_LINE = "non-finite value in df calculation\n"
PRINTED = {
    ("euler_binary_flux", 0): "",
    ("euler_binary_flux", 1): "Cr = 2.0, t=0.0, dt=1.0, system may be unstable\n",
    ("euler_binary_flux", 2): "\rCr = 2.0, t=0.0, dt=1.0\n",
    ("euler_binary_flux", 3): "Cr = 2.0,Cr_Bin = 2.0, Cr_Emi = 0.0, Cr_Flux = 0.5, t=0.0, "
                              "t_save =10.0, dt=1.0\n",
    ("euler_nonfinite", 0): NONFINITE_LINE,
    ("euler_nonfinite", 1): NONFINITE_LINE,
    ("euler_nonfinite", 2): NONFINITE_LINE + "\rCr = 0.5, t=0.0, dt=1.0\n",
    ("euler_nonfinite", 3): NONFINITE_LINE + "Cr = 0.5,Cr_Bin = 0.0, Cr_Emi = 0.0, "
                            "Cr_Flux = 0.5, t=0.0, t_save =10.0, dt=1.0\n",
    ("symplectic_pflux_nan", 0): "",
    ("symplectic_pflux_nan", 1): "",
    ("symplectic_pflux_nan", 2): "Cr = 0.25, t=0.0, dt=1.0\n",
    ("symplectic_pflux_nan", 3): "Cr = 0.25, Cr_Bin = 0.0, Cr_Emi = 0.0, Cr_PFlux = -0.0, "
                                 "Cr_Momentum = -0.5, Cr_Space = 0.5,  t=0.0, t_save=10.0 , "
                                 "dt=1.0\n",
}


@pytest.mark.parametrize("verbose", [0, 1, 2, 3])
@pytest.mark.parametrize("name", ["euler_binary_flux", "euler_nonfinite", "symplectic_pflux_nan"])
def test_printed_lines(name, verbose, capfd):
    case = next(c for c in H.CASES if c.name == name)
    m = _hand_struct(case)
    capfd.readouterr()
    m(0.0, 10.0, 1.0, verbose, stats=True)  # the stats call reads and prints the step's record
    out = capfd.readouterr().out
    m.close()
    assert out == PRINTED[(name, verbose)]


FINITE = [-1.0, -1e-12, -0.0, 0.0, H.THR10, float(np.nextafter(H.THR10, 0.0)),
          float(np.nextafter(H.THR10, 1.0)), H.THR20, float(np.nextafter(H.THR20, 0.0)), 1e-300,
          5e-324, 3.0]
NONFINITE = [math.nan, math.inf, -math.inf, -1.0]


def _plant(w, cells):
    """f0 with the given values planted in the given cells (distinct states of each cell)."""
    f = np.array(w.f0)
    N = w.ps.n_momentum
    for c, vals in cells.items():
        for q, v in enumerate(vals):
            f[c * N + (7 * q + 3) % N] = v
    return f


def _compare(w, f_gpu, f_ora, pre, exact_cells, tainted_cells):
    """exact_cells: flux and injection only (no collisions, no emission), bit for bit.
    tainted_cells: hold or border a NaN / Inf, so their zeros may carry either sign and a zero
    clamped from a NaN is a zero of either sign.  Elsewhere zeros keep their sign and non-zero
    values agree to 1e-12 relative to the state's largest value."""
    N = w.ps.n_momentum
    scale = np.max(np.abs(f_ora[np.isfinite(f_ora)]))
    bad = []
    for i, (g, o) in enumerate(zip(f_gpu, f_ora)):
        cell = i // N
        if cell in exact_cells:
            ok = H.same_bits(float(g), float(o))
        elif math.isnan(pre[i]) and not math.isnan(o):
            ok = g == 0.0
        elif o == 0.0:
            ok = g == 0.0 and (cell in tainted_cells or H.same_bits(float(g), float(o)))
        elif math.isnan(o) or math.isinf(o):
            ok = H.same_bits(float(g), float(o))
        else:
            ok = abs(g - o) <= 1e-12 * scale
        if not ok:
            bad.append((i, g, o, pre[i]))
    assert not bad, bad[:10]


@pytest.mark.parametrize("emission", [False, True], ids=["no_emission", "emission"])
@pytest.mark.parametrize("method", ["euler", "symplectic"])
def test_special_values_against_oracle(method, emission):
    """Five periodic cells with x advection and synchrotron cooling; collisions only in cells 0
    and 4 (Domain).  Cell 1 holds the finite specials (negatives, -0.0, values one ulp around the
    clamp thresholds, subnormals): without emission its update is flux and injection
    only, so it must match the oracle bit for bit.  Cell 3 holds NaN, +-Inf and a negative value;
    its neighbours 2 and 4 receive them through the flux.  One step at Verbose 3."""
    w = workloads.tiny(x_num=5, steps=3, emission=emission, injection=True, domain=[0, 4])
    f0 = _plant(w, {1: FINITE, 3: NONFINITE})
    m = workloads.make_struct(w, precision="fp64", method=method)
    m.f = f0
    # the library keeps M_Emi as per-cell blocks, which behaves as the reference's sparse M_Emi:
    # a NaN or Inf reaches only its own cell (the dense n x n product would carry it into every
    # row through the stored zeros of the other blocks)
    st = O.Stepper(O.problem_from_workload(w, sparse_emission=True), f0,
                   method=0 if method == "euler" else 1)
    dt_g, save_g, sg = m(0.0, 1.0, w.dt, 3, stats=True)
    dt_o, save_o, so = st(0.0, 1.0, w.dt, 3)
    assert dt_g == dt_o and save_g == save_o
    _compare(w, m.f, st.f, st.pre_clamp, exact_cells=set() if emission else {1},
             tainted_cells={2, 3, 4})
    assert sg.finite == so.finite
    if method == "euler":
        assert sg.finite == 0  # NaN and Inf reach df
    keys = ["cr", "cr_bin", "cr_emi", "cr_flux"]
    if method == "symplectic":
        keys += ["cr_momentum", "cr_space"]
    for k in keys:
        a, b = getattr(sg, k), getattr(so, k)
        if math.isfinite(b) and b != 0.0:
            assert abs(a - b) <= 1e-12 * abs(b), (k, a, b)
        else:
            assert H.same_bits(a, b) or (a == 0.0 and b == 0.0), (k, a, b)
    if method == "euler":
        # the clamps removed every non-finite value: the next step is finite again and within
        # 1e-12 of the oracle everywhere
        m(dt_g, 1.0, w.dt, 1)
        _, _, so = st(dt_o, 1.0, w.dt, 1)
        _, _, sg = m(2 * dt_g, 1.0, w.dt, 1, stats=True)
        _, _, so = st(2 * dt_o, 1.0, w.dt, 1)
        assert sg.finite == so.finite == 1
        f_g, f_o = m.f, st.f
        assert np.all(np.isfinite(f_g))
        assert np.max(np.abs(f_g - f_o)) <= 1e-12 * np.max(np.abs(f_o))
    m.close()


def test_threshold_neighbourhood_values_bitwise():
    """Values a few ulps around the clamp thresholds after a real update: a one-cell problem
    with flux-free collisions where f0 is scaled so that many states land near 1f-10 after the
    step; the clamped states and their signs match the oracle exactly."""
    w = workloads.tiny(x_num=1, steps=3, with_flux=False)
    f0 = np.array(w.f0)
    f0[::3] = H.THR10 * (1.0 + np.linspace(-1e-15, 1e-15, f0[::3].size))
    f0[1::5] = -f0[1::5]
    m = workloads.make_struct(w, precision="fp64")
    m.f = f0
    st = O.Stepper(O.problem_from_workload(w), f0)
    m(0.0, 1.0, w.dt, 3)
    st(0.0, 1.0, w.dt, 3)
    for g, o in zip(m.f, st.f):
        if o == 0.0:
            assert H.same_bits(g, o), (g, o)
    assert np.max(np.abs(m.f - st.f)) <= 1e-12 * np.max(np.abs(st.f))
    m.close()
