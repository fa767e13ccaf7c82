"""Pins the FP64 CPU oracle: against the pure-Python brute force, closed forms, invariants and
hand-computed sequences.  CPU only (no GPU needed)."""
import math

import numpy as np
import pytest

import brute
import oracle as O
from dtx import grids, synth, workloads


@pytest.fixture(scope="module")
def tiny():
    return workloads.tiny(steps=12, injection=True)


def test_binary_two_gemv_bitwise_and_closed_form(tiny):
    """Oracle update_Big_Bin! == brute-force two-GEMV loop bit for bit, and == the closed form
    Vol * sum_{i,j} M[k,i,j] f_i f_j within rounding."""
    w = tiny
    N, C = w.ps.n_momentum, w.ps.n_space
    prob = O.problem_from_workload(w)
    dB = prob.binary_term(w.f0)
    st = brute.make_state(w)
    dB2 = brute.binary_term_two_gemv(st["M_Bin"], list(w.f0), N, C, st["Vol"])
    np.testing.assert_array_equal(dB, dB2)
    M3 = w.M_Bin.tolist()
    dB3 = brute.binary_term_triple(M3, list(w.f0), N, C, st["Vol"])
    scale = np.max(np.abs(dB3))
    assert np.max(np.abs(dB - np.array(dB3))) <= 1e-13 * scale


def test_monolithic_equals_channels(tiny):
    """The channel form (gain blocks + loss blocks) contracts to the same dB as the monolithic
    tensor assembled by the reference's writers."""
    w = tiny
    a = O.problem_from_workload(w, use_monolithic=True).binary_term(w.f0)
    b = O.problem_from_workload(w, use_monolithic=False).binary_term(w.f0)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))


def test_binary_invariants(tiny):
    w = tiny
    N, C = w.ps.n_momentum, w.ps.n_space
    M = w.M_Bin
    # symmetric in the two input indices (gains split val/2 into (k,i,j) and (k,j,i))
    np.testing.assert_allclose(M, M.transpose(0, 2, 1), rtol=0, atol=1e-15)
    prob = O.problem_from_workload(w)
    f = w.f0
    dB = prob.binary_term(f)
    # number conservation per cell: losses are the column sums of the gains
    for c in range(C):
        blk = dB[c * N:(c + 1) * N]
        terms = np.abs(np.einsum("kij,i,j->k", np.abs(M), f[c*N:(c+1)*N], f[c*N:(c+1)*N]))
        assert abs(blk.sum()) <= 1e-13 * terms.sum()
    # bilinearity: f -> 3 f gives dB -> 9 dB
    np.testing.assert_allclose(prob.binary_term(3.0 * f), 9.0 * dB, rtol=1e-13,
                               atol=1e-13 * np.abs(dB).max())
    # Vol scaling: dB is proportional to Vol of the cell
    w2 = workloads.tiny(steps=12, injection=True)
    w2.Vol = w2.Vol * 2.5
    dB2 = O.problem_from_workload(w2).binary_term(f)
    np.testing.assert_allclose(dB2, 2.5 * dB, rtol=1e-14, atol=1e-14 * np.abs(dB).max())


def test_loss_is_column_sum():
    ps = grids.PhaseSpace(workloads.species_pair(3, 2, 1))
    gains, losses = synth.binary_channels(ps, seed=11)
    for g in gains:
        Ls = [l for l in losses if (l.s1, l.s2) in ((g.s1, g.s2), (g.s2, g.s1))]
        col = g.G.sum(axis=0)
        assert any(np.allclose(l.L, col) or np.allclose(l.L, col.T) for l in Ls)


def test_domain_cells_without_binary():
    """Cells outside Domain (with non-zero f) keep df_Bin = 0 (src/SteppingMethods.jl:369)."""
    w = workloads.tiny(steps=3, domain=[1])
    dB = O.problem_from_workload(w).binary_term(w.f0)
    N = w.ps.n_momentum
    assert np.all(dB[0:N] == 0.0) and np.all(dB[2 * N:3 * N] == 0.0)
    assert np.any(dB[N:2 * N] != 0.0)


def test_flux_csc_matches_dense_and_conserves():
    w = workloads.transport_1d(p_num=4, u_num=4, x_num=6, steps=1)
    n = w.ps.n
    F = w.F
    D = synth.csc_to_dense(F, n)
    x = synth.normal(np.arange(n, dtype=np.uint64), 123)
    y = brute.csc_matvec(F, list(x))
    np.testing.assert_allclose(y, D @ x, rtol=1e-13, atol=1e-16)
    # column sums vanish: periodic advection and closed-boundary cooling move particles only
    assert np.max(np.abs(D.sum(axis=0))) <= 1e-12 * np.max(np.abs(D))
    # CSC is Julia-shaped: 1-based, rows ascending inside each column
    colptr, rowval, _ = F
    assert colptr[0] == 1 and rowval.min() >= 1 and rowval.max() <= n
    for j in range(n):
        r = rowval[colptr[j] - 1:colptr[j + 1] - 1]
        assert np.all(np.diff(r) > 0)
    # stored entries per row: self, x +-1, p +-1 and u +-1 except at the closed p and u
    # boundaries, explicit zeros kept
    p, u = 4, 4
    assert (colptr[-1] - 1) / n == 1 + 2 + 2 * (p - 1) / p + 2 * (u - 1) / u


def test_courant_one_pulse_advects_one_cell():
    """Upwind advection at Courant number 1 shifts a one-bin pulse by exactly one cell."""
    sp = [grids.Species("Pho", 0.0, 0.0, 1, 1, 1, p_low=0.0, p_up=1.0, u_low=0.5, u_up=1.0)]
    ps = grids.PhaseSpace(sp, x_num=8)
    X, P = synth.flux_matrices(ps, axes="x")
    Ap, Vol = synth.ap_vol(ps)
    v = 1.0 * (1.0 - 0.25) / (2 * 0.5)  # massless: <v> = 1; <u> over [0.5, 1] = 0.75
    dt = (1.0 / 8) / v  # v dt / dx = 1
    f0 = np.zeros(ps.n)
    f0[3] = 1.0
    prob = O.OracleProblem(ps, Ap, Vol, np.zeros(ps.n), F=X)
    st = O.Stepper(prob, f0)
    st(0.0, 10.0, dt, 0)
    expect = np.zeros(ps.n)
    expect[4] = 1.0
    np.testing.assert_allclose(st.f, expect, atol=1e-15)


def test_clamp_semantics():
    """f*(f >= 1f-10) then f*(f != Inf) with Julia's strong-zero Bool product
    (src/SteppingMethods.jl:99-101): negatives -> -0.0, small positives -> +0.0,
    NaN -> signed zero, +Inf -> +0.0, Float64(1f-10) itself survives."""
    ps = grids.PhaseSpace([grids.Species("Ele", 1.0, -1.0, 8, 1, 1)])
    thr = float(np.float32(1e-10))
    f0 = np.array([-1.0, 5e-11, thr, np.nextafter(thr, 0), np.inf, 2.0, np.nan, -np.inf])
    Ap, Vol = synth.ap_vol(ps)
    prob = O.OracleProblem(ps, Ap, Vol, np.zeros(ps.n))
    st = O.Stepper(prob, f0)
    st(0.0, 1.0, 0.1, 0)
    f = st.f
    assert f[0] == 0.0 and math.copysign(1.0, f[0]) == -1.0
    assert f[1] == 0.0 and math.copysign(1.0, f[1]) == 1.0
    assert f[2] == thr
    assert f[3] == 0.0
    assert f[4] == 0.0 and math.copysign(1.0, f[4]) == 1.0
    assert f[5] == 2.0
    assert f[6] == 0.0
    assert f[7] == 0.0 and math.copysign(1.0, f[7]) == -1.0
    # the brute force agrees bit for bit
    bst = {"dt0": 1.0, "F": None, "invAp": [1.0] * 8, "df_inj": [0.0] * 8, "f": list(f0)}
    brute.forward_euler(bst, 0.0, 1.0, 0.1, 0)
    for a, b in zip(bst["f"], f):
        assert (a == b) and math.copysign(1.0, a) == math.copysign(1.0, b)


def test_dt_clip_and_carry_over():
    """Solve's dt sequence (src/Solver.jl:25-37, src/SteppingMethods.jl:18-23) by hand:
    t_save = [0, 0.25, 1.0], dt = 0.1 -> 0.1, 0.1, 0.05 (clipped, save) and then 0.05 carried
    into the second interval for 15 steps."""
    ps = grids.PhaseSpace([grids.Species("Ele", 1.0, -1.0, 2, 1, 1)])
    Ap, Vol = synth.ap_vol(ps)
    prob = O.OracleProblem(ps, Ap, Vol, np.zeros(ps.n))
    fo, to, stats, nsteps = prob.solve(np.ones(ps.n), 0.1, [0.0, 0.25, 1.0], verbose=0)
    dts = [s.dt for s in stats]
    assert nsteps == 3 + 15
    np.testing.assert_allclose(dts[:3], [0.1, 0.1, 0.25 - 0.2], rtol=1e-14)
    assert all(abs(d - dts[2]) <= 1e-15 for d in dts[3:-1])
    assert [s.save for s in stats].count(1) == 2
    np.testing.assert_array_equal(to, [0.0, 0.25, 1.0])


def test_cr_definition():
    """Cr = -min(isnan(df/f) ? Inf : df/f) when sum(f) != 0 (src/SteppingMethods.jl:60-83)."""
    r = brute.neg_min_ratio([1.0, -2.0, 0.0], [1.0, 4.0, 0.0])
    assert r == 0.5
    assert brute.neg_min_ratio([1.0], [-0.0]) == math.inf


@pytest.mark.parametrize("method", [0, 1])
def test_oracle_solve_equals_brute_force(method):
    """Whole Solve runs: the C oracle and the pure-Python brute force agree bit for bit on the
    tiny workload with flux, injection and collisions (both methods)."""
    w = workloads.tiny(steps=8, injection=True)
    ts = w.t_save(3)
    fo, to, stats, nsteps = O.problem_from_workload(w).solve(w.f0, w.dt, ts, verbose=1,
                                                             method=method)
    st = brute.make_state(w)
    outs, dts = brute.solve(st, list(w.f0), w.dt, list(ts), verbose=1, method=method)
    assert len(dts) == nsteps
    for a, b in zip(outs, fo):
        np.testing.assert_array_equal(np.array(a), b)


def test_oracle_with_emission_equals_brute_force():
    w = workloads.tiny(steps=6, emission=True)
    ts = w.t_save(2)
    fo, *_ = O.problem_from_workload(w).solve(w.f0, w.dt, ts, verbose=1)
    st = brute.make_state(w)
    outs, _ = brute.solve(st, list(w.f0), w.dt, list(ts), verbose=1)
    np.testing.assert_array_equal(np.array(outs[-1]), fo[-1])


def test_moments_closed_form():
    s = grids.Species("Ele", 1.0, -1.0, 6, 3, 2)
    f = synth.uniform(np.arange(s.size, dtype=np.uint64), 5)
    num, en = O.moments(f, s)
    wgt = np.tile(grids.energy_weight_T00(s.pr, 1.0), s.u_num * s.h_num)
    assert abs(num - f.sum()) <= 1e-14 * f.sum()
    assert abs(en - f @ wgt) <= 1e-13 * abs(f @ wgt)
    # a single-bin state gives exactly that bin's weight
    g = np.zeros(s.size)
    g[4] = 1.0
    assert O.moments(g, s)[1] == grids.energy_weight_T00(s.pr, 1.0)[4]


def test_synchrotron_j_flux_aligns_pitch_angle():
    """The J (pitch-angle) part of the synchrotron radiation reaction alone (one p bin, so no I
    neighbours): conserves number, keeps f mirror-symmetric in u, moves particles towards
    |u| = 1 (alignment with B), and passes nothing through u = +-1."""
    sp = [grids.Species("Ele", 1.0, -1.0, 1, 8, 1, p_low=0.0, p_up=0.5)]
    ps = grids.PhaseSpace(sp, x_num=1)
    _, P = synth.flux_matrices(ps, sync_K=1.0, axes="")
    D = synth.csc_to_dense(P, ps.n)
    Ap, Vol = synth.ap_vol(ps)
    assert np.count_nonzero(D) > 0
    assert np.max(np.abs(D.sum(axis=0))) <= 1e-14 * np.max(np.abs(D))   # conservation
    f0 = np.ones(ps.n)
    prob = O.OracleProblem(ps, Ap, Vol, np.zeros(ps.n), F=P)
    st = O.Stepper(prob, f0)
    dt = 0.1 / np.max(np.abs(np.diag(D)) / Ap)
    st(0.0, 1.0, dt, 0)
    f1 = st.f
    assert abs(f1.sum() - f0.sum()) <= 1e-14 * f0.sum()
    np.testing.assert_allclose(f1, f1[::-1], rtol=1e-14)               # u -> -u symmetry
    umid = np.abs(0.5 * (sp[0].ur[1:] + sp[0].ur[:-1]))
    assert umid @ f1 > umid @ f0                                       # towards |u| = 1
    assert f1[0] > f0[0] and f1[-1] > f0[-1]                            # outer bins gain
    assert f1[3] < f0[3] and f1[4] < f0[4]                              # u ~ 0 bins lose
