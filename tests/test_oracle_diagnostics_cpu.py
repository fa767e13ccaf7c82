"""Pins the oracle's step diagnostics (finite flag, Cr, Cr_Bin, Cr_Emi, Cr_Flux / Cr_PFlux,
Cr_Momentum, Cr_Space) and the symplectic method's elementwise semantics: against hand-computed
cases (tests/hand_cases.py) and, step by step over whole solves, against the pure-Python brute
force.  Also pins the oracle's cell-batched channel evaluation to its cell-by-cell loop.
CPU only."""
import math

import numpy as np
import pytest

import brute
import hand_cases as H
import oracle as O
from dtx import workloads

STAT_KEYS = ("finite", "cr_evaluated", "cr", "cr_bin", "cr_emi", "cr_flux")
SYMPLECTIC_KEYS = STAT_KEYS + ("cr_momentum", "cr_space")


def _oracle_step(case, verbose):
    ps = case.phase_space()
    n = case.n
    prob = O.OracleProblem(ps, np.ones(n), np.ones(1), case.injection, F=case.F, X=case.X,
                           P=case.P, M_Bin=case.M_Bin)
    st = O.Stepper(prob, np.asarray(case.f, dtype=np.float64), method=case.method)
    dt, save, stats = st(0.0, 10.0, 1.0, verbose)
    return st.f, stats


def _brute_step(case, verbose):
    n = case.n
    state = {"dt0": 1.0, "N": n, "C": 1, "Vol": [1.0], "invAp": [1.0] * n,
             "df_inj": [float(v) for v in case.injection], "df_Bin": [0.0] * n,
             "F": case.F, "X": case.X, "P": case.P, "M_Emi": None,
             "M_Bin": ([float(v) for v in case.M_Bin.reshape(-1, order="F")]
                       if case.M_Bin is not None else None),
             "f": [float(v) for v in case.f]}
    step = brute.forward_euler if case.method == 0 else brute.forward_symplectic_euler
    _, _, stats = step(state, 0.0, 10.0, 1.0, verbose)
    return state["f"], stats


@pytest.mark.parametrize("verbose", [1, 3])
@pytest.mark.parametrize("case", H.CASES, ids=lambda c: c.name)
def test_oracle_hand_cases(case, verbose):
    f, st = _oracle_step(case, verbose)
    H.check_state(case, f)
    H.check_stats(case, st, verbose)


@pytest.mark.parametrize("verbose", [1, 3])
@pytest.mark.parametrize("case", H.CASES, ids=lambda c: c.name)
def test_brute_force_hand_cases(case, verbose):
    f, st = _brute_step(case, verbose)
    H.check_state(case, f)
    H.check_stats(case, {k: (int(v) if isinstance(v, bool) else v) for k, v in st.items()},
                  verbose)


def test_cr_pflux_replaces_nan_before_the_minimum():
    """Cr_PFlux's replace!(NaN => 0.0) (src/SteppingMethods.jl:211-213) changes the result
    against the isnan -> Inf map that every other Cr uses: -0.0 here, not -0.5."""
    _, st = _oracle_step(H.SYMPLECTIC, 3)
    assert st.cr_flux == 0.0 and math.copysign(1.0, st.cr_flux) == -1.0
    assert brute.neg_min_ratio([0.5, math.nan], [1.0, 1.0]) == -0.5


def _stats_equal(a, b, keys):
    for k in keys:
        x, y = getattr(a, k), b[k]
        if isinstance(y, bool):
            y = int(y)
        assert H.same_bits(float(x), float(y)), (k, x, y)


@pytest.mark.parametrize("verbose", [1, 3])
@pytest.mark.parametrize("method", [0, 1])
def test_oracle_stats_equal_brute_force_every_step(method, verbose):
    """Every field of every step's stats, oracle against brute force, bit for bit, on a tiny
    workload with flux, emission, binary collisions and injection."""
    w = workloads.tiny(steps=6, emission=True, injection=True)
    ts = w.t_save(3)
    prob = O.problem_from_workload(w)
    fo, to, stats, nsteps = prob.solve(w.f0, w.dt, ts, verbose=verbose, method=method)
    st = brute.make_state(w)
    step = brute.forward_euler if method == 0 else brute.forward_symplectic_euler
    st["f"] = list(w.f0)
    dt, k = w.dt, 0
    for i in range(1, len(ts)):
        t, save = ts[i - 1], False
        while not save:
            dt, save, bs = step(st, t, ts[i], dt, verbose)
            _stats_equal(stats[k], bs, STAT_KEYS if method == 0 else SYMPLECTIC_KEYS)
            assert stats[k].dt == dt and bool(stats[k].save) == save
            t += dt
            k += 1
    assert k == nsteps
    np.testing.assert_array_equal(np.array(st["f"]), fo[-1])


def test_non_finite_df_reported_over_a_solve():
    """A state holding +Inf reports finite = 0 at its first step, in the oracle and the brute
    force alike, and the state is cleaned by the clamps so later steps are finite again."""
    w = workloads.tiny(steps=4, with_flux=True)
    f0 = np.array(w.f0)
    f0[5] = math.inf
    prob = O.problem_from_workload(w)
    fo, _, stats, n = prob.solve(f0, w.dt, w.t_save(2), verbose=1)
    st = brute.make_state(w)
    outs, _ = brute.solve(st, list(f0), w.dt, list(w.t_save(2)), verbose=1)
    assert stats[0].finite == 0
    assert all(s.finite == 1 for s in stats[1:])
    np.testing.assert_array_equal(np.array(outs[-1]), fo[-1])


@pytest.mark.parametrize("kw", [dict(x_num=5, domain=[0, 2, 3]),
                                dict(num_species=3, x_num=4, emission=True)],
                         ids=["domain", "3species"])
def test_batched_channels_equal_cell_loop(kw):
    """The cell-batched channel evaluation gives the cell-by-cell result bit for bit."""
    w = workloads.tiny(steps=6, injection=True, monolithic=False, **kw)
    a = O.problem_from_workload(w, use_monolithic=False)
    b = O.problem_from_workload(w, use_monolithic=False, batched=True)
    np.testing.assert_array_equal(a.binary_term(w.f0), b.binary_term(w.f0))
    fa = a.solve(w.f0, w.dt, w.t_save(3))[0]
    fb = b.solve(w.f0, w.dt, w.t_save(3))[0]
    np.testing.assert_array_equal(fa, fb)


def test_sparse_and_dense_emission():
    """The oracle's sparse M_Emi (Emi_sparse, src/Collisions/BuildEmissionMatrices.jl:28-39)
    gives the dense product's df_Emi bit for bit on finite states; with a NaN in one cell the
    dense n x n product spreads it to every cell through the stored zeros of the other blocks
    (0 * NaN = NaN), the sparse one only to its own cell."""
    w = workloads.tiny(steps=3, emission=True, x_num=3, with_flux=False)
    dense = O.problem_from_workload(w)
    sparse = O.problem_from_workload(w, sparse_emission=True)
    fa = dense.solve(w.f0, w.dt, w.t_save(2))[0]
    fb = sparse.solve(w.f0, w.dt, w.t_save(2))[0]
    np.testing.assert_array_equal(fa, fb)
    f0 = np.array(w.f0)
    N = w.ps.n_momentum
    f0[2 * N + 1] = math.nan
    a, b = O.Stepper(dense, f0), O.Stepper(sparse, f0)
    a(0.0, 1.0, w.dt, 0)
    b(0.0, 1.0, w.dt, 0)
    assert np.all(np.isnan(a.pre_clamp))
    assert not np.any(np.isnan(b.pre_clamp[:2 * N]))   # cells 0, 1 untouched by the sparse one
    assert np.all(np.isnan(b.pre_clamp[2 * N:]))


def test_binary_csc_is_the_monolithic_tensor():
    """The sparse M_Bin handed to dtx_set_binary_csc holds exactly the dense tensor's non-zero
    entries at M_Bin[(i-1)N + k, j] (1-based), rows ascending in each column."""
    w = workloads.tiny(steps=2)
    colptr, rowval, nzval = workloads.binary_csc(w)
    N = w.ps.n_momentum
    D = np.zeros((N * N, N))
    for j in range(N):
        for q in range(colptr[j] - 1, colptr[j + 1] - 1):
            D[rowval[q] - 1, j] += nzval[q]
        assert np.all(np.diff(rowval[colptr[j] - 1:colptr[j + 1] - 1]) > 0)
    np.testing.assert_array_equal(D, w.M_Bin.reshape(N * N, N, order="F"))
    assert nzval.size == np.count_nonzero(w.M_Bin)
