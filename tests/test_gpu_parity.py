"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle.

Bar: FP64 GPU vs oracle within 1e-12 relative (L-inf over the state, relative to the state's
L-inf norm) after whole Solve runs; elementwise semantics (clamps, Cr, finite flag) exact.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from dtx import workloads
from dtx.api import Solve

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def rel_linf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def _unpack(raw, rowstep):
    """Place each lane's accumulator element at its (row, col) of the 16x16 tile."""
    D = np.full((16, 16), np.nan)
    for l in range(64):
        for e in range(4):
            row = 4 * (l // 16) + e if rowstep == 1 else l // 16 + 4 * e
            D[row, l % 16] = raw[l * 4 + e]
    return D


def test_mfma_lane_layouts():
    """The result layouts the contraction kernels' epilogues assume: FP64 16x16x4 holds
    D[(l/16)+4e][l%16]; FP32 16x16x4 and BF16 16x16x32 hold D[4(l/16)+e][l%16]."""
    lib = C.CDLL(str(Path(__file__).parent / "gpu_probe" / "libmfma_probe.so"))
    lib.mfma_probe.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    raw = np.zeros(256)
    assert lib.mfma_probe(0, A.ctypes.data, B.ctypes.data, raw.ctypes.data) == 0
    np.testing.assert_allclose(_unpack(raw, 4), A @ B, rtol=1e-14, atol=1e-14)
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    raw32 = np.zeros(256, np.float32)
    assert lib.mfma_probe(1, A32.ctypes.data, B32.ctypes.data, raw32.ctypes.data) == 0
    np.testing.assert_allclose(_unpack(raw32, 1), A32.astype(np.float64) @ B32.astype(np.float64),
                               rtol=1e-5, atol=1e-5)
    Ab = rng.standard_normal((16, 32)).astype(np.float32)
    Bb = rng.standard_normal((32, 16)).astype(np.float32)
    to_bits = lambda x: (x.view(np.uint32) >> 16).astype(np.uint16)  # noqa: E731
    Ab = (to_bits(Ab).astype(np.uint32) << 16).view(np.float32)  # exactly representable
    Bb = (to_bits(Bb).astype(np.uint32) << 16).view(np.float32)
    rawb = np.zeros(256, np.float32)
    Abits = np.ascontiguousarray(to_bits(Ab))
    Bbits = np.ascontiguousarray(to_bits(Bb).T)  # column-major B
    assert lib.mfma_probe(2, Abits.ctypes.data, Bbits.ctypes.data, rawb.ctypes.data) == 0
    np.testing.assert_allclose(_unpack(rawb, 1), Ab.astype(np.float64) @ Bb.astype(np.float64),
                               rtol=1e-5, atol=1e-4)


def test_lane_exchanges():
    """The DPP / permlane lane exchanges of csrc/wave_ops.h move exactly what __shfl_xor and
    the quad broadcast would: lane l receives lane l ^ m (m = 1..32) or lane 4 (l // 4) + q,
    bit for bit, on doubles, 64-bit integers and ints."""
    lib = C.CDLL(str(Path(__file__).parent / "gpu_probe" / "libmfma_probe.so"))
    lib.lane_probe.argtypes = [C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(1)
    x = rng.standard_normal(64) * 10.0 ** rng.integers(-30, 30, 64)
    out = np.zeros((20, 64))
    assert lib.lane_probe(x.ctypes.data, out.ctypes.data) == 0
    lanes = np.arange(64)
    u = (lanes.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
    for i, m in enumerate((1, 2, 4, 8, 16, 32)):
        assert np.array_equal(out[i].view(np.uint64), x[lanes ^ m].view(np.uint64)), m
        assert np.array_equal(out[10 + i], u[lanes ^ m].astype(np.float64)), m
    for q in range(4):
        src = 4 * (lanes // 4) + q
        assert np.array_equal(out[6 + q].view(np.uint64), x[src].view(np.uint64)), q
        assert np.array_equal(out[16 + q], (1000 + 7 * src).astype(np.float64)), q


def _solve_both(w, n_save=3, verbose=1, method="euler", mono=None, precision="fp64"):
    m = workloads.make_struct(w, precision=precision, method=method, use_monolithic=mono)
    out = Solve(m, w.dt, w.t_save(n_save), Verbose=verbose)
    prob = O.problem_from_workload(w, use_monolithic=mono)
    fo, to, stats, nsteps = prob.solve(w.f0, w.dt, w.t_save(n_save), verbose=verbose,
                                       method=0 if method == "euler" else 1)
    return m, out, fo, to, stats


@pytest.mark.parametrize("mono", [True, False])
def test_tiny_solve_parity_fp64(mono):
    w = workloads.tiny(steps=30)
    m, out, fo, to, _ = _solve_both(w, mono=mono)
    np.testing.assert_array_equal(out.t, to)
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, i


def test_tiny_single_step_binary_and_stats():
    w = workloads.tiny(steps=5)
    m = workloads.make_struct(w)
    prob = O.problem_from_workload(w)
    st = O.Stepper(prob, w.f0)
    t = 0.0
    for k in range(5):
        dt_g, save_g, sg = m(t, 1.0, w.dt, 3, stats=True)
        dt_o, save_o, so = st(t, 1.0, w.dt, 3)
        assert dt_g == dt_o and save_g == save_o
        assert sg.finite == so.finite
        for name in ("cr", "cr_bin", "cr_flux"):
            a, b = getattr(sg, name), getattr(so, name)
            assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), (k, name, a, b)
        assert rel_linf(m.f, st.f) < TOL64
        t += dt_g


@pytest.mark.parametrize("kw", [dict(emission=True), dict(injection=True),
                                dict(domain=[0, 2]), dict(num_species=1),
                                dict(num_species=3, x_num=2)])
def test_tiny_variants_parity(kw):
    w = workloads.tiny(steps=25, **kw)
    m, out, fo, to, _ = _solve_both(w)
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, (kw, i)


def test_tiny_symplectic_parity():
    w = workloads.tiny(steps=25, injection=True)
    m, out, fo, to, _ = _solve_both(w, method="symplectic")
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, i


def test_moments_match_oracle():
    w = workloads.tiny(steps=10)
    m, out, fo, to, _ = _solve_both(w)
    num, en = m.moments()
    N = w.ps.n_momentum
    for c in range(w.ps.n_space):
        for s, (off, sp) in enumerate(zip(w.ps.species_offset, w.ps.species)):
            blk = fo[-1][c * N + off: c * N + off + sp.size]
            n0, e0 = O.moments(blk, sp)
            assert abs(num[c, s] - n0) <= 1e-12 * abs(n0)
            assert abs(en[c, s] - e0) <= 1e-12 * abs(e0)


def test_medium_transport_parity_fp64():
    # 2 species x (8 p x 4 u) = 64 states per cell, 24 cells, 200 steps, x advection + cooling
    w = workloads.transport_1d(p_num=8, u_num=4, x_num=24, steps=200, monolithic=True)
    m, out, fo, to, _ = _solve_both(w, n_save=5)
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, i


@pytest.mark.parametrize("shape", [(12, 8, 130), (8, 4, 100)], ids=["2ktiles", "1ktile"])
def test_many_cells_parity_fp64(shape):
    # more than 64 cells: the FP64 gain items take two k tiles x 128 cells.  (12, 8, 130): 96
    # states per species, so the second k tile of each item is partial, and the second cell
    # tile holds 2 cells; (8, 4, 100): 32 states per species, so the second k tile is absent
    p_num, u_num, x_num = shape
    w = workloads.transport_1d(p_num=p_num, u_num=u_num, x_num=x_num, steps=12, monolithic=False)
    m, out, fo, to, _ = _solve_both(w, n_save=2, mono=False)
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, i


@pytest.mark.parametrize("x_num", [20, 32])
def test_32_cell_tiles_parity_fp64(x_num):
    # 17-32 cells: 32-cell tiles (20: a partial tile), 160 states per species in channel form
    # (three 64-row k tiles, the last partial), with flux
    w = workloads.transport_1d(p_num=20, u_num=8, x_num=x_num, steps=12, monolithic=False)
    m, out, fo, to, _ = _solve_both(w, n_save=2, mono=False)
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, (x_num, i)


@pytest.mark.parametrize("kw", [dict(p_num=20, u_num=8, monolithic=False),
                                dict(num_species=3, emission=True, injection=True)],
                         ids=["3ktiles", "3species"])
def test_single_cell_parity_fp64(kw):
    # one cell: the contraction runs as GEMV items.  (20, 8): 160 states per species, so three
    # 64-row k tiles, the last partial
    w = workloads.tiny(x_num=1, steps=20, **kw)
    m, out, fo, to, _ = _solve_both(w, n_save=2, mono=kw.get("monolithic", True))
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], fo[i]) < TOL64, (kw, i)


@pytest.mark.parametrize("case", ["odd_N", "transport_1d_grid", "single_cell", "symplectic"])
def test_emission_gemv_parity(case):
    """The emission GEMV (16-byte loads of two rows for even N, 8-byte loads for odd N; with
    few cells the columns split over several workgroups whose sums are added by a second
    launch) against the oracle's block GEMV, in both step methods."""
    mono, method = None, "euler"
    if case == "odd_N":  # 3 species x 3 states: N = 9
        w = workloads.tiny(num_species=3, p_num=3, u_num=1, x_num=3, steps=10, emission=True,
                           injection=True)
    elif case == "transport_1d_grid":  # N = 512 in 4 cells: column ranges over workgroups
        w = workloads.transport_1d(x_num=4, steps=6, emission=True)
        mono = False
    elif case == "single_cell":  # N = 256, one cell
        w = workloads.tiny(p_num=16, u_num=8, x_num=1, steps=10, emission=True,
                           monolithic=False)
        mono = False
    else:
        w = workloads.tiny(steps=20, emission=True, injection=True)
        method = "symplectic"
    assert (w.ps.n_momentum % 2 == 1) == (case == "odd_N")
    m, out, fo, to, _ = _solve_both(w, n_save=2, method=method, mono=mono)
    for i in range(len(out.f)):
        assert np.all(np.isfinite(out.f[i]))
        assert rel_linf(out.f[i], fo[i]) < TOL64, (case, i)


@pytest.mark.parametrize("path", __import__("fixture_util").fixtures(), ids=lambda p: p.stem)
def test_golden_fixtures_fp64(path):
    """The HIP path against the committed golden outputs (oracle + brute force agreed on them)."""
    from fixture_util import load
    w, z = load(path)
    method = "euler" if int(z["method"][0]) == 0 else "symplectic"
    m = workloads.make_struct(w, precision="fp64", method=method)
    out = Solve(m, w.dt, z["t_save"], Verbose=1)
    np.testing.assert_array_equal(out.t, z["t_out"])
    for i in range(len(out.f)):
        assert rel_linf(out.f[i], z["f_out"][i]) < TOL64, (path.stem, i)


def test_sampled_kernel_timing():
    """Kernel timing of every k-th step records k-times fewer launches, and timing does not
    change the state."""
    w = workloads.tiny(steps=12)
    a = workloads.make_struct(w)
    b = workloads.make_struct(w)
    a.kernel_timing(enable=True, reset=True, every=4)
    t = 0.0
    for _ in range(12):
        dt, _ = a(t, 1e300, w.dt, 0)
        b(t, 1e300, w.dt, 0)
        t += dt
    kt = a.kernel_timing(enable=False, reset=True)
    assert kt.bin_launches == 3 and kt.update_launches == 3
    assert kt.bin_ms > 0.0 and kt.update_ms > 0.0
    np.testing.assert_array_equal(a.f, b.f)
    a.close()
    b.close()


def test_status_of_unsynchronised_steps():
    """Steps enqueued without host synchronisation (each step's finalisation then runs inside
    the next step's update launch) report the same per-step status, moments and state as steps
    synchronised one by one."""
    w = workloads.tiny(steps=10, injection=True)
    a = workloads.make_struct(w)
    b = workloads.make_struct(w)
    t = 0.0
    sync_stats = []
    for _ in range(10):
        dt, _, st = a(t, 1e300, w.dt, 3, stats=True)
        b(t, 1e300, w.dt, 3)
        sync_stats.append(st)
        t += dt
    later = b.recent_stats(10)
    assert len(later) == 10
    for x, y in zip(sync_stats, later):
        for name, _ in x._fields_:
            assert getattr(x, name) == getattr(y, name), name
    np.testing.assert_array_equal(a.f, b.f)
    for u, v in zip(a.moments(), b.moments()):
        np.testing.assert_array_equal(u, v)
    a.close()
    b.close()


def test_solve_writes_solution_file(tmp_path):
    """Solve with fileName saves the output and the PhaseSpace (src/Solver.jl:53-59); loading
    the file gives back the same states and times."""
    from dtx import SolutionFileLoad
    w = workloads.tiny(steps=4)
    m = workloads.make_struct(w)
    out = Solve(m, w.dt, w.t_save(3), Verbose=0, fileName="tiny.jld2",
                fileLocation=str(tmp_path))
    ps, sol = SolutionFileLoad(str(tmp_path), "tiny.jld2")
    assert ps == w.ps
    assert len(sol.f) == len(out.f) == 3
    for a, b in zip(sol.f, out.f):
        assert np.array_equal(a, b)
    assert np.array_equal(sol.t, np.asarray(out.t))
