"""dtx_step_noalloc (the entry point that allocates nothing) against dtx_step through the C ABI:
two runs in a row on the same device, the second on a smaller problem than the first, each
compared bit for bit (state and step statistics) with identical handles stepped by dtx_step;
then the refusals (operators not packed, a step whose partial sums exceed the device array)."""
import ctypes as C

import numpy as np
import pytest

from dtx import _lib, workloads

pytestmark = pytest.mark.gpu


def _noalloc_step(m, t, dt, verbose, stats=False):
    lib = _lib.load()
    dt_out, save = C.c_double(), C.c_int32()
    st = _lib.StepStats() if stats else None
    rc = lib.dtx_step_noalloc(m._h, t, 1e300, dt, verbose, C.byref(dt_out), C.byref(save),
                              C.byref(st) if st is not None else None)
    return rc, dt_out.value, st


def _compare_runs(w, precision, method, steps, verbose):
    ref = workloads.make_struct(w, precision=precision, method=method, use_monolithic=False)
    m = workloads.make_struct(w, precision=precision, method=method, use_monolithic=False)
    m.work()  # packs the operators, so the no-alloc steps allocate nothing
    t = 0.0
    for k in range(steps):
        dt_r, _ = ref(t, 1e300, w.dt, verbose)
        rc, dt_m, _ = _noalloc_step(m, t, w.dt, verbose)
        assert rc == 0, _lib.load().dtx_last_error(m._h)
        assert dt_m == dt_r
        t += dt_r
    np.testing.assert_array_equal(m.f, ref.f)
    for x, y in zip(m.recent_stats(steps), ref.recent_stats(steps)):
        for name, _ in x._fields_:
            assert getattr(x, name) == getattr(y, name), name
    nm, em = m.moments()
    nr, er = ref.moments()
    np.testing.assert_array_equal(nm, nr)
    np.testing.assert_array_equal(em, er)
    m.close()
    ref.close()


def test_noalloc_steps_equal_dtx_step_two_runs_second_smaller():
    # first run: the full transport_1d problem (65,536 unknowns) with the FP32 contraction and
    # its FP64 moment-rate correction, whose slot sums and moment-rate items use the array
    big = workloads.transport_1d(steps=10)
    _compare_runs(big, "fp32", "euler", 10, verbose=1)
    # second run, smaller, right after: three cells, emission and injection, symplectic, FP64;
    # the array still holds the first run's partial sums
    small = workloads.tiny(steps=10, emission=True, injection=True, monolithic=False)
    _compare_runs(small, "fp64", "symplectic", 10, verbose=3)


def test_noalloc_refusals():
    lib = _lib.load()
    w = workloads.tiny(steps=3, monolithic=False)
    m = workloads.make_struct(w, use_monolithic=False)
    rc, _, _ = _noalloc_step(m, 0.0, w.dt, 0)
    assert rc == -4  # DTX_ERR_STATE: operators not packed yet
    assert b"not packed" in lib.dtx_last_error(m._h)
    m.close()
    # transport_2d with the split-BF16 contraction: ~54 partial slots of C x N x 8 = 8 MiB
    big = workloads.transport_2d(steps=1)
    h = workloads.make_struct(big, precision="bf16x3", device_flux=True)
    h.work()
    rc, _, _ = _noalloc_step(h, 0.0, big.dt, 0)
    assert rc == -5, rc  # DTX_ERR_UNSUPPORTED, refused before any launch
    assert b"more than the" in lib.dtx_last_error(h._h)
    # the handle is untouched and still steps through dtx_step
    h(0.0, 1e300, big.dt, 0)
    assert np.all(np.isfinite(h.f))
    h.close()
