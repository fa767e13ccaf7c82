"""Flux operators built on the device (dtx_synth_flux) against the host-assembled CSC uploaded
through dtx_set_flux_csc (dtx/synth.py flux_matrices + ap_vol, the restatement of
src/Fluxes/BuildFluxMatrices.jl:10-184 that the CPU tests pin).

Both handles step the same state with the same binary operators, so any difference in a flux
entry, in 1/Ap or in Vol shows in f.  The ELL width (the longest row after duplicate sums) shows
in the device bytes, which must agree.  Bar: 1e-13 relative (L-inf); the x stencil uses only
sqrt and IEEE division and is expected to agree bit for bit, the y stencil and the synchrotron
forces (I along p, J along u) call cos, asin, pow and asinh, whose device versions may differ
from the host's by an ulp.
"""
import numpy as np
import pytest

from dtx import synth, workloads
from dtx.api import (BinaryMatricesStruct, EmissionMatricesStruct, FluxMatricesStruct,
                     ForwardEulerStruct, Solve)
from dtx.grids import PhaseSpace

pytestmark = pytest.mark.gpu
TOL = 1e-13


def rel_linf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _compare(host, dev, dt, steps):
    bh = host.lib.dtx_device_bytes(host._h)
    bd = dev.lib.dtx_device_bytes(dev._h)
    t_save = np.array([0.0, steps * dt])
    fh = Solve(host, dt, t_save, Verbose=0).f[-1]
    fd = Solve(dev, dt, t_save, Verbose=0).f[-1]
    err = rel_linf(fd, fh)
    print(f"device bytes {bd} / {bh}, rel err {err:.3e}, bitwise {np.array_equal(fd, fh)}",
          flush=True)
    assert bd == bh, "ELL widths differ"
    assert np.all(np.isfinite(fd))
    assert err <= TOL


@pytest.mark.parametrize("case", ["x5_euler", "x2_euler", "x5_symplectic", "xy_euler",
                                  "transport_1d_full"])
def test_device_flux_matches_host_csc(case):
    method = "symplectic" if "symplectic" in case else "euler"
    if case == "transport_1d_full":  # 65,536 rows: the full transport_1d operator
        w = workloads.transport_1d(steps=6)
    elif case.startswith("xy"):
        w = workloads.transport_2d(p_num=4, u_num=4, h_num=2, x_num=4, y_num=3, steps=6)
    else:
        w = workloads.tiny(x_num=5 if case.startswith("x5") else 2, steps=6, monolithic=False)
    print(f"case {case}: n {w.ps.n}, spec {w.flux_spec}", flush=True)
    host = workloads.make_struct(w, method=method)
    dev = workloads.make_struct(w, method=method, device_flux=True)
    _compare(host, dev, w.dt, w.steps)


def test_device_flux_closed_boundaries_and_z_cells():
    """Closed x / y boundaries (clamped neighbours drop the boundary faces), z_num = 2 (the
    transverse widths and the cell order), non-uniform x cells; fluxes only."""
    ps = PhaseSpace(workloads.species_pair(4, 4, 2), x_num=3, y_num=4, z_num=2, x_up=2.0)
    spec = dict(sync_K=0.3, axes="xy", periodic=False)
    X, P = synth.flux_matrices(ps, **spec)
    Ap, Vol = synth.ap_vol(ps)
    F = synth.csc_to_dense(synth.csc_add(ps.n, X, P), ps.n)
    dt = 0.2 / float(np.max(np.abs(np.diag(F)) / Ap))
    f0 = synth.initial_state(ps, theta=1.0, seed=3, noise=0.05)
    print(f"closed xy z2: n {ps.n}, dt {dt:.3e}", flush=True)

    def build(FluxM):
        return ForwardEulerStruct(ps, f0, np.zeros(ps.n), BinaryMatricesStruct(Binary_list=[]),
                                  EmissionMatricesStruct(), FluxM)

    host = build(FluxMatricesStruct(Ap, X, P, Vol))
    dev = build(FluxMatricesStruct(None, None, None, None, synth=spec))
    _compare(host, dev, dt, 5)
