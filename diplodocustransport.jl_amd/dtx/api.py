"""Python mirror of DiplodocusTransport.jl's stepping API on the HIP backend.

Same names, argument meaning and error behaviour as the reference's Julia surface:

* ``ForwardEulerStruct(PhaseSpace, Initial, Injection, BinM, EmiM, FluxM; Adaptive)``
  (src/SteppingStructs.jl:33) and ``ForwardSymplecticEulerStruct`` (:142),
* the functor ``method(t_start, t_stop, dt, Verbose) -> (dt, save)``
  (src/SteppingMethods.jl:12, 120),
* ``Solve(method, dt_initial, t_save; Verbose)`` (src/Solver.jl:1) returning ``OutputStruct``
  (src/StructsAndDictionaries.jl:362-377),
* the matrix containers BinaryMatricesStruct / EmissionMatricesStruct / FluxMatricesStruct
  (src/StructsAndDictionaries.jl:316-359),
* number N^0 and energy T^00 moments (src/DistributionMoments.jl:8-42, 94-212).

Every numerical call goes through libdtx.so (the C ABI of include/dtx.h); this module only
marshals arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .grids import PhaseSpace


class BackendType:
    pass


class CPUBackend(BackendType):
    """Present for API parity; this package only implements the HIP backend."""


class HIPBackend(BackendType):
    """The MI355X backend: state and operators live in HBM, steps run as gfx950 kernels."""

    def __init__(self, device: int = 0):
        self.device = device


@dataclass
class BinaryMatricesStruct:
    """M_Bin as the (N^2 x N) column-major tensor (any array whose Fortran-order memory is
    M_Bin, e.g. a 3-D M[k,i,j] Fortran array) or as a sparse N^2 x N matrix (Bin_sparse: a
    1-based (colptr, rowval, nzval) CSC tuple or a scipy sparse matrix), and/or the channel form
    (gains, losses).
    Binary_list non-empty <=> binary interactions are on (src/SteppingStructs.jl:46)."""
    M_Bin: Optional[np.ndarray] = None
    Binary_list: list = field(default_factory=list)
    Domain: Optional[list] = None
    gains: list = field(default_factory=list)     # dtx.synth.GainChannel
    losses: list = field(default_factory=list)    # dtx.synth.LossChannel
    synth_seed: Optional[int] = None              # generate on the device instead
    synth_scale: float = 1.0
    # collision arrays per interaction (dtx.collisions.BuildBinaryMatrices): the operator is
    # assembled from them by dtx_add_binary_interaction
    interactions: list = field(default_factory=list)
    Bin_Mode: object = None
    Bin_sparse: bool = False


@dataclass
class EmissionMatricesStruct:
    """M_Emi dense n x n (column-major, block diagonal per cell), sparse n x n (Emi_sparse: a
    1-based CSC tuple or a scipy sparse matrix), or per-cell blocks."""
    M_Emi: Optional[np.ndarray] = None
    Emission_list: list = field(default_factory=list)
    blocks: Optional[np.ndarray] = None


@dataclass
class FluxMatricesStruct:
    """Ap_Flux (n), X_Flux / P_Flux as (colptr, rowval, nzval) 1-based CSC tuples, Vol (C).
    With `synth` = dict(sync_K=..., axes="x"|"y"|"xy", periodic=...), the synthetic Cartesian
    stencils (dtx/synth.py flux_matrices, ap_vol) are built on the device instead
    (dtx_synth_flux); the matrices and Ap / Vol given here are then ignored and may be None."""
    Ap_Flux: Optional[np.ndarray]
    X_Flux: Optional[tuple]
    P_Flux: Optional[tuple]
    Vol: Optional[np.ndarray]
    synth: Optional[dict] = None


@dataclass
class OutputStruct:
    f: list
    t: np.ndarray


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _is_sparse(M) -> bool:
    """A SparseMatrixCSC given as a 1-based (colptr, rowval, nzval) tuple, or a scipy sparse
    matrix."""
    return isinstance(M, tuple) or hasattr(M, "tocsc")


def _as_julia_csc(M, shape):
    """(colptr, rowval, nzval) 1-based Int64 CSC with rows ascending in each column."""
    if isinstance(M, tuple):
        return _csc_args(M)
    if tuple(M.shape) != tuple(shape):
        raise ValueError(f"sparse matrix is {M.shape}, expected {shape}")
    A = M.tocsc(copy=True)
    A.sort_indices()
    return _csc_args((A.indptr.astype(np.int64) + 1, A.indices.astype(np.int64) + 1, A.data))


def _csc_args(csc):
    colptr, rowval, nzval = csc
    colptr = np.ascontiguousarray(colptr, dtype=np.int64)
    rowval = np.ascontiguousarray(rowval, dtype=np.int64)
    nzval = _f64(nzval)
    return colptr, rowval, nzval


class _SteppingStruct:
    METHOD = _lib.METHOD_FORWARD_EULER

    def __init__(self, PhaseSpace: PhaseSpace, Initial, Injection, BinM: BinaryMatricesStruct,
                 EmiM: EmissionMatricesStruct, FluxM: FluxMatricesStruct,
                 Adaptive: bool = False, Backend: BackendType = None,
                 Precision: str = "fp64", MomentCorrection: bool = True):
        if Backend is None:
            Backend = HIPBackend()
        if not isinstance(Backend, HIPBackend):
            raise ValueError("Backend type not recognized.")  # src/SteppingStructs.jl:74
        self.lib = _lib.load()
        self.PhaseSpace = PhaseSpace
        self.Adaptive = Adaptive
        self.Implicit = False
        self.dt0 = float(PhaseSpace.dt0)
        self.Binary_Interactions = bool(BinM.Binary_list)
        self.Emission_Interactions = bool(EmiM.Emission_list)
        self.Bin_Domain = BinM.Domain
        self.f_init = _f64(Initial).copy()
        self.precision = Precision
        n = PhaseSpace.n
        if self.f_init.size != n:
            raise ValueError(f"Initial has length {self.f_init.size}, expected {n}")
        self._h = C.c_void_p()
        cfg = _lib.make_config(PhaseSpace, Precision, self.METHOD, Backend.device)
        _lib.check(self.lib.dtx_create(C.byref(cfg), C.byref(self._h)), None, "dtx_create")
        h = self._h
        # fluxes
        if FluxM.synth is not None:
            fs, keep = _lib.make_flux_synth(PhaseSpace, **FluxM.synth)
            _lib.check(self.lib.dtx_synth_flux(h, C.byref(fs)), h, "dtx_synth_flux")
            del keep
        elif self.METHOD == _lib.METHOD_FORWARD_EULER:
            F = self._flux_sum(n, FluxM)
            if F is not None:
                cp, rv, nz = _csc_args(F)
                _lib.check(self.lib.dtx_set_flux_csc(h, _lib.FLUX_F, n, _lib.i64ptr(cp),
                                                     _lib.i64ptr(rv), _lib.dptr(nz)), h,
                           "dtx_set_flux_csc")
        else:
            for which, M in ((_lib.FLUX_X, FluxM.X_Flux), (_lib.FLUX_P, FluxM.P_Flux)):
                if M is not None:
                    cp, rv, nz = _csc_args(M)
                    _lib.check(self.lib.dtx_set_flux_csc(h, which, n, _lib.i64ptr(cp),
                                                         _lib.i64ptr(rv), _lib.dptr(nz)), h,
                               "dtx_set_flux_csc")
        # binary
        if self.Binary_Interactions:
            if BinM.synth_seed is not None:
                _lib.check(self.lib.dtx_synth_binary(h, int(BinM.synth_seed),
                                                     float(BinM.synth_scale)), h,
                           "dtx_synth_binary")
            elif BinM.interactions:
                self._add_interactions(h, PhaseSpace, BinM)
            elif BinM.M_Bin is not None and _is_sparse(BinM.M_Bin):
                # Bin_sparse: the CSC goes to the library as is (no dense N^2 x N copy)
                N = PhaseSpace.n_momentum
                cp, rv, nz = _as_julia_csc(BinM.M_Bin, (N * N, N))
                _lib.check(self.lib.dtx_set_binary_csc(h, N, _lib.i64ptr(cp), _lib.i64ptr(rv),
                                                       _lib.dptr(nz)), h, "dtx_set_binary_csc")
            elif BinM.M_Bin is not None:
                M = np.asarray(BinM.M_Bin, dtype=np.float64)
                Mf = M.reshape(-1, order="F") if M.flags.f_contiguous else \
                    np.asfortranarray(M).reshape(-1, order="F")
                N = PhaseSpace.n_momentum
                if Mf.size != N ** 3:
                    raise ValueError("M_Bin is not the correct size")  # src/SteppingMethods.jl:349
                _lib.check(self.lib.dtx_set_binary_dense(h, N, _lib.dptr(Mf)), h,
                           "dtx_set_binary_dense")
            else:
                for g in BinM.gains:
                    G = np.asfortranarray(g.G, dtype=np.float64).reshape(-1, order="F")
                    _lib.check(self.lib.dtx_add_binary_gain(h, g.s_out, g.s1, g.s2,
                                                            _lib.dptr(G)), h,
                               "dtx_add_binary_gain")
                for l in BinM.losses:
                    L = np.asfortranarray(l.L, dtype=np.float64).reshape(-1, order="F")
                    _lib.check(self.lib.dtx_add_binary_loss(h, l.s1, l.s2, _lib.dptr(L)), h,
                               "dtx_add_binary_loss")
            if BinM.Domain is not None:
                dom = np.ascontiguousarray(BinM.Domain, dtype=np.int64)
                _lib.check(self.lib.dtx_set_binary_domain(h, dom.size, _lib.i64ptr(dom)), h,
                           "dtx_set_binary_domain")
        # emission
        if self.Emission_Interactions:
            if EmiM.blocks is not None:
                B = np.ascontiguousarray(EmiM.blocks, dtype=np.float64)
                _lib.check(self.lib.dtx_set_emission_blocks(h, _lib.dptr(B)), h,
                           "dtx_set_emission_blocks")
            elif _is_sparse(EmiM.M_Emi):  # Emi_sparse
                cp, rv, nz = _as_julia_csc(EmiM.M_Emi, (n, n))
                _lib.check(self.lib.dtx_set_emission_csc(h, n, _lib.i64ptr(cp), _lib.i64ptr(rv),
                                                         _lib.dptr(nz)), h,
                           "dtx_set_emission_csc")
            else:
                M = np.asfortranarray(EmiM.M_Emi, dtype=np.float64).reshape(-1, order="F")
                _lib.check(self.lib.dtx_set_emission_dense(h, n, _lib.dptr(M)), h,
                           "dtx_set_emission_dense")
        if FluxM.synth is not None:  # 1/Ap and Vol are already on the device
            ap = vol = None
        else:
            ap, vol = _f64(FluxM.Ap_Flux), _f64(FluxM.Vol)
        inj = _f64(Injection)
        _lib.check(self.lib.dtx_set_vectors(h, None if ap is None else _lib.dptr(ap),
                                            None if vol is None else _lib.dptr(vol),
                                            _lib.dptr(self.f_init), _lib.dptr(inj)), h,
                   "dtx_set_vectors")
        w = _f64(PhaseSpace.energy_weights())
        _lib.check(self.lib.dtx_set_moment_weights(h, _lib.dptr(w)), h,
                   "dtx_set_moment_weights")
        # FP32 / BF16x3: FP64 number and energy rates of the gain term (no effect in FP64);
        # MomentCorrection=2 keeps its three separate launches (the same arithmetic)
        mc = 2 if (not isinstance(MomentCorrection, bool) and MomentCorrection == 2) else \
            int(bool(MomentCorrection))
        _lib.check(self.lib.dtx_set_moment_correction(h, mc), h, "dtx_set_moment_correction")

    def _add_interactions(self, h, PhaseSpace, BinM):
        """Each resolved interaction through dtx_add_binary_interaction, in Binary_list order
        (the Fill_M_Bin! calls of LoadMatrices_Binary, src/Collisions/LoadMatrices.jl:136)."""
        from .collisions import Ani, angle_widths
        dpy, dpz = angle_widths(PhaseSpace)
        mode = (BinM.Bin_Mode or Ani()).code
        for it in BinM.interactions:
            arrs = [np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F")
                    for a in (it.GainMatrix3, it.GainMatrix4, it.LossMatrix1, it.LossMatrix2)]
            x = _lib.BinaryInteraction()
            x.species[:] = [int(v) for v in it.name_locs]
            x.mode = mode
            x.sparse = int(bool(BinM.Bin_sparse))
            x.gain_scale = float(it.gain_scale)
            x.loss_scale = float(it.loss_scale)
            x.gain3, x.gain4, x.loss1, x.loss2 = (_lib.dptr(a) for a in arrs)
            x.dpy, x.dpz = _lib.dptr(dpy), _lib.dptr(dpz)
            _lib.check(self.lib.dtx_add_binary_interaction(h, C.byref(x)), h,
                       "dtx_add_binary_interaction")
            del arrs  # the library has read them

    @staticmethod
    def _flux_sum(n, FluxM):
        from .synth import csc_add
        X, P = FluxM.X_Flux, FluxM.P_Flux
        if X is None and P is None:
            return None
        if X is None:
            return P
        if P is None:
            return X
        return csc_add(n, X, P)  # F_Flux = X_Flux + P_Flux (src/SteppingStructs.jl:60)

    # ---- functor ----------------------------------------------------------------------
    def __call__(self, t_start, t_stop, dt, Verbose: int, stats: bool = False):
        dt_out = C.c_double()
        save = C.c_int32()
        st = _lib.StepStats() if stats else None
        _lib.check(self.lib.dtx_step(self._h, float(t_start), float(t_stop), float(dt),
                                     int(Verbose), C.byref(dt_out), C.byref(save),
                                     C.byref(st) if st is not None else None), self._h,
                   "dtx_step")
        if stats:
            return dt_out.value, bool(save.value), st
        return dt_out.value, bool(save.value)

    # ---- state --------------------------------------------------------------------------
    @property
    def f(self) -> np.ndarray:
        out = np.empty(self.PhaseSpace.n)
        _lib.check(self.lib.dtx_get_f(self._h, _lib.dptr(out)), self._h, "dtx_get_f")
        return out

    @f.setter
    def f(self, value):
        v = _f64(value)
        _lib.check(self.lib.dtx_set_f(self._h, _lib.dptr(v)), self._h, "dtx_set_f")

    def reset(self):
        _lib.check(self.lib.dtx_reset(self._h), self._h, "dtx_reset")

    def synchronize(self):
        _lib.check(self.lib.dtx_synchronize(self._h), self._h, "dtx_synchronize")

    def moments(self):
        """(number, energy), each (C, num_species): N^0 and T^00 of the current f."""
        ps = self.PhaseSpace
        num = np.empty(ps.n_space * ps.num_species)
        en = np.empty_like(num)
        _lib.check(self.lib.dtx_moments(self._h, _lib.dptr(num), _lib.dptr(en)), self._h,
                   "dtx_moments")
        return num.reshape(ps.n_space, ps.num_species), en.reshape(ps.n_space, ps.num_species)

    def recent_stats(self, count: int):
        arr = (_lib.StepStats * count)()
        got = C.c_int32()
        _lib.check(self.lib.dtx_recent_stats(self._h, arr, count, C.byref(got)), self._h,
                   "dtx_recent_stats")
        return list(arr)[:got.value]

    def work(self) -> _lib.WorkInfo:
        w = _lib.WorkInfo()
        _lib.check(self.lib.dtx_work(self._h, C.byref(w)), self._h, "dtx_work")
        return w

    def kernel_timing(self, enable: bool, reset: bool = True, every: int = 1) -> _lib.KernelTimes:
        """Collect the kernel times of the window so far; from now on time the kernels of every
        `every`-th step (enable) or none."""
        t = _lib.KernelTimes()
        _lib.check(self.lib.dtx_kernel_timing(self._h, max(1, int(every)) if enable else 0,
                                              int(reset), C.byref(t)),
                   self._h, "dtx_kernel_timing")
        return t

    def device_bytes(self) -> int:
        return int(self.lib.dtx_device_bytes(self._h))

    def solve_native(self, dt_initial, t_save, Verbose=1):
        """Solve entirely inside the library (no per-step host synchronisation)."""
        t_save = _f64(t_save)
        n = self.PhaseSpace.n
        fout = np.empty((t_save.size, n))
        tout = np.empty(t_save.size)
        steps = C.c_int64()
        _lib.check(self.lib.dtx_solve(self._h, float(dt_initial), t_save.size,
                                      _lib.dptr(t_save), int(Verbose), _lib.dptr(fout),
                                      _lib.dptr(tout), C.byref(steps)), self._h, "dtx_solve")
        return OutputStruct([fout[i] for i in range(t_save.size)], tout), steps.value

    def release(self):
        """Give up ownership of the library handle (e.g. to dtx_server_submit_handle); this
        object is closed afterwards."""
        h, self._h = self._h, C.c_void_p()
        return h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.dtx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ForwardEulerStruct(_SteppingStruct):
    METHOD = _lib.METHOD_FORWARD_EULER


class ForwardSymplecticEulerStruct(_SteppingStruct):
    METHOD = _lib.METHOD_FORWARD_SYMPLECTIC_EULER


def Solve(method: _SteppingStruct, dt_initial: float, t_save, progress: bool = False,
          fileName=None, fileLocation=None, Verbose: int = 1) -> OutputStruct:
    """Solve (src/Solver.jl:1-64): the time loop runs inside libdtx (dtx_solve) with the same
    dt clipping and carry-over; with fileName, the output and the PhaseSpace are written by
    SolutionFileSave (:53-59).  A missing fileLocation is created before stepping, as the
    reference's mkpath (:3-5) does."""
    import os
    if fileLocation is not None and not os.path.isdir(fileLocation):
        os.makedirs(fileLocation, exist_ok=True)
    out, _ = method.solve_native(dt_initial, t_save, Verbose)
    if fileName is not None:
        print("Saving Simulation output")
        SolutionFileSave(out, method.PhaseSpace, "." if fileLocation is None else fileLocation,
                         fileName)
    return out


def SolutionFileSave(sol: OutputStruct, PhaseSpace: PhaseSpace, fileLocation: str,
                     fileName: str) -> None:
    """SolutionFileSave (src/DataReading.jl:31-41): the entries "PhaseSpace" and "sol" in one
    file at fileLocation/fileName, overwriting an existing file.  The container is a NumPy .npz
    without pickled objects (JLD2 is a Julia format): "sol.f" (n_save x n), "sol.t" (n_save)
    and "PhaseSpace" (the grid as JSON text)."""
    import json
    import os
    from dataclasses import asdict
    path = os.path.join(fileLocation, fileName)
    f = np.stack([np.asarray(x, dtype=np.float64) for x in sol.f]) if len(sol.f) else \
        np.zeros((0, PhaseSpace.n))
    with open(path, "wb") as fh:  # an explicit handle: np.savez would append ".npz"
        np.savez(fh, **{"sol.f": f, "sol.t": np.asarray(sol.t, dtype=np.float64),
                        "PhaseSpace": np.array(json.dumps(asdict(PhaseSpace)))})


def SolutionFileLoad(fileLocation: str, fileName: str):
    """SolutionFileLoad (src/DataReading.jl:6-28): (PhaseSpace, sol) from a file written by
    SolutionFileSave, with the reference's error for a missing file.  The path is joined with
    the platform's separator (the reference joins with a literal backslash, :8)."""
    import json
    import os
    from .grids import Species
    path = os.path.join(fileLocation, fileName)
    if not os.path.isfile(path):
        raise RuntimeError(f"no file with name {fileName} found at location {fileLocation}")
    with np.load(path, allow_pickle=False) as z:
        ps = json.loads(str(z["PhaseSpace"]))
        f, t = z["sol.f"], z["sol.t"]
    ps["species"] = [Species(**sp) for sp in ps["species"]]
    return PhaseSpace(**ps), OutputStruct(f=[row.copy() for row in f], t=t.copy())


def FourFlow_N0(state: np.ndarray, ps: PhaseSpace, species: int, x_idx=1, y_idx=1, z_idx=1):
    """N^0 of FourFlow (src/DistributionMoments.jl:33) for a host state vector."""
    return float(np.sum(ps.view(state, species, x_idx, y_idx, z_idx)))
