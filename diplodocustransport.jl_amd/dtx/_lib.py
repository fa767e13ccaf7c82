"""ctypes binding of libdtx.so (the C ABI declared in include/dtx.h).

The library is loaded from the package directory.  There is no fallback: if libdtx.so is missing
or fails to load, every entry point raises, so a GPU run can never silently use another path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

MAX_SPECIES = 16
PKG_DIR = Path(__file__).resolve().parent.parent
# the library next to the package, or the one DTX_LIBRARY names (as the Julia glue reads it)
LIB_PATH = Path(os.environ.get("DTX_LIBRARY", PKG_DIR / "libdtx.so"))

DTX_OK = 0
ERRORS = {-1: "DTX_ERR_INVALID", -2: "DTX_ERR_OOM", -3: "DTX_ERR_HIP", -4: "DTX_ERR_STATE",
          -5: "DTX_ERR_UNSUPPORTED", -6: "DTX_ERR_BUSY"}
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}
METHOD_FORWARD_EULER, METHOD_FORWARD_SYMPLECTIC_EULER = 0, 1
FLUX_F, FLUX_X, FLUX_P = 0, 1, 2

I32 = C.c_int32
I64 = C.c_int64
U64 = C.c_uint64
D = C.c_double
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)
PI32 = C.POINTER(C.c_int32)


class Config(C.Structure):
    _fields_ = [("num_species", I32), ("px_num", I32 * MAX_SPECIES),
                ("py_num", I32 * MAX_SPECIES), ("pz_num", I32 * MAX_SPECIES),
                ("x_num", I32), ("y_num", I32), ("z_num", I32), ("precision", I32),
                ("method", I32), ("device", I32), ("dt0", D)]


class StepStats(C.Structure):
    _fields_ = [("finite", I32), ("cr_evaluated", I32), ("save", I32), ("step_index", I32),
                ("cr", D), ("cr_bin", D), ("cr_emi", D), ("cr_flux", D), ("cr_momentum", D),
                ("cr_space", D), ("dt", D), ("t_start", D), ("sum_f", D)]


class WorkInfo(C.Structure):
    _fields_ = [("bin_flops", D), ("bin_bytes", D), ("update_bytes", D), ("gain_bytes", D),
                ("pairs", I64), ("bin_launch_items", I32), ("fused_correction", I32),
                ("emission_bytes", D)]


class KernelTimes(C.Structure):
    _fields_ = [("bin_ms", D), ("update_ms", D), ("bin_launches", I64),
                ("update_launches", I64), ("emission_ms", D), ("emission_launches", I64)]


class ServerConfig(C.Structure):
    _fields_ = [("device", I32), ("precision", I32), ("hbm_budget", I64), ("max_active", I32),
                ("no_moment_correction", I32)]


class Request(C.Structure):
    _fields_ = [("cfg", Config), ("seed", U64), ("steps", I64), ("theta", D),
                ("with_flux", I32)]


class RequestResult(C.Structure):
    _fields_ = [("id", I64), ("status", I32), ("submit_ms", D), ("start_ms", D),
                ("finish_ms", D), ("number_total", D), ("energy_total", D),
                ("steps_done", I64)]


class FluxSynth(C.Structure):
    _fields_ = [("xr", PD), ("yr", PD), ("zr", PD), ("pr", PD), ("ur", PD), ("hr", PD),
                ("mass", D * MAX_SPECIES), ("charge", D * MAX_SPECIES), ("sync_K", D),
                ("axes", I32), ("periodic", I32)]


BIN_ANI, BIN_AXI, BIN_ISO = 0, 1, 2


class BinaryInteraction(C.Structure):
    _fields_ = [("species", I32 * 4), ("mode", I32), ("sparse", I32), ("gain_scale", D),
                ("loss_scale", D), ("gain3", PD), ("gain4", PD), ("loss1", PD), ("loss2", PD),
                ("dpy", PD), ("dpz", PD)]


# (name, restype, argtypes) for every entry point of include/dtx.h
SIGNATURES = [
    ("dtx_create", I32, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    ("dtx_destroy", None, [C.c_void_p]),
    ("dtx_last_error", C.c_char_p, [C.c_void_p]),
    ("dtx_set_flux_csc", I32, [C.c_void_p, I32, I64, PI64, PI64, PD]),
    ("dtx_set_binary_dense", I32, [C.c_void_p, I64, PD]),
    ("dtx_set_binary_csc", I32, [C.c_void_p, I64, PI64, PI64, PD]),
    ("dtx_add_binary_gain", I32, [C.c_void_p, I32, I32, I32, PD]),
    ("dtx_add_binary_loss", I32, [C.c_void_p, I32, I32, PD]),
    ("dtx_add_binary_interaction", I32, [C.c_void_p, C.POINTER(BinaryInteraction)]),
    ("dtx_set_binary_domain", I32, [C.c_void_p, I64, PI64]),
    ("dtx_set_emission_dense", I32, [C.c_void_p, I64, PD]),
    ("dtx_set_emission_blocks", I32, [C.c_void_p, PD]),
    ("dtx_set_emission_csc", I32, [C.c_void_p, I64, PI64, PI64, PD]),
    ("dtx_set_vectors", I32, [C.c_void_p, PD, PD, PD, PD]),
    ("dtx_set_moment_weights", I32, [C.c_void_p, PD]),
    ("dtx_set_moment_correction", I32, [C.c_void_p, I32]),
    ("dtx_synth_binary", I32, [C.c_void_p, U64, D]),
    ("dtx_synth_flux", I32, [C.c_void_p, C.POINTER(FluxSynth)]),
    ("dtx_step", I32, [C.c_void_p, D, D, D, I32, PD, PI32, C.POINTER(StepStats)]),
    ("dtx_step_noalloc", I32, [C.c_void_p, D, D, D, I32, PD, PI32, C.POINTER(StepStats)]),
    ("dtx_solve", I32, [C.c_void_p, D, I64, PD, I32, PD, PD, PI64]),
    ("dtx_reset", I32, [C.c_void_p]),
    ("dtx_get_f", I32, [C.c_void_p, PD]),
    ("dtx_set_f", I32, [C.c_void_p, PD]),
    ("dtx_moments", I32, [C.c_void_p, PD, PD]),
    ("dtx_synchronize", I32, [C.c_void_p]),
    ("dtx_recent_stats", I32, [C.c_void_p, C.POINTER(StepStats), I32, PI32]),
    ("dtx_device_bytes", I64, [C.c_void_p]),
    ("dtx_work", I32, [C.c_void_p, C.POINTER(WorkInfo)]),
    ("dtx_kernel_timing", I32, [C.c_void_p, I32, I32, C.POINTER(KernelTimes)]),
    ("dtx_server_create", I32, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    ("dtx_server_submit", I32, [C.c_void_p, C.POINTER(Request), PI64]),
    ("dtx_server_submit_handle", I32, [C.c_void_p, C.c_void_p, I64, D, PI64]),
    ("dtx_server_poll", I32, [C.c_void_p, I64, C.POINTER(RequestResult)]),
    ("dtx_server_fetch", I32, [C.c_void_p, I64, PD, I64]),
    ("dtx_server_drain", I32, [C.c_void_p]),
    ("dtx_server_destroy", None, [C.c_void_p]),
    ("dtx_server_last_error", C.c_char_p, [C.c_void_p]),
    ("dtx_request_create", I32, [C.POINTER(Request), C.POINTER(ServerConfig),
                                 C.POINTER(C.c_void_p), PD]),
]

_lib = None


class DtxError(RuntimeError):
    pass


def load():
    """Load libdtx.so (once).  Raises if it is missing: there is no other implementation."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise DtxError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                       f"(make -C diplodocustransport.jl_amd/csrc)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return [s[0] for s in SIGNATURES]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous or a.flags.f_contiguous
    return a.ctypes.data_as(PD)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64
    return a.ctypes.data_as(PI64)


def check(rc: int, handle=None, what: str = ""):
    if rc != DTX_OK:
        lib = load()
        msg = lib.dtx_last_error(handle)
        msg = msg.decode() if msg else ""
        raise DtxError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")


def make_config(ps, precision="fp64", method=METHOD_FORWARD_EULER, device=0) -> Config:
    cfg = Config()
    cfg.num_species = ps.num_species
    for i, s in enumerate(ps.species):
        cfg.px_num[i] = s.p_num
        cfg.py_num[i] = s.u_num
        cfg.pz_num[i] = s.h_num
    cfg.x_num, cfg.y_num, cfg.z_num = ps.x_num, ps.y_num, ps.z_num
    cfg.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    cfg.method = method
    cfg.device = device
    cfg.dt0 = ps.dt0
    return cfg


def make_flux_synth(ps, sync_K=0.0, axes="x", periodic=True):
    """dtx_flux_synth for a PhaseSpace; returns (struct, arrays it points into)."""
    import numpy as np
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (
        ps.xr, ps.yr, ps.zr,
        np.concatenate([s.pr for s in ps.species]),
        np.concatenate([s.ur for s in ps.species]),
        np.concatenate([s.hr for s in ps.species]))]
    fs = FluxSynth()
    fs.xr, fs.yr, fs.zr, fs.pr, fs.ur, fs.hr = (dptr(a) for a in keep)
    for i, s in enumerate(ps.species):
        fs.mass[i] = s.mass
        fs.charge[i] = s.charge
    fs.sync_K = float(sync_K)
    fs.axes = (1 if "x" in axes else 0) | (2 if "y" in axes else 0)
    fs.periodic = 1 if periodic else 0
    return fs, keep
