"""The C-ABI library: it loads, exports every entry point include/dtx.h declares, and the ctypes
binding covers them all.  No compute calls (CPU only)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from dtx import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "dtx.h"


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-zA-Z_][a-zA-Z0-9_]*\s*\*?\s*(dtx_[a-z0-9_]+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = header_functions()
    for required in ("dtx_create", "dtx_destroy", "dtx_set_flux_csc", "dtx_set_binary_dense",
                     "dtx_add_binary_gain", "dtx_add_binary_loss", "dtx_set_vectors",
                     "dtx_step", "dtx_solve", "dtx_get_f", "dtx_moments", "dtx_server_submit"):
        assert required in names


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    for name in header_functions():
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\bT\s+(dtx_[a-z0-9_]+)", nm))
    assert set(header_functions()) <= exported


def test_ctypes_binding_covers_the_header():
    assert set(header_functions()) == set(_lib.exported_symbols())


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """Every struct of the ABI has the same size and field offsets in ctypes as in C."""
    structs = {"dtx_config": _lib.Config, "dtx_step_stats": _lib.StepStats,
               "dtx_work_info": _lib.WorkInfo, "dtx_kernel_times": _lib.KernelTimes,
               "dtx_server_config": _lib.ServerConfig, "dtx_request": _lib.Request,
               "dtx_request_result": _lib.RequestResult, "dtx_flux_synth": _lib.FluxSynth,
               "dtx_binary_interaction": _lib.BinaryInteraction}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, value = line.split()
        py = structs[cname]
        if field == "size":
            assert C.sizeof(py) == int(value), cname
        else:
            assert getattr(py, field).offset == int(value), (cname, field)


def test_create_without_gpu_fails_cleanly():
    """Without a GPU, dtx_create returns an error code (no crash) and explains why."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    from dtx import grids
    ps = grids.PhaseSpace([grids.Species("Ele", 1.0, -1.0, 2, 1, 1)])
    cfg = _lib.make_config(ps)
    h = C.c_void_p()
    rc = lib.dtx_create(C.byref(cfg), C.byref(h))
    assert rc != 0 and not h.value
    assert lib.dtx_last_error(None)


def test_create_rejects_invalid_config():
    lib = _lib.load()
    from dtx import grids
    ps = grids.PhaseSpace([grids.Species("Ele", 1.0, -1.0, 2, 1, 1)])
    cfg = _lib.make_config(ps)
    cfg.num_species = 0
    h = C.c_void_p()
    assert lib.dtx_create(C.byref(cfg), C.byref(h)) == -1
    assert b"num_species" in lib.dtx_last_error(None)
    cfg = _lib.make_config(ps)
    cfg.dt0 = 0.0
    assert lib.dtx_create(C.byref(cfg), C.byref(h)) == -1
