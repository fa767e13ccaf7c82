"""dtx -- MI355X-native explicit-Euler stepping for DiplodocusTransport.jl.

The numerical work runs in libdtx.so (gfx950 HIP kernels behind the C ABI of include/dtx.h).
This package holds the Python mirror of the reference's stepping API (api), the grid helpers
(grids) and the seeded synthetic workload generators (synth).
"""
from . import grids, synth  # noqa: F401
from ._lib import DtxError, load  # noqa: F401
from .api import (BinaryMatricesStruct, CPUBackend, EmissionMatricesStruct,  # noqa: F401
                  FluxMatricesStruct, ForwardEulerStruct, ForwardSymplecticEulerStruct,
                  HIPBackend, OutputStruct, Solve, SolutionFileLoad, SolutionFileSave)
