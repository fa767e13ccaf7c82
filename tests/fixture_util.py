"""Rebuild a problem from a golden fixture (tests/golden/*.npz) without the generators."""
from pathlib import Path

import numpy as np

from dtx import grids
from dtx.workloads import Workload

GOLDEN = Path(__file__).resolve().parent / "golden"


def fixtures():
    return sorted(GOLDEN.glob("*.npz"))


def load(path):
    z = np.load(path, allow_pickle=False)
    names = ["Ele", "Pho", "Pos"]
    sp = [grids.Species(names[i], float(z["mass"][i]), -1.0 if z["mass"][i] else 0.0,
                        int(z["p_num"][i]), int(z["u_num"][i]), int(z["h_num"][i]))
          for i in range(len(z["p_num"]))]
    ps = grids.PhaseSpace(sp, x_num=int(z["x_num"][0]))
    csc = lambda p: (z[f"{p}_colptr"], z[f"{p}_rowval"], z[f"{p}_nzval"]) \
        if f"{p}_colptr" in z else None  # noqa: E731
    N = int(z["N"][0])
    w = Workload("golden", ps, z["f0"], z["inj"], z["Ap"], z["Vol"], csc("X"), csc("P"),
                 float(z["dt"][0]), 0, 0, M_Bin=z["M_Bin"].reshape((N, N, N), order="F"),
                 emission_blocks=z["emission_blocks"] if "emission_blocks" in z else None,
                 domain=list(z["domain"]) if "domain" in z else None)
    return w, z
