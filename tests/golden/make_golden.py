"""Writes the golden fixtures of tests/golden/*.npz.

Each fixture holds the inputs of a small explicit-Euler problem (state, operators, time grid)
and the states Solve saves.  The expected outputs are produced by the FP64 C oracle and
accepted only if the independent pure-Python brute force reproduces them bit for bit.
Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "diplodocustransport.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import brute  # noqa: E402
import oracle as O  # noqa: E402
from dtx import workloads  # noqa: E402

CASES = {
    "euler_flux_injection": dict(kw=dict(steps=12, injection=True), method=0, n_save=4),
    "euler_emission_domain": dict(kw=dict(steps=10, emission=True, domain=[0, 2]), method=0,
                                  n_save=3),
    "symplectic_flux_injection": dict(kw=dict(steps=12, injection=True), method=1, n_save=4),
    "euler_three_species": dict(kw=dict(steps=8, num_species=3, x_num=2), method=0, n_save=3),
}


def csc_arrays(prefix, csc):
    if csc is None:
        return {}
    return {f"{prefix}_colptr": csc[0], f"{prefix}_rowval": csc[1], f"{prefix}_nzval": csc[2]}


def main():
    out_dir = Path(__file__).resolve().parent
    for name, case in CASES.items():
        w = workloads.tiny(**case["kw"])
        ts = w.t_save(case["n_save"])
        prob = O.problem_from_workload(w)
        fo, to, stats, nsteps = prob.solve(w.f0, w.dt, ts, verbose=1, method=case["method"])
        st = brute.make_state(w)
        outs, dts = brute.solve(st, list(w.f0), w.dt, list(ts), verbose=1,
                                method=case["method"])
        assert len(dts) == nsteps, name
        for a, b in zip(outs, fo):
            assert np.array_equal(np.array(a), b), name
        arrays = dict(f0=w.f0, inj=w.inj, Ap=w.Ap, Vol=w.Vol, M_Bin=w.M_Bin.reshape(-1, order="F"),
                      dt=np.array([w.dt]), t_save=ts, f_out=fo, t_out=to,
                      cr=np.array([s.cr for s in stats]), method=np.array([case["method"]]),
                      N=np.array([w.ps.n_momentum]), C=np.array([w.ps.n_space]),
                      species_sizes=np.array([s.size for s in w.ps.species]),
                      p_num=np.array([s.p_num for s in w.ps.species]),
                      u_num=np.array([s.u_num for s in w.ps.species]),
                      h_num=np.array([s.h_num for s in w.ps.species]),
                      mass=np.array([s.mass for s in w.ps.species]),
                      x_num=np.array([w.ps.x_num]))
        arrays.update(csc_arrays("X", w.X))
        arrays.update(csc_arrays("P", w.P))
        if w.emission_blocks is not None:
            arrays["emission_blocks"] = w.emission_blocks
        if w.domain is not None:
            arrays["domain"] = np.array(w.domain, dtype=np.int64)
        np.savez_compressed(out_dir / f"{name}.npz", **arrays)
        print(f"{name}: N={w.ps.n_momentum} C={w.ps.n_space} steps={nsteps} -> {name}.npz")


if __name__ == "__main__":
    main()
