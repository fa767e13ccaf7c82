"""SolutionFileSave / SolutionFileLoad (src/DataReading.jl:6-41): the saved PhaseSpace and
solution come back unchanged, and a missing file raises the reference's error."""
import numpy as np
import pytest

from dtx import OutputStruct, SolutionFileLoad, SolutionFileSave
from dtx.grids import PhaseSpace
from dtx.workloads import species_pair


def test_save_load_round_trip(tmp_path):
    ps = PhaseSpace(species_pair(4, 2, 2), x_num=3, y_num=2, x_up=2.5, dt0=0.5)
    rng = np.random.default_rng(5)
    sol = OutputStruct(f=[rng.standard_normal(ps.n) for _ in range(3)],
                       t=np.array([0.0, 0.1, 0.30000000000000004]))
    SolutionFileSave(sol, ps, str(tmp_path), "run.jld2")
    ps2, sol2 = SolutionFileLoad(str(tmp_path), "run.jld2")
    assert ps2 == ps
    assert ps2.n == ps.n and np.array_equal(ps2.species[0].pr, ps.species[0].pr)
    assert len(sol2.f) == 3
    for a, b in zip(sol2.f, sol.f):
        assert np.array_equal(a, b)
    assert np.array_equal(sol2.t, sol.t)
    # overwrites an existing file, as jldopen(path, "w") does
    SolutionFileSave(OutputStruct(f=[sol.f[0]], t=sol.t[:1]), ps, str(tmp_path), "run.jld2")
    assert len(SolutionFileLoad(str(tmp_path), "run.jld2")[1].f) == 1


def test_load_missing_file(tmp_path):
    with pytest.raises(RuntimeError, match="no file with name nope.jld2 found at location"):
        SolutionFileLoad(str(tmp_path), "nope.jld2")


def test_solve_creates_missing_file_location(tmp_path):
    """Solve creates a missing fileLocation before stepping (src/Solver.jl:3-5), so the save at
    the end cannot fail on the directory; a stand-in method records that the directory was
    already there when the time loop started."""
    from dtx import Solve
    ps = PhaseSpace(species_pair(2, 1, 1), x_num=1)
    where = tmp_path / "a" / "b"

    class Method:
        PhaseSpace = ps
        seen = None

        def solve_native(self, dt, t_save, Verbose):
            Method.seen = where.is_dir()
            return OutputStruct(f=[np.zeros(ps.n) for _ in t_save], t=np.asarray(t_save)), 1

    out = Solve(Method(), 0.1, [0.0, 0.1], fileName="x.jld2", fileLocation=str(where),
                Verbose=0)
    assert Method.seen is True
    ps2, sol = SolutionFileLoad(str(where), "x.jld2")
    assert ps2 == ps and len(sol.f) == len(out.f) == 2
