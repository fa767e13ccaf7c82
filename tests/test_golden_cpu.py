"""The oracle reproduces the committed golden fixtures bit for bit (regression guard).  CPU."""
import numpy as np
import pytest

import oracle as O
from fixture_util import fixtures, load


@pytest.mark.parametrize("path", fixtures(), ids=lambda p: p.stem)
def test_oracle_reproduces_fixture(path):
    w, z = load(path)
    prob = O.problem_from_workload(w)
    fo, to, stats, _ = prob.solve(w.f0, w.dt, z["t_save"], verbose=1, method=int(z["method"][0]))
    np.testing.assert_array_equal(fo, z["f_out"])
    np.testing.assert_array_equal(to, z["t_out"])
    np.testing.assert_array_equal([s.cr for s in stats], z["cr"])
