"""Known-answer tests from the reference's own docstrings (src/ValuesOnTheGrid.jl:52-267) and
closed forms of the moment weights.  CPU only."""
import numpy as np

from dtx import grids


def test_deltaVector_kat():
    # src/ValuesOnTheGrid.jl:59-65
    np.testing.assert_array_equal(grids.deltaVector([1.0, 10.0, 100.0, 1000.0]), [9.0, 90.0, 900.0])


def test_meanVector_kat():
    # src/ValuesOnTheGrid.jl:87-93
    np.testing.assert_array_equal(grids.meanVector([1.0, 10.0, 100.0, 1000.0]), [5.5, 55.0, 550.0])


def test_p_range_kat():
    # src/ValuesOnTheGrid.jl:113-126
    np.testing.assert_allclose(grids.p_range(-5.0, 4.0, 9), 10.0 ** np.arange(-5, 5), rtol=1e-15)


def test_u_range_kat():
    # src/ValuesOnTheGrid.jl:139-151
    np.testing.assert_array_equal(grids.u_range(-1.0, 1.0, 8),
                                  [-1.0, -0.75, -0.5, -0.25, 0.0, 0.25, 0.5, 0.75, 1.0])


def test_deltaEVector_kat():
    # src/ValuesOnTheGrid.jl:191-197 (printed to Float32 precision in the docstring)
    np.testing.assert_allclose(grids.deltaEVector(np.array([1.0, 10.0, 100.0, 1000.0]), 1.0),
                               [50.600693, 4951.15, 495001.16], rtol=2e-7)


def test_deltaEkinVector_kat():
    # src/ValuesOnTheGrid.jl:234-240 (full Float64 digits)
    np.testing.assert_allclose(grids.deltaEkinVector(np.array([1.0, 10.0, 100.0, 1000.0]), 1.0),
                               [46.10069600605712, 4906.1506753523645, 494551.15128635924],
                               rtol=1e-15)


def test_T00_weight_equals_dE_over_dp():
    """StressEnergyTensor's T^00 bin weight (src/DistributionMoments.jl:145) equals
    deltaEVector/dp (Grids.dE_list, src/StructsAndDictionaries.jl:259): two different formulas
    of the bin-averaged energy."""
    pr = grids.p_range(-3.0, 3.0, 32)
    for m in (1.0, 1836.15):
        w = grids.energy_weight_T00(pr, m)
        ref = grids.deltaEVector(pr, m) / grids.deltaVector(pr)
        np.testing.assert_allclose(w, ref, rtol=1e-12)
        # bin-mean energy lies between the energies at the bin edges
        E = np.sqrt(m * m + pr * pr)
        assert np.all(w >= E[:-1] * (1 - 1e-15)) and np.all(w <= E[1:] * (1 + 1e-15))
    np.testing.assert_array_equal(grids.energy_weight_T00(pr, 0.0), (pr[:-1] + pr[1:]) / 2)


def test_state_index_layout():
    """GlobalIndices_To_StateIndex (src/GlobalToStateIndices.jl:7-25): cell-major, species
    blocks, px fastest; the view of (x, species) covers exactly its block."""
    sp = [grids.Species("Ele", 1.0, -1.0, 3, 2, 2), grids.Species("Pho", 0.0, 0.0, 4, 1, 1)]
    ps = grids.PhaseSpace(sp, x_num=2, y_num=3, z_num=1)
    assert ps.n_momentum == 12 + 4 and ps.n == 16 * 6
    seen = set()
    for x in range(1, 3):
        for y in range(1, 4):
            for si, s in enumerate(sp, start=1):
                for pz in range(1, s.h_num + 1):
                    for py in range(1, s.u_num + 1):
                        for px in range(1, s.p_num + 1):
                            seen.add(ps.state_index(x, y, 1, px, py, pz, si))
    assert seen == set(range(1, ps.n + 1))
    assert ps.state_index(1, 1, 1, 2, 1, 1, 1) == 2
    assert ps.state_index(1, 1, 1, 1, 2, 1, 1) == 4
    assert ps.state_index(1, 2, 1, 1, 1, 1, 2) == 16 + 12 + 1
    state = np.arange(1, ps.n + 1, dtype=float)
    v = ps.view(state, 2, x_idx=2, y_idx=3)
    off = ((2 - 1) * 3 + 3 - 1) * 16 + 12
    np.testing.assert_array_equal(v, np.arange(off + 1, off + 5))
