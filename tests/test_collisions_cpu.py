"""Collision-matrix ingest on the host (no GPU): the oracle's restatement of Fill_M_Bin!
(oracle/fill_m_bin.py, src/Collisions/BuildBinaryMatrices.jl:81-284) and the name / scale rules
of LoadMatrices_Binary (dtx/collisions.py, src/Collisions/LoadMatrices.jl:6-146).

The restatement is pinned against an independent, vectorised statement of the collision term
the arrays define: for an interaction 1 + 2 -> 3 + 4 with (scaled) arrays G3, G4, L1, L2,

    df[a in 3] += sum_{b in 1, c in 2} Gbar3[a, b, c] f_b f_c      (likewise G4 into 4)
    df[b in 1] -= f_b sum_{c in 2} Lbar1[b, c] f_c                (likewise L2 for 2 against 1)

where Gbar is G without its under/overflow px3 bins (Ani), or its weighted average over phi
(Axi) or over u and phi (Iso) spread back over those bins, formed here with tensordot; the
contraction of the restated M_Bin (sum_{i,j} M[k,i,j] f_i f_j) must give the same df.  A
dropped half of the symmetrisation, a transposed array, an off-by-one px3 bin or a wrong
weight fails it.  Further pins: Iso of angle-independent arrays equals Ani; the sparse
assembly drops exactly the entries below 1e-16 of their column maximum; number conservation
when the losses are the gains summed over the outgoing states.
"""
import numpy as np
import pytest

import fill_m_bin as OC
from dtx import collisions as DC
from dtx.grids import PhaseSpace, Species

TOL = 1e-13


def phase_space(bin_norm=1.0):
    """Two species with different (px, py, pz) grids, so a transposed index cannot pass."""
    return PhaseSpace([Species("Ele", 1.0, -1.0, 3, 2, 2), Species("Pho", 0.0, 0.0, 2, 3, 1)],
                      x_num=1, Bin_Norm=bin_norm)


def grid(ps, s):
    sp = ps.species[s]
    return sp.p_num, sp.u_num, sp.h_num


def random_arrays(ps, locs, seed, zero_frac=0.1):
    rng = np.random.default_rng(seed)
    n1, n2, n3, n4 = locs
    d1, d2, d3, d4 = (grid(ps, s) for s in locs)

    def arr(shape):
        a = rng.lognormal(sigma=0.5, size=shape)
        a[rng.random(shape) < zero_frac] = 0.0
        return np.asfortranarray(a)

    G3 = arr((d3[0] + 2,) + d3[1:] + d1 + d2)
    G4 = arr((d4[0] + 2,) + d4[1:] + d1 + d2)
    L1 = arr(d1 + d2)
    L2 = arr(d2 + d1)
    return G3, G4, L1, L2


def widths(ps):
    dpy, dpz = DC.angle_widths(ps)
    out_y, out_z, oy, oz = [], [], 0, 0
    for s in ps.species:
        out_y.append(dpy[oy:oy + s.u_num])
        out_z.append(dpz[oz:oz + s.h_num])
        oy += s.u_num
        oz += s.h_num
    return out_y, out_z


def gbar(G, mode, dpy, dpz, dims):
    """(n3, n1, n2) effective gain of a 9-D array: Ani drops the px3 under/overflow bins;
    Axi / Iso replace every entry by the weighted angle average (tensordot over the averaged
    axes, then broadcast back)."""
    (d3, d1, d2), (y3, y1, y2), (z3, z1, z2) = dims, dpy, dpz
    G = np.asarray(G)[1:-1]
    if mode == OC.AXI:
        w = 1.0 / (z1.sum() * z2.sum() * z3.sum())
        A = np.tensordot(G, z3, axes=([2], [0]))        # (x3, y3, x1, y1, z1, x2, y2, z2)
        A = np.tensordot(A, z1, axes=([4], [0]))        # (x3, y3, x1, y1, x2, y2, z2)
        A = np.tensordot(A, z2, axes=([6], [0])) * w    # (x3, y3, x1, y1, x2, y2)
        G = np.broadcast_to(A[:, :, None, :, :, None, :, :, None], G.shape)
    elif mode == OC.ISO:
        w = 1.0 / (z1.sum() * z2.sum() * z3.sum() * y1.sum() * y2.sum() * y3.sum())
        A = np.einsum("abcdefghi,b,c,e,f,h,i->adg", G, y3, z3, y1, z1, y2, z2) * w
        G = np.broadcast_to(A[:, None, None, :, None, None, :, None, None], G.shape)
    n3, n1, n2 = (int(np.prod(d)) for d in (d3, d1, d2))
    return np.asarray(G).reshape(n3, n1, n2, order="F")


def lbar(L, mode, dpy, dpz, dims):
    (d1, d2), (y1, y2), (z1, z2) = dims, dpy, dpz
    L = np.asarray(L)
    if mode == OC.AXI:
        w = 1.0 / (z1.sum() * z2.sum())
        A = np.einsum("abcdef,c,f->abde", L, z1, z2) * w
        L = np.broadcast_to(A[:, :, None, :, :, None], L.shape)
    elif mode == OC.ISO:
        w = 1.0 / (z1.sum() * z2.sum() * y1.sum() * y2.sum())
        A = np.einsum("abcdef,b,c,e,f->ad", L, y1, z1, y2, z2) * w
        L = np.broadcast_to(A[:, None, None, :, None, None], L.shape)
    n1, n2 = (int(np.prod(d)) for d in (d1, d2))
    return np.asarray(L).reshape(n1, n2, order="F")


def direct_df(ps, interactions, mode, f):
    dpy, dpz = widths(ps)
    off = ps.species_offset
    size = [s.size for s in ps.species]
    df = np.zeros(ps.n_momentum)
    for locs, G3, G4, L1, L2, gs, ls in interactions:
        n1, n2, n3, n4 = locs
        f1 = f[off[n1]:off[n1] + size[n1]]
        f2 = f[off[n2]:off[n2] + size[n2]]
        for G, no in ((G3, n3), (G4, n4)):
            Gb = gbar(np.asarray(G) * gs, mode, (dpy[no], dpy[n1], dpy[n2]),
                      (dpz[no], dpz[n1], dpz[n2]), (grid(ps, no), grid(ps, n1), grid(ps, n2)))
            df[off[no]:off[no] + size[no]] += np.einsum("abc,b,c->a", Gb, f1, f2)
        for L, na, nb in ((L1, n1, n2), (L2, n2, n1)):
            Lb = lbar(np.asarray(L) * ls, mode, (dpy[na], dpy[nb]), (dpz[na], dpz[nb]),
                      (grid(ps, na), grid(ps, nb)))
            fa = f[off[na]:off[na] + size[na]]
            fb = f[off[nb]:off[nb] + size[nb]]
            df[off[na]:off[na] + size[na]] -= fa * (Lb @ fb)
    return df


def contract(M, f):
    N = f.size
    return np.einsum("kij,i,j->k", M.reshape(N, N, N, order="F"), f, f)


def fill(ps, interactions, mode, sparse=False):
    dpy, dpz = widths(ps)
    return OC.fill_m_bin(ps.n_momentum, ps.species_offset, dpy, dpz, interactions, mode, sparse)


INTERACTIONS = {"EleEle_EleEle": (0, 0, 0, 0), "ElePho_ElePho": (0, 1, 0, 1),
                "PhoPho_EleEle": (1, 1, 0, 0), "PhoEle_PhoEle": (1, 0, 1, 0),
                "ElePho_PhoPho": (0, 1, 1, 1)}


@pytest.mark.parametrize("mode", [OC.ANI, OC.AXI, OC.ISO], ids=["Ani", "Axi", "Iso"])
@pytest.mark.parametrize("name", list(INTERACTIONS))
def test_restated_fill_equals_direct_collision_term(name, mode):
    ps = phase_space()
    locs = INTERACTIONS[name]
    G3, G4, L1, L2 = random_arrays(ps, locs, seed=sum(locs) + 3 * mode)
    its = [(locs, G3, G4, L1, L2, 0.7, 1.3)]
    M = fill(ps, its, mode)
    f = np.random.default_rng(5).lognormal(size=ps.n_momentum)
    got, want = contract(M, f), direct_df(ps, its, mode, f)
    assert np.max(np.abs(got - want)) <= TOL * np.max(np.abs(want))


@pytest.mark.parametrize("mode", [OC.ANI, OC.AXI, OC.ISO], ids=["Ani", "Axi", "Iso"])
def test_restated_fill_of_an_interaction_list(mode):
    """Several interactions accumulate into one M_Bin (LoadMatrices_Binary's loop)."""
    ps = phase_space()
    its = []
    for q, locs in enumerate(INTERACTIONS.values()):
        its.append((locs, *random_arrays(ps, locs, seed=11 + q), 1.0 + q, 0.5 + q))
    M = fill(ps, its, mode)
    f = np.random.default_rng(9).lognormal(size=ps.n_momentum)
    got, want = contract(M, f), direct_df(ps, its, mode, f)
    assert np.max(np.abs(got - want)) <= TOL * np.max(np.abs(want))


def test_iso_of_angle_independent_arrays_equals_ani():
    """Averaging an array that does not depend on u or phi returns it: Iso == Axi == Ani."""
    ps = phase_space()
    locs = (0, 1, 0, 1)
    G3, G4, L1, L2 = random_arrays(ps, locs, seed=4, zero_frac=0.0)
    # make every array constant along its py and pz axes
    G3 = np.asfortranarray(np.broadcast_to(G3[:, :1, :1, :, :1, :1, :, :1, :1], G3.shape))
    G4 = np.asfortranarray(np.broadcast_to(G4[:, :1, :1, :, :1, :1, :, :1, :1], G4.shape))
    L1 = np.asfortranarray(np.broadcast_to(L1[:, :1, :1, :, :1, :1], L1.shape))
    L2 = np.asfortranarray(np.broadcast_to(L2[:, :1, :1, :, :1, :1], L2.shape))
    its = [(locs, G3, G4, L1, L2, 1.0, 1.0)]
    Mani = fill(ps, its, OC.ANI)
    for mode in (OC.AXI, OC.ISO):
        M = fill(ps, its, mode)
        assert np.max(np.abs(M - Mani)) <= 1e-14 * np.max(np.abs(Mani)), mode


def test_sparse_assembly_drops_small_ani_gains():
    """Bin_sparse: an Ani gain entry below 1e-16 of its column's maximum (over every px3 bin,
    under/overflow included) is not pushed (BuildBinaryMatrices.jl:163-173); everything else
    equals the dense fill."""
    ps = phase_space()
    locs = (0, 1, 0, 1)
    G3, G4, L1, L2 = random_arrays(ps, locs, seed=8, zero_frac=0.0)
    G3[2, 0, 0, 1, 0, 0, 1, 0, 0] = 1e-18        # dropped: below 1e-16 x the column max
    G3[3, 1, 0, 0, 0, 0, 0, 0, 0] = 1e-15        # kept
    G4[0, 0, 0, 0, 0, 0, 0, 0, 0] = 1e30         # underflow bin: never used, but sets the max
    its = [(locs, G3, G4, L1, L2, 1.0, 1.0)]
    Ms = fill(ps, its, OC.ANI, sparse=True)
    G3z, G4z = G3.copy(), G4.copy()
    G3z[2, 0, 0, 1, 0, 0, 1, 0, 0] = 0.0
    G4z[1:-1, 0, 0, 0, 0, 0, 0, 0, 0] = 0.0        # the whole column is below 1e30 x 1e-16
    Md = fill(ps, [(locs, G3z, G4z, L1, L2, 1.0, 1.0)], OC.ANI)
    np.testing.assert_array_equal(Ms, Md)
    assert not np.array_equal(Ms, fill(ps, its, OC.ANI))


def test_number_conservation_with_consistent_losses():
    """Each collision 1 + 2 -> 3 + 4 removes one particle of 1 and of 2 and adds one of 3 and
    of 4: with L1[b, c] = sum_a G3[a; b, c] and L2[c, b] = sum_a G4[a; b, c] the summed df
    vanishes."""
    ps = phase_space()
    for locs in INTERACTIONS.values():
        G3, G4, _, _ = random_arrays(ps, locs, seed=21, zero_frac=0.0)
        n1, n2, n3, n4 = locs
        d1, d2 = grid(ps, n1), grid(ps, n2)
        s1, s2 = int(np.prod(d1)), int(np.prod(d2))
        g3 = G3[1:-1].reshape(-1, s1, s2, order="F").sum(axis=0)
        g4 = G4[1:-1].reshape(-1, s1, s2, order="F").sum(axis=0)
        L1 = np.asfortranarray(g3.reshape(d1 + d2, order="F"))
        L2 = np.asfortranarray(g4.T.reshape(d2 + d1, order="F"))
        M = fill(ps, [(locs, G3, G4, L1, L2, 1.0, 1.0)], OC.ANI)
        f = np.random.default_rng(2).lognormal(size=ps.n_momentum)
        df = contract(M, f)
        assert abs(df.sum()) <= 1e-13 * np.abs(df).sum(), locs


# ---------------------------------------------------------------- LoadMatrices_Binary rules --

def _provider(ps, seed=0):
    calls = []

    def arrays(names, corrected):
        calls.append((names, corrected))
        locs = tuple(ps.name_list.index(n if n in ps.name_list else "Ele") for n in names)
        return random_arrays(ps, locs, seed)
    return arrays, calls


def test_load_matrices_locations_and_bin_norm():
    ps = phase_space(bin_norm=2.5)
    arrays, calls = _provider(ps)
    its = DC.LoadMatrices_Binary([DC.BinaryStruct("Ele", "Pho", "Ele", "Pho"),
                                  DC.BinaryStruct("Pho", "Pho", "Ele", "Ele")], arrays, ps,
                                 DC.Axi(), corrected=False)
    assert [it.name_locs for it in its] == [(0, 1, 0, 1), (1, 1, 0, 0)]
    assert all(it.gain_scale == 2.5 and it.loss_scale == 2.5 for it in its)
    assert calls == [(("Ele", "Pho", "Ele", "Pho"), False), (("Pho", "Pho", "Ele", "Ele"), False)]


def test_load_matrices_merged_positrons():
    """"Pos" not a species: Ele stands for Ele + Pos (LoadMatrices.jl:47-70)."""
    ps = phase_space(bin_norm=2.0)
    arrays, calls = _provider(ps)
    its = DC.LoadMatrices_Binary([DC.BinaryStruct("Ele", "Pos", "Pho", "Pho"),
                                  DC.BinaryStruct("Pho", "Pho", "Ele", "Pos")], arrays, ps)
    ann, pair = its
    assert ann.name_locs == (0, 0, 1, 1) and ann.gain_scale == 0.5 and ann.loss_scale == 0.5
    assert pair.name_locs == (1, 1, 0, 0) and pair.gain_scale == 2.0
    # Pos Compton ends the list: what follows it is not loaded (the reference returns)
    arrays, calls = _provider(ps)
    its = DC.LoadMatrices_Binary([DC.BinaryStruct("Ele", "Pho", "Ele", "Pho"),
                                  DC.BinaryStruct("Pos", "Pho", "Pos", "Pho"),
                                  DC.BinaryStruct("Pho", "Pho", "Ele", "Ele")], arrays, ps)
    assert [it.name_locs for it in its] == [(0, 1, 0, 1)]
    assert len(calls) == 1
    with pytest.raises(ValueError, match="no Ele/Pos rule"):
        DC.LoadMatrices_Binary([DC.BinaryStruct("Pos", "Pos", "Pho", "Pho")], arrays, ps)


def test_load_matrices_positron_species_and_compton_swap():
    """With Pos a species of its own, Pos Compton keeps its locations but reads the Ele Compton
    arrays (LoadMatrices.jl:73-76)."""
    ps = PhaseSpace([Species("Ele", 1.0, -1.0, 2, 2, 1), Species("Pho", 0.0, 0.0, 2, 2, 1),
                     Species("Pos", 1.0, 1.0, 2, 2, 1)])
    seen = []

    def arrays(names, corrected):
        seen.append(names)
        return random_arrays(ps, (0, 1, 0, 1), 0)
    its = DC.LoadMatrices_Binary([DC.BinaryStruct("Pos", "Pho", "Pos", "Pho")], arrays, ps)
    assert its[0].name_locs == (2, 1, 2, 1)
    assert seen == [("Ele", "Pho", "Ele", "Pho")]


def test_build_binary_matrices_arguments():
    ps = phase_space()
    with pytest.raises(ValueError, match="requires Bin_corrected"):
        DC.BuildBinaryMatrices(ps, [DC.BinaryStruct("Ele", "Ele", "Ele", "Ele")], None, ".",
                               Bin_sparse=True, Bin_corrected=False)
    with pytest.raises(NotImplementedError, match="JLD2"):
        DC.BuildBinaryMatrices(ps, [DC.BinaryStruct("Ele", "Ele", "Ele", "Ele")], None, ".")
    empty = DC.BuildBinaryMatrices(ps, [], None, ".")
    assert empty.Binary_list == [] and not empty.interactions
    arrays, _ = _provider(ps)
    bad = lambda names, corrected: random_arrays(ps, (0, 0, 0, 0), 0)  # noqa: E731
    with pytest.raises(ValueError, match="GainMatrix3 has shape"):
        DC.BuildBinaryMatrices(ps, [DC.BinaryStruct("Ele", "Pho", "Ele", "Pho")], None,
                               arrays=bad)
    ok = DC.BuildBinaryMatrices(ps, [DC.BinaryStruct("Ele", "Pho", "Ele", "Pho")], [0], ".",
                                Bin_Mode=DC.Iso(), Bin_sparse=True, arrays=arrays)
    assert ok.Bin_Mode == DC.Iso() and ok.Bin_sparse and ok.Domain == [0]
    assert len(ok.interactions) == 1
