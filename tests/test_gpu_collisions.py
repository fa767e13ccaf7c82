"""Collision-matrix ingest through the C ABI (dtx_add_binary_interaction): the operator assembled
from DiplodocusCollisions-shaped arrays on the handle steps the state as the oracle does with the
M_Bin that the restated Fill_M_Bin! builds (oracle/fill_m_bin.py, pinned on the CPU against an
independent statement of the collision term), at 1e-12 relative (L-inf) over the steps, and as
the dense upload of that same M_Bin (dtx_set_binary_dense) does, at 1e-14: the two differ only
in the order of the few additions that meet in one operator entry.

Cases: Ani, Axi and Iso; the sparse (Bin_sparse) assembly; an interaction list over two species
with different (px, py, pz) grids, Bin_Norm and the Ele/Pos rules; both stepping methods; and the
refusals of bad descriptors.
"""
import ctypes as C

import numpy as np
import pytest

import fill_m_bin as OC
import oracle as O
from dtx import _lib, synth
from dtx import collisions as DC
from dtx.api import (BinaryMatricesStruct, EmissionMatricesStruct, FluxMatricesStruct,
                     ForwardEulerStruct, ForwardSymplecticEulerStruct, Solve)
from dtx.grids import PhaseSpace, Species
from dtx.workloads import _exact_dt

pytestmark = pytest.mark.gpu
TOL_ORACLE = 1e-12
TOL_DENSE = 1e-14

LISTS = {
    "compton": [("Ele", "Pho", "Ele", "Pho")],
    "like": [("Ele", "Ele", "Ele", "Ele")],
    "mixed": [("Ele", "Ele", "Ele", "Ele"), ("Ele", "Pho", "Ele", "Pho"),
              ("Pho", "Pho", "Ele", "Pos"), ("Ele", "Pos", "Pho", "Pho")],
}


def rel_linf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def phase_space(x_num=3):
    return PhaseSpace([Species("Ele", 1.0, -1.0, 4, 2, 2), Species("Pho", 0.0, 0.0, 3, 3, 1)],
                      x_num=x_num, Bin_Norm=0.75)


def provider(ps, seed):
    """Arrays of the shapes DiplodocusCollisions writes, per looked-up names: random gains and
    the losses they imply (L1[b, c] = sum_a G3[a; b, c], L2[c, b] = sum_a G4[a; b, c], the
    under/overflow bins included), so a collision conserves number and f stays bounded."""
    def arrays(names, corrected):
        locs = [ps.name_list.index(n if n in ps.name_list else "Ele") for n in names]
        rng = np.random.default_rng([seed] + locs)
        d = [(ps.species[s].p_num, ps.species[s].u_num, ps.species[s].h_num) for s in locs]

        def arr(shape):
            a = rng.lognormal(sigma=0.7, size=shape)
            a[rng.random(shape) < 0.1] = 0.0
            return np.asfortranarray(a)
        G3 = arr((d[2][0] + 2,) + d[2][1:] + d[0] + d[1])
        G4 = arr((d[3][0] + 2,) + d[3][1:] + d[0] + d[1])
        n1, n2 = int(np.prod(d[0])), int(np.prod(d[1]))
        g3 = G3.reshape(-1, n1, n2, order="F").sum(axis=0)
        g4 = G4.reshape(-1, n1, n2, order="F").sum(axis=0)
        L1 = np.asfortranarray(g3.reshape(d[0] + d[1], order="F"))
        L2 = np.asfortranarray(g4.T.reshape(d[1] + d[0], order="F"))
        # 1e30 in the underflow px3 bin of 5% of the columns (after the losses): never an
        # entry, but the column maximum of the sparse assembly, which then drops the column
        for G in (G3, G4):
            G[0][rng.random(G.shape[1:]) < 0.05] = 1e30
        return G3, G4, L1, L2
    return arrays


def restated_m_bin(ps, interactions, mode, sparse):
    dpy, dpz = DC.angle_widths(ps)
    py, pz, oy, oz = [], [], 0, 0
    for s in ps.species:
        py.append(dpy[oy:oy + s.u_num])
        pz.append(dpz[oz:oz + s.h_num])
        oy += s.u_num
        oz += s.h_num
    its = [(it.name_locs, it.GainMatrix3, it.GainMatrix4, it.LossMatrix1, it.LossMatrix2,
            it.gain_scale, it.loss_scale) for it in interactions]
    M = OC.fill_m_bin(ps.n_momentum, ps.species_offset, py, pz, its, mode.code, sparse)
    N = ps.n_momentum
    return M.reshape(N, N, N, order="F")


def stable_dt(ps, M, f0, Ap):
    """0.1 / the fastest loss rate sum_j |M[k,k,j]| f_j + sum_i |M[k,i,k]| f_i of any cell."""
    N, worst = ps.n_momentum, 0.0
    k = np.arange(N)
    A = np.abs(M)
    D = A[k, k, :] + A[k, :, k]  # (k, other)
    for c in range(ps.n_space):
        f = f0[c * N:(c + 1) * N]
        worst = max(worst, float(np.max(D @ f / Ap[c * N:(c + 1) * N])))
    return _exact_dt(0.1 / worst)


def build(ps, method, BinM, f0, X, P, Ap, Vol):
    cls = ForwardEulerStruct if method == "euler" else ForwardSymplecticEulerStruct
    return cls(ps, f0, np.zeros(ps.n), BinM, EmissionMatricesStruct(),
               FluxMatricesStruct(Ap, X, P, Vol))


@pytest.mark.parametrize("method", ["euler", "symplectic"])
@pytest.mark.parametrize("mode,sparse", [(DC.Ani(), False), (DC.Ani(), True), (DC.Axi(), False),
                                         (DC.Iso(), False)], ids=["Ani", "Ani_sparse", "Axi",
                                                                  "Iso"])
@pytest.mark.parametrize("which", list(LISTS))
def test_ingested_operator_against_oracle_and_dense(which, mode, sparse, method):
    ps = phase_space()
    blist = [DC.BinaryStruct(*n) for n in LISTS[which]]
    BinM = DC.BuildBinaryMatrices(ps, blist, None, arrays=provider(ps, 3), Bin_Mode=mode,
                                  Bin_sparse=sparse)
    M = restated_m_bin(ps, BinM.interactions, mode, sparse)
    f0 = synth.initial_state(ps, theta=1.0, seed=4, noise=0.05)
    Ap, Vol = synth.ap_vol(ps)
    X, P = synth.flux_matrices(ps, sync_K=0.05, axes="x", periodic=True)
    dt = stable_dt(ps, M, f0, Ap)
    steps = 12
    ts = np.array([0.0, 6 * dt, steps * dt])
    print(f"{which} {mode} sparse={sparse} {method}: N {ps.n_momentum}, dt {dt:.3e}",
          flush=True)
    ing = build(ps, method, BinM, f0, X, P, Ap, Vol)
    dense = build(ps, method, BinaryMatricesStruct(M_Bin=M, Binary_list=["dense"]), f0, X, P,
                  Ap, Vol)
    out_i = Solve(ing, dt, ts, Verbose=0)
    out_d = Solve(dense, dt, ts, Verbose=0)
    F = None
    if method == "euler":
        F = synth.csc_add(ps.n, X, P)
    prob = O.OracleProblem(ps, Ap, Vol, np.zeros(ps.n), F=F,
                           X=X if method != "euler" else None,
                           P=P if method != "euler" else None, M_Bin=M)
    fo, *_ = prob.solve(f0, dt, ts, verbose=0, method=0 if method == "euler" else 1)
    for i in range(len(ts)):
        assert np.all(np.isfinite(out_i.f[i]))
        assert rel_linf(out_i.f[i], fo[i]) <= TOL_ORACLE, i
        assert rel_linf(out_i.f[i], out_d.f[i]) <= TOL_DENSE, i
    assert rel_linf(out_i.f[-1], f0) > 1e-6  # the collisions did change the state
    ing.close()
    dense.close()


def test_sparse_assembly_drops_columns_on_the_handle():
    """Bin_sparse (Ani): columns whose maximum (the planted underflow bin) exceeds their entries
    by more than 1e16 are dropped by the library as by the restated sparse fill, and that changes
    the state (so the drop is seen)."""
    ps = phase_space()
    blist = [DC.BinaryStruct("Ele", "Pho", "Ele", "Pho")]
    f0 = synth.initial_state(ps, theta=1.0, seed=4, noise=0.05)
    Ap, Vol = synth.ap_vol(ps)
    out = {}
    for sparse in (False, True):
        BinM = DC.BuildBinaryMatrices(ps, blist, None, arrays=provider(ps, 3), Bin_sparse=sparse)
        M = restated_m_bin(ps, BinM.interactions, DC.Ani(), sparse)
        dt = stable_dt(ps, restated_m_bin(ps, BinM.interactions, DC.Ani(), False), f0, Ap)
        ts = np.array([0.0, 8 * dt])
        m = build(ps, "euler", BinM, f0, None, None, Ap, Vol)
        out[sparse] = Solve(m, dt, ts, Verbose=0).f[-1]
        fo, *_ = O.OracleProblem(ps, Ap, Vol, np.zeros(ps.n), M_Bin=M).solve(f0, dt, ts,
                                                                            verbose=0)
        assert rel_linf(out[sparse], fo[-1]) <= TOL_ORACLE, sparse
        m.close()
    assert rel_linf(out[True], out[False]) > 1e-4


def test_full_size_ingest_against_direct_collision_term():
    """transport_1d's momentum grid (2 species x 256 states), three interactions, 0.86 GB of
    arrays (workloads.collision_ingest): one Forward Euler step of the ingested operator (one
    cell, no flux) moves f by (dt / dt0) Vol / Ap * df, df the vectorised collision term of the
    arrays (test_collisions_cpu.direct_df), to 1e-12 of the largest move, wherever f stays clear
    of the 1f-10 clamp.  (The arrays conserve number: sum Ap * move vanishes to 1e-12.)"""
    from dtx import workloads
    from test_collisions_cpu import direct_df
    ps, blist, arrays, f0, Ap, Vol = workloads.collision_ingest()
    BinM = DC.BuildBinaryMatrices(ps, blist, None, arrays=arrays)
    its = [(it.name_locs, it.GainMatrix3, it.GainMatrix4, it.LossMatrix1, it.LossMatrix2,
            it.gain_scale, it.loss_scale) for it in BinM.interactions]
    rate = direct_df(ps, its, OC.ANI, f0) * Vol[0] / Ap
    dt_scale = 0.05 * np.max(f0) / np.max(np.abs(rate))  # the largest move is 5% of max f
    dt = _exact_dt(dt_scale * ps.dt0)
    move = rate * (dt / ps.dt0)
    print(f"full-size ingest: N {ps.n_momentum}, dt {dt:.3e}", flush=True)
    m = build(ps, "euler", BinM, f0, None, None, Ap, Vol)
    m(0.0, 1e300, dt, 0)
    f1 = m.f
    m.close()
    keep = (f0 >= 2e-10) & (f0 + move >= 2e-10)
    assert keep.sum() > 300
    err = np.max(np.abs((f1 - f0)[keep] - move[keep]))
    assert err <= 1e-12 * np.max(np.abs(move)), err
    assert abs(Ap @ move) <= 1e-12 * (np.abs(Ap) @ np.abs(move))


def test_interaction_descriptor_refusals():
    ps = phase_space(x_num=1)
    BinM = DC.BuildBinaryMatrices(ps, [DC.BinaryStruct("Ele", "Pho", "Ele", "Pho")], None,
                                  arrays=provider(ps, 1))
    f0 = synth.initial_state(ps, theta=1.0, seed=2, noise=0.0)
    Ap, Vol = synth.ap_vol(ps)
    m = build(ps, "euler", BinM, f0, None, None, Ap, Vol)
    it = BinM.interactions[0]
    arrs = [np.asfortranarray(a).reshape(-1, order="F")
            for a in (it.GainMatrix3, it.GainMatrix4, it.LossMatrix1, it.LossMatrix2)]
    dpy, dpz = DC.angle_widths(ps)

    def call(**kw):
        x = _lib.BinaryInteraction()
        x.species[:] = kw.get("species", [0, 1, 0, 1])
        x.mode = kw.get("mode", 0)
        x.sparse = 0
        x.gain_scale = x.loss_scale = 1.0
        ptrs = [_lib.dptr(a) for a in arrs]
        if kw.get("null"):
            ptrs[2] = C.cast(None, _lib.PD)
        x.gain3, x.gain4, x.loss1, x.loss2 = ptrs
        x.dpy, x.dpz = _lib.dptr(dpy), _lib.dptr(dpz)
        return m.lib.dtx_add_binary_interaction(m._h, C.byref(x))

    for kw, msg in ((dict(species=[0, 2, 0, 1]), b"species index out of range"),
                    (dict(mode=3), b"unknown mode"), (dict(null=True), b"null array")):
        assert call(**kw) == -1  # DTX_ERR_INVALID
        assert msg in m.lib.dtx_last_error(m._h)
    assert call() == 0  # a valid descriptor still goes through afterwards
    m.close()
