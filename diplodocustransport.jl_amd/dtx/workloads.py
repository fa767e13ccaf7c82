"""The benchmark workloads as concrete, seeded inputs.

Each builder returns a Workload: phase space, initial and injection vectors, flux matrices,
Ap/Vol, the binary operator (host channels, monolithic tensor or a device-synthesis seed), the
time step and the step count.

  homogeneous_box  -- 2 species (electron, photon), 32 log p bins on [1e-3, 1e3] x 8 u x 1 phi,
                      one cell, dense binary tensor, 1000 steps (CPU-runnable case).
  transport_1d     -- the same momentum grid on 128 periodic x cells on [0,1], x advection plus
                      synchrotron cooling (cooling time ~100 steps), collisions per cell,
                      2000 steps.
  precision_box    -- 3 species x (64 p x 8 u x 2 phi) = 1024 states each, one cell, device-
                      generated gain tensors, 1000 steps.
  transport_2d     -- 32 x 32 periodic x-y cells, 2 species x (32 p x 8 u x 2 phi), advection in
                      x and y, cooling in p, device-generated gain tensors shared by all cells,
                      500 steps.
The time step is 0.1 / (fastest loss rate of the initial state), rounded down to 12 significant
bits so that save times are exact multiples of it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import synth
from .grids import PhaseSpace, Species

E_HALF = math.exp(0.5)  # mean of lognormal(0, 1)


@dataclass
class Workload:
    name: str
    ps: PhaseSpace
    f0: np.ndarray
    inj: np.ndarray
    Ap: np.ndarray
    Vol: np.ndarray
    X: Optional[tuple]
    P: Optional[tuple]
    dt: float
    steps: int
    seed: int
    gains: list = field(default_factory=list)
    losses: list = field(default_factory=list)
    M_Bin: Optional[np.ndarray] = None
    synth_seed: Optional[int] = None
    emission_blocks: Optional[np.ndarray] = None
    domain: Optional[list] = None
    flux_spec: Optional[dict] = None  # synth.flux_matrices arguments of X and P, for the device

    @property
    def F(self):
        n = self.ps.n
        if self.X is None and self.P is None:
            return None
        if self.X is None:
            return self.P
        if self.P is None:
            return self.X
        return synth.csc_add(n, self.X, self.P)

    def t_save(self, n_save: int = 2):
        """Save times after 0, steps/(n_save-1), ..., steps steps of size dt, accumulated as
        Solve accumulates t_start (src/Solver.jl:36).  With the workloads' 12-bit dt every sum
        is exact: each interval ends on a step and the clipped dt equals dt."""
        marks = np.linspace(0, self.steps, n_save).round().astype(int)
        out, t, k = [0.0], 0.0, 0
        for m in marks[1:]:
            while k < m:
                t += self.dt
                k += 1
            out.append(t)
        return np.array(out)


def species_pair(p_num=32, u_num=8, h_num=1, p_low=-3.0, p_up=3.0):
    return [Species("Ele", 1.0, -1.0, p_num, u_num, h_num, p_low, p_up),
            Species("Pho", 0.0, 0.0, p_num, u_num, h_num, p_low, p_up)]


def _exact_dt(dt: float) -> float:
    """dt rounded down to 12 significant bits.  Every save time k*dt (k < 2^41) and every
    Solve accumulation t_start + dt is then exact, so each save interval ends exactly on a step
    and the clipped dt that Solve carries into the next interval (src/Solver.jl:35) equals dt;
    with an unrounded dt the carried value drifts by an ulp per interval until an interval ends
    a rounding error after a step."""
    m, e = math.frexp(dt)
    return math.ldexp(math.floor(m * 4096.0) / 4096.0, e)


def _dt_from_losses(ps, losses, f0, Ap, Vol):
    N = ps.n_momentum
    worst = 0.0
    for c in range(ps.n_space):
        worst = max(worst, synth.max_loss_rate(ps, losses, f0[c * N:(c + 1) * N], Vol[c],
                                               Ap[c * N]))
    return _exact_dt(0.1 / worst)


def _dt_expected(ps, f0, Ap, Vol):
    """0.1 / fastest loss rate with the expected loss matrix L[i,j] = No * e^(1/2)."""
    N = ps.n_momentum
    off = ps.species_offset
    worst = 0.0
    for c in range(ps.n_space):
        fc = f0[c * N:(c + 1) * N]
        tot = [float(np.sum(fc[o:o + s.size])) for o, s in zip(off, ps.species)]
        for a, sa in enumerate(ps.species):
            rate = 0.0
            for (x, y) in synth.interactions(ps.num_species):
                if a == x:
                    rate += sa.size * E_HALF * tot[y] if x != y else sa.size * E_HALF * tot[x]
                elif a == y:
                    rate += sa.size * E_HALF * tot[x]
            worst = max(worst, rate * Vol[c] / Ap[c * N])
    return _exact_dt(0.1 / worst)


def _sync_K(ps, dt, cooling_steps=100.0):
    _, P1 = synth.flux_matrices(ps, sync_K=1.0, axes="")
    c1 = synth.sync_outflow_coefficient(ps, P1)
    return 1.0 / (cooling_steps * dt * c1) if c1 > 0 else 0.0


def homogeneous_box(p_num=32, u_num=8, h_num=1, steps=1000, seed=None, monolithic=True):
    seed = synth.config_seed(0) if seed is None else seed
    ps = PhaseSpace(species_pair(p_num, u_num, h_num))
    f0 = synth.initial_state(ps, theta=1.0, seed=seed)
    Ap, Vol = synth.ap_vol(ps)
    gains, losses = synth.binary_channels(ps, seed)
    dt = _dt_from_losses(ps, losses, f0, Ap, Vol)
    M = synth.assemble_M_Bin(ps, gains, losses) if monolithic else None
    return Workload("homogeneous_box", ps, f0, np.zeros(ps.n), Ap, Vol, None, None, dt, steps,
                    seed, gains, losses, M)


def emission_blocks(ps, seed, dt, Ap=None):
    """Dense per-cell emission blocks [cell][column j][row i] (each block column-major), entries
    uniform in (-0.1, 0.1) / (dt N), and with Ap given each row i times Ap_i, so that the term
    changes f_i by at most a tenth of the mean f per step whatever the cell volume."""
    N = ps.n_momentum
    key = synth.stream_key(seed, 77)
    raw = synth.uniform(np.arange(ps.n_space * N * N, dtype=np.uint64), key)
    blocks = (raw.reshape(ps.n_space, N, N) - 0.5) * (0.2 / dt / N)   # [cell][row][column]
    if Ap is not None:
        blocks *= np.asarray(Ap, dtype=np.float64).reshape(ps.n_space, N, 1)
    return np.ascontiguousarray(blocks.transpose(0, 2, 1))


def transport_1d(p_num=32, u_num=8, h_num=1, x_num=128, steps=2000, seed=None,
                 monolithic=False, with_sync=True, emission=False):
    """emission: dense per-cell emission blocks as well (emission_blocks)."""
    seed = synth.config_seed(1) if seed is None else seed
    ps = PhaseSpace(species_pair(p_num, u_num, h_num), x_num=x_num)
    f0 = synth.initial_state(ps, theta=1.0, seed=seed)
    Ap, Vol = synth.ap_vol(ps)
    gains, losses = synth.binary_channels(ps, seed)
    dt = _dt_from_losses(ps, losses, f0, Ap, Vol)
    K = _sync_K(ps, dt) if with_sync else 0.0
    X, P = synth.flux_matrices(ps, sync_K=K, axes="x")
    M = synth.assemble_M_Bin(ps, gains, losses) if monolithic else None
    blocks = emission_blocks(ps, seed, dt, Ap) if emission else None
    return Workload("transport_1d", ps, f0, np.zeros(ps.n), Ap, Vol, X, P, dt, steps, seed,
                    gains, losses, M, emission_blocks=blocks,
                    flux_spec=dict(sync_K=K, axes="x", periodic=True))


# Interactions of the collision-ingest case: like-particle, Compton-like and photon pair
# production, over the two species of transport_1d.
INGEST_INTERACTIONS = (("Ele", "Ele", "Ele", "Ele"), ("Ele", "Pho", "Ele", "Pho"),
                       ("Pho", "Pho", "Ele", "Ele"))


def collision_ingest(p_num=32, u_num=8, h_num=1, seed=None, theta=30.0):
    """Collision arrays at transport_1d's momentum grid (2 species x p_num u_num h_num states),
    one cell: (PhaseSpace, Binary_list, arrays provider for BuildBinaryMatrices, f0, Ap, Vol).
    The arrays (synth.collision_arrays) conserve number; theta keeps f above the clamp."""
    from .collisions import BinaryStruct
    seed = synth.config_seed(5) if seed is None else seed
    ps = PhaseSpace(species_pair(p_num, u_num, h_num), x_num=1)
    f0 = synth.initial_state(ps, theta=theta, seed=seed, noise=0.0)
    Ap, Vol = synth.ap_vol(ps)
    blist = [BinaryStruct(*n) for n in INGEST_INTERACTIONS]

    def arrays(names, corrected):
        return synth.collision_arrays(ps, tuple(ps.name_list.index(n) for n in names), seed)
    return ps, blist, arrays, f0, Ap, Vol


def precision_box(p_num=64, u_num=8, h_num=2, num_species=3, steps=1000, seed=None,
                  emission=False):
    """emission: a dense emission block as well (emission_blocks)."""
    seed = synth.config_seed(2) if seed is None else seed
    sp = [Species("Ele", 1.0, -1.0, p_num, u_num, h_num),
          Species("Pos", 1.0, 1.0, p_num, u_num, h_num),
          Species("Pho", 0.0, 0.0, p_num, u_num, h_num)][:num_species]
    ps = PhaseSpace(sp)
    f0 = synth.initial_state(ps, theta=1.0, seed=seed)
    Ap, Vol = synth.ap_vol(ps)
    dt = _dt_expected(ps, f0, Ap, Vol)
    blocks = emission_blocks(ps, seed, dt, Ap) if emission else None
    return Workload("precision_box", ps, f0, np.zeros(ps.n), Ap, Vol, None, None, dt, steps,
                    seed, emission_blocks=blocks, synth_seed=seed)


def transport_2d(p_num=32, u_num=8, h_num=2, x_num=32, y_num=32, steps=500, seed=None):
    seed = synth.config_seed(4) if seed is None else seed
    ps = PhaseSpace(species_pair(p_num, u_num, h_num), x_num=x_num, y_num=y_num)
    f0 = synth.initial_state(ps, theta=1.0, seed=seed)
    Ap, Vol = synth.ap_vol(ps)
    dt = _dt_expected(ps, f0, Ap, Vol)
    K = _sync_K(ps, dt)
    X, P = synth.flux_matrices(ps, sync_K=K, axes="xy")
    return Workload("transport_2d", ps, f0, np.zeros(ps.n), Ap, Vol, X, P, dt, steps, seed,
                    synth_seed=seed, flux_spec=dict(sync_K=K, axes="xy", periodic=True))


def tiny(p_num=4, u_num=2, h_num=1, x_num=3, seed=7, steps=20, monolithic=True,
         with_flux=True, emission=False, injection=False, domain=None, num_species=2):
    """Small case for parity tests (N = 2*p*u*h states per cell)."""
    sp = species_pair(p_num, u_num, h_num)
    if num_species == 1:
        sp = sp[:1]
    elif num_species == 3:
        sp = sp + [Species("Pos", 1.0, 1.0, p_num, u_num, h_num)]
        sp[1], sp[2] = sp[2], sp[1]
    ps = PhaseSpace(sp, x_num=x_num)
    f0 = synth.initial_state(ps, theta=1.0, seed=seed, noise=0.05)
    Ap, Vol = synth.ap_vol(ps)
    gains, losses = synth.binary_channels(ps, seed)
    dt = _dt_from_losses(ps, losses, f0, Ap, Vol)
    X = P = spec = None
    if with_flux and x_num > 1:
        K = _sync_K(ps, dt, cooling_steps=20.0)
        spec = dict(sync_K=K, axes="x", periodic=True)
        X, P = synth.flux_matrices(ps, **spec)
    M = synth.assemble_M_Bin(ps, gains, losses) if monolithic else None
    inj = np.zeros(ps.n)
    if injection:
        inj = 0.01 * synth.initial_state(ps, theta=0.5, seed=seed + 1, noise=0.0)
    blocks = emission_blocks(ps, seed, dt) if emission else None
    return Workload("tiny", ps, f0, inj, Ap, Vol, X, P, dt, steps, seed, gains, losses, M,
                    emission_blocks=blocks, domain=domain, flux_spec=spec)


def emission_dense(w: Workload) -> Optional[np.ndarray]:
    """The per-cell emission blocks as the reference's dense global n x n M_Emi."""
    if w.emission_blocks is None:
        return None
    n, N = w.ps.n, w.ps.n_momentum
    M = np.zeros((n, n), order="F")
    for c in range(w.ps.n_space):
        # blocks are stored column-major: element (i, j) at [c, j, i] of the C-order array
        M[c * N:(c + 1) * N, c * N:(c + 1) * N] = w.emission_blocks[c].T
    return M


def binary_csc(w: Workload):
    """The monolithic M_Bin as the reference's sparse form (Bin_sparse): the (N^2 x N) matrix as
    a 1-based (colptr, rowval, nzval) CSC of its non-zero entries."""
    import scipy.sparse as sp
    N = w.ps.n_momentum
    A = sp.csc_matrix(np.asarray(w.M_Bin).reshape(N * N, N, order="F"))
    A.sort_indices()
    return (A.indptr.astype(np.int64) + 1, A.indices.astype(np.int64) + 1, A.data.copy())


def emission_csc(w: Workload):
    """The per-cell emission blocks as the reference's sparse M_Emi (SparseMatrixCSC, 1-based
    Int64, every block entry stored): (colptr, rowval, nzval), or None."""
    if w.emission_blocks is None:
        return None
    n, N, C = w.ps.n, w.ps.n_momentum, w.ps.n_space
    colptr = np.empty(n + 1, np.int64)
    colptr[0] = 1
    rowval = np.empty(n * N, np.int64)
    nzval = np.empty(n * N)
    for c in range(C):
        for j in range(N):
            col = c * N + j
            k0 = col * N
            rowval[k0:k0 + N] = np.arange(c * N, (c + 1) * N) + 1
            nzval[k0:k0 + N] = w.emission_blocks[c, j, :]  # column j of the block (rows i)
            colptr[col + 1] = k0 + N + 1
    return colptr, rowval, nzval


def make_struct(w: Workload, precision="fp64", method="euler", use_monolithic=None,
                moment_correction=True, device_flux=False, sparse=False):
    """ForwardEulerStruct / ForwardSymplecticEulerStruct on the HIP backend for a workload;
    device_flux: the flux stencils, Ap and Vol built on the device from w.flux_spec;
    sparse: the monolithic M_Bin and the emission passed as CSC (Bin_sparse / Emi_sparse)."""
    from .api import (BinaryMatricesStruct, EmissionMatricesStruct, FluxMatricesStruct,
                      ForwardEulerStruct, ForwardSymplecticEulerStruct)
    mono = (w.M_Bin is not None) if use_monolithic is None else use_monolithic
    if w.synth_seed is not None:
        BinM = BinaryMatricesStruct(Binary_list=["synthetic"], Domain=w.domain,
                                    synth_seed=w.synth_seed)
    elif mono:
        BinM = BinaryMatricesStruct(M_Bin=binary_csc(w) if sparse else w.M_Bin,
                                    Binary_list=["synthetic"], Domain=w.domain)
    else:
        BinM = BinaryMatricesStruct(Binary_list=["synthetic"], Domain=w.domain, gains=w.gains,
                                    losses=w.losses)
    has_emi = w.emission_blocks is not None
    EmiM = EmissionMatricesStruct(Emission_list=["synthetic"] if has_emi else [],
                                  blocks=None if (sparse and has_emi) else w.emission_blocks,
                                  M_Emi=emission_csc(w) if (sparse and has_emi) else None)
    if device_flux:
        if w.flux_spec is None:
            raise ValueError(f"workload {w.name} has no flux stencil to build on the device")
        FluxM = FluxMatricesStruct(None, None, None, None, synth=w.flux_spec)
    else:
        FluxM = FluxMatricesStruct(w.Ap, w.X, w.P, w.Vol)
    cls = ForwardEulerStruct if method == "euler" else ForwardSymplecticEulerStruct
    return cls(w.ps, w.f0, w.inj, BinM, EmiM, FluxM, Precision=precision,
               MomentCorrection=moment_correction)
