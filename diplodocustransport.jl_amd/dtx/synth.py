"""Seeded synthetic workloads for the explicit-Euler hot path.

The reference builds its operators from DiplodocusCollisions data files, which are not available;
these generators produce inputs of the same shapes and layouts:

* counter-based random numbers: SplitMix64 of (key, index) -> uniform in (0,1) -> Box-Muller
  normal -> lognormal.  The same hash runs on the device (csrc/synth.hip) so large tensors can be
  generated in HBM; integer hashing is bit-identical on both sides.
* binary gain/loss channels with number conservation per input pair (i, j):
  sum_k G3[k,i,j] = L1[i,j] and sum_k G4[k,i,j] = L2[j,i], assembled into the monolithic
  (N^2 x N) tensor exactly as GainMatrix_to_M_Bin! / LossMatrix_to_M_Bin! do
  (src/Collisions/BuildBinaryMatrices.jl:181-196, 261-276);
* upwind finite-volume flux matrices in Julia CSC form (1-based Int64): x-advection with the
  bin-averaged velocity and synchrotron cooling in p, following the stencil of Fill_B_Flux! /
  Fill_I_Flux! (src/Fluxes/BuildFluxMatrices.jl:275-413, 708-837);
* Ap (Cartesian A flux / MomentumSpaceNorm = dx dy dz) and Vol = dt0 dx dy dz
  (src/Fluxes/FluxFunctionsCoordinate.jl:43-48, 311-315);
* Maxwell-Juttner initial conditions integrated with the 10-sample trapezium rule
  (src/DistributionFunctions.jl:112-163, 177-193) and renormalised to num_Init
  (src/InitialConditions.jl:276-277).  For massless species the reference's formula is singular
  (invtheta / m^3); the massless limit exp(-p/Theta) is used.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grids import PhaseSpace, Species, deltaVector

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
SEED_BASE = 20261203


def config_seed(config: int, request: int = 0) -> int:
    """Seed convention: base 20261203 + 1000*config + request id."""
    return SEED_BASE + 1000 * config + request


def stream_key(seed: int, stream: int) -> int:
    """Key of an independent stream: splitmix of (seed, stream)."""
    return int(splitmix64(np.array([stream], dtype=np.uint64), seed)[0])


def splitmix64(idx: np.ndarray, key: int) -> np.ndarray:
    """z = key + (idx+1)*GOLDEN, then the SplitMix64 finaliser (wrapping uint64 arithmetic)."""
    with np.errstate(over="ignore"):
        z = np.uint64(key & 0xFFFFFFFFFFFFFFFF) + (idx.astype(np.uint64) + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(idx: np.ndarray, key: int) -> np.ndarray:
    """Uniform in (0,1): ((z >> 11) + 0.5) * 2^-53."""
    z = splitmix64(idx, key)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(idx: np.ndarray, key: int) -> np.ndarray:
    """Standard normal by Box-Muller (cosine branch) from uniforms 2*idx and 2*idx+1."""
    idx = np.asarray(idx, dtype=np.uint64)
    u1 = uniform(idx * np.uint64(2), key)
    u2 = uniform(idx * np.uint64(2) + np.uint64(1), key)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def lognormal(count: int, key: int, start: int = 0) -> np.ndarray:
    idx = np.arange(start, start + count, dtype=np.uint64)
    return np.exp(normal(idx, key))


# ------------------------------------------------------------------ binary channels ---------

@dataclass
class GainChannel:
    s_out: int          # 0-based species
    s1: int
    s2: int
    G: np.ndarray       # shape (No, N1, N2) in Fortran order: G[k, i, j], k fastest


@dataclass
class LossChannel:
    s1: int
    s2: int
    L: np.ndarray       # shape (N1, N2) Fortran order


def interactions(num_species: int) -> list:
    """Synthetic interaction list: every unordered species pair (a <= b) scatters into (a, b)."""
    return [(a, b) for a in range(num_species) for b in range(a, num_species)]


def binary_channels(ps: PhaseSpace, seed: int, scale: float = 1.0):
    """Lognormal(0,1) gain blocks per interaction with number-conserving losses.

    Interaction a+b -> a+b (a < b): G3 (out a) and G4 (out b), each (No x Na x Nb), with
    L1[i,j] = sum_k G3[k,i,j] (loss of species a) and L2[j,i] = sum_k G4[k,i,j] (loss of b).
    Interaction a+a -> a+a: one block G (Na^3) with L[i,j] = sum_k G[k,i,j].
    So for every input pair the total gain equals the total loss: number is conserved.
    """
    gains, losses = [], []
    stream = 0
    for (a, b) in interactions(ps.num_species):
        Na, Nb = ps.species[a].size, ps.species[b].size
        outs = [a] if a == b else [a, b]
        for so in outs:
            No = ps.species[so].size
            key = stream_key(seed, stream)
            stream += 1
            G = (scale * lognormal(No * Na * Nb, key)).reshape((No, Na, Nb), order="F")
            gains.append(GainChannel(so, a, b, G))
            colsum = G.sum(axis=0)  # (Na, Nb)
            if a == b:
                losses.append(LossChannel(a, a, np.asfortranarray(colsum)))
            elif so == a:
                losses.append(LossChannel(a, b, np.asfortranarray(colsum)))
            else:
                losses.append(LossChannel(b, a, np.asfortranarray(colsum.T)))
    return gains, losses


def assemble_M_Bin(ps: PhaseSpace, gains, losses) -> np.ndarray:
    """Monolithic (N^2 x N) column-major tensor exactly as the reference's writers build it:
    M_Bin[(b-1)N+a, c] += val/2 and M_Bin[(c-1)N+a, b] += val/2 for gains (a out, b in1, c in2);
    M_Bin[(a-1)N+a, c] -= val/2 and M_Bin[(c-1)N+a, a] -= val/2 for losses.
    Returned as a 3-D Fortran array M[a, b, c] whose memory is M_Bin."""
    N = ps.n_momentum
    off = ps.species_offset
    M = np.zeros((N, N, N), order="F")
    for g in gains:
        o, p, q = off[g.s_out], off[g.s1], off[g.s2]
        No, N1, N2 = g.G.shape
        half = g.G / 2
        M[o:o + No, p:p + N1, q:q + N2] += half
        M[o:o + No, q:q + N2, p:p + N1] += half.transpose(0, 2, 1)
    for l in losses:
        p, q = off[l.s1], off[l.s2]
        N1, N2 = l.L.shape
        half = l.L / 2
        for a in range(N1):
            M[p + a, p + a, q:q + N2] -= half[a, :]
            M[p + a, q:q + N2, p + a] -= half[a, :]
    return M


def collision_arrays(ps: PhaseSpace, locs, seed: int, scale: float = 1.0):
    """DiplodocusCollisions-shaped arrays of one interaction 1 + 2 -> 3 + 4 (species `locs`,
    0-based): GainMatrix3 (px3+2, py3, pz3, px1, py1, pz1, px2, py2, pz2), GainMatrix4 likewise
    for species 4, LossMatrix1 (px1, py1, pz1, px2, py2, pz2) and LossMatrix2 (2 against 1), all
    Fortran-ordered.  Gains are lognormal values from numpy's generator seeded with (seed,
    locs) (under/overflow px3 bins zero; host arrays only, nothing regenerates them on the
    device); the losses are the gains summed over the outgoing states, L1[b, c] =
    sum_a G3[a; b, c] and L2[c, b] = sum_a G4[a; b, c], so every collision conserves the number
    of particles."""
    dims = [(ps.species[s].p_num, ps.species[s].u_num, ps.species[s].h_num) for s in locs]
    d1, d2, d3, d4 = dims
    n1, n2 = int(np.prod(d1)), int(np.prod(d2))
    rng = np.random.default_rng([seed] + [int(v) for v in locs])
    out = []
    for d in (d3, d4):
        G = np.zeros((d[0] + 2,) + d[1:] + d1 + d2, order="F")
        G[1:-1] = rng.lognormal(sigma=1.0, size=d + d1 + d2) * scale
        out.append(G)
    G3, G4 = out
    g3 = G3.reshape(-1, n1, n2, order="F").sum(axis=0)
    g4 = G4.reshape(-1, n1, n2, order="F").sum(axis=0)
    L1 = np.asfortranarray(g3.reshape(d1 + d2, order="F"))
    L2 = np.asfortranarray(g4.T.reshape(d2 + d1, order="F"))
    return G3, G4, L1, L2


def max_loss_rate(ps: PhaseSpace, losses, f_cell: np.ndarray, vol: float, ap: float) -> float:
    """max_a  Vol/Ap * sum_c L[a,c] f_c  over one cell (the stiff diagonal rate)."""
    off = ps.species_offset
    rate = np.zeros(ps.n_momentum)
    for l in losses:
        p, q = off[l.s1], off[l.s2]
        N1, N2 = l.L.shape
        rate[p:p + N1] += l.L @ f_cell[q:q + N2]
    return float(rate.max() * vol / ap)


# ------------------------------------------------------------------ initial conditions ------

def maxwell_juttner_block(s: Species, theta: float, num_init: float = 1.0) -> np.ndarray:
    """DistributionToDIP_TrapeziumIntegration! of a Maxwell-Juttner distribution for one species,
    renormalised to num_init; returned flattened with p fastest, then u, then phi."""
    samples = 10
    pr, ur, hr = s.pr, s.ur, s.hr
    w1 = np.ones(samples)
    w1[0] = w1[-1] = 0.5
    fp = np.zeros(s.p_num)
    for px in range(s.p_num):
        x = np.linspace(pr[px], pr[px + 1], samples)
        if s.mass != 0.0:
            E = np.sqrt(s.mass ** 2 + x ** 2)
            dis = np.exp((s.mass - E) / theta)
        else:
            dis = np.exp(-x / theta)
        fp[px] = np.sum(w1 * dis * x ** 2) * (pr[px + 1] - pr[px]) / (samples - 1)
    du = deltaVector(ur)
    dh = deltaVector(hr)
    f3 = fp[:, None, None] * du[None, :, None] * dh[None, None, :]
    f3 *= num_init / f3.sum()
    return np.asarray(f3, order="F").reshape(-1, order="F")


def initial_state(ps: PhaseSpace, theta: float, seed: int, noise: float = 0.01,
                  num_init: float = 1.0) -> np.ndarray:
    """Maxwell-Juttner at Theta in every cell and species, times (1 + noise*N(0,1))."""
    f = np.zeros(ps.n)
    N = ps.n_momentum
    blocks = [maxwell_juttner_block(s, theta, num_init) for s in ps.species]
    base = np.concatenate(blocks)
    for c in range(ps.n_space):
        f[c * N:(c + 1) * N] = base
    if noise:
        z = normal(np.arange(ps.n, dtype=np.uint64), stream_key(seed, 1000))
        f = f * (1.0 + noise * z)
    return f


# ------------------------------------------------------------------ flux matrices -----------

def ap_vol(ps: PhaseSpace):
    """Ap (length n) and Vol (length C) for Cartesian space: Ap = dx dy dz, Vol = dt0 dx dy dz."""
    dx, dy, dz = deltaVector(ps.xr), deltaVector(ps.yr), deltaVector(ps.zr)
    vol = np.empty(ps.n_space)
    ap = np.empty(ps.n)
    N = ps.n_momentum
    for x in range(ps.x_num):
        for y in range(ps.y_num):
            for z in range(ps.z_num):
                off = x * ps.y_num * ps.z_num + y * ps.z_num + z
                v = dx[x] * dy[y] * dz[z]
                vol[off] = ps.dt0 * v
                ap[off * N:(off + 1) * N] = v
    return ap, vol


def coo_to_csc(n: int, I, J, V):
    """Julia `sparse(I, J, V, n, n)` for 1-based (I, J): duplicates summed in push order,
    explicit zeros kept.  Returns (colptr, rowval, nzval) as 1-based Int64 / Float64."""
    I = np.asarray(I, dtype=np.int64)
    J = np.asarray(J, dtype=np.int64)
    V = np.asarray(V, dtype=np.float64)
    if I.size == 0:
        return np.ones(n + 1, dtype=np.int64), np.zeros(0, np.int64), np.zeros(0)
    order = np.lexsort((np.arange(I.size), I, J))  # by column, then row, then push order
    I, J, V = I[order], J[order], V[order]
    key = J * (n + 1) + I
    first = np.ones(I.size, dtype=bool)
    first[1:] = key[1:] != key[:-1]
    grp = np.cumsum(first) - 1
    nz = np.zeros(int(grp[-1]) + 1)
    np.add.at(nz, grp, V)  # unbuffered, applied in (sorted = push) order
    rows = I[first]
    cols = J[first]
    colptr = np.ones(n + 1, dtype=np.int64)
    colptr[1:] = 1 + np.cumsum(np.bincount(cols - 1, minlength=n))
    return colptr, rows.astype(np.int64), nz


def csc_to_coo(csc):
    colptr, rowval, nzval = csc
    n = colptr.size - 1
    cols = np.repeat(np.arange(1, n + 1, dtype=np.int64), np.diff(colptr))
    return rowval, cols, nzval


def csc_add(n, A, B):
    """X + P for two CSCs: union pattern, explicit zeros kept."""
    Ia, Ja, Va = csc_to_coo(A)
    Ib, Jb, Vb = csc_to_coo(B)
    return coo_to_csc(n, np.concatenate([Ia, Ib]), np.concatenate([Ja, Jb]),
                      np.concatenate([Va, Vb]))


def csc_to_dense(csc, n):
    out = np.zeros((n, n))
    I, J, V = csc_to_coo(csc)
    np.add.at(out, (I - 1, J - 1), V)
    return out


def _species_momentum(s: Species):
    """Per-state (p fastest, then u, then phi) bin data of one species."""
    pr, ur, hr = s.pr, s.ur, s.hr
    P, U, H = np.meshgrid(np.arange(s.p_num), np.arange(s.u_num), np.arange(s.h_num),
                          indexing="ij")
    P, U, H = (A.reshape(-1, order="F") for A in (P, U, H))
    return pr, ur, hr, P, U, H


def _upwind(flux_plus, flux_minus):
    """h weights of the upwind scheme (src/Fluxes/BuildFluxMatrices.jl:338-352)."""
    hpr = np.where(np.sign(flux_plus) == 1, 0.0, 1.0)
    hpl = 1.0 - hpr
    hmr = np.where(np.sign(flux_minus) == 1, 1.0, 0.0)
    hml = 1.0 - hmr
    return hpr, hpl, hmr, hml


def flux_matrices(ps: PhaseSpace, sync_K: float = 0.0, periodic: bool = True,
                  axes: str = "x"):
    """X_Flux (space) and P_Flux (momentum) as 1-based CSC, Cartesian space, upwind.

    Space axis 'x': advection with the bin-averaged velocity component along the u axis,
    <v><u> = (E1-E0)/dp * (u1^2-u0^2)/(2du) (the surface integral of
    src/Fluxes/FluxFunctionsCoordinate.jl:264-267 with the advection axis along u).
    Space axis 'y': the perpendicular component <v><sqrt(1-u^2)><sin(phi)>.
    Each face flux carries (t1-t0) and the two transverse widths; the entries are divided by
    MomentumSpaceNorm (dp du dphi), exactly as Fill_B_Flux! pushes them (:801-833).
    Momentum axes p and u (sync_K > 0): synchrotron radiation reaction,
    I(p) = -K/(3m) p sqrt(m^2+p^2) (g(u1)-g(u0)) dphi dt dx dy dz with g(u) = 3u-u^3
    (src/Fluxes/FluxFunctionsEmissionForces.jl:46-70), stencil of Fill_I_Flux! (:308-408), and
    J(u) = K (1/2) m u (u^2-1) (h0-h1) (2 asinh(p1/m) - 2 asinh(p0/m)) dt dx dy dz (:84-150),
    stencil of Fill_J_Flux! (:420-557), pushed after I; closed boundaries, zero for massless or
    neutral species (zero-valued entries still pushed).  The K (phi) flux of the synchrotron
    force is zero in every mode (:152-215); its stored zeros are not pushed.
    Zero-weighted upwind entries are pushed too, so the CSC keeps them as explicit zeros."""
    n, N, C = ps.n, ps.n_momentum, ps.n_space
    dx, dy, dz = deltaVector(ps.xr), deltaVector(ps.yr), deltaVector(ps.zr)
    dt0 = ps.dt0
    XI, XJ, XV, PI, PJ, PV = [], [], [], [], [], []
    cx, cy, cz = np.meshgrid(np.arange(ps.x_num), np.arange(ps.y_num), np.arange(ps.z_num),
                             indexing="ij")
    cx, cy, cz = cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)  # z fastest = cell order
    cell = cx * ps.y_num * ps.z_num + cy * ps.z_num + cz
    for si, s in enumerate(ps.species):
        pr, ur, hr, P, U, H = _species_momentum(s)
        so = ps.species_offset[si]
        m = s.mass
        E = np.sqrt(m * m + pr * pr)
        vbar = (E[P + 1] - E[P]) / (pr[P + 1] - pr[P])
        u0, u1, h0, h1 = ur[U], ur[U + 1], hr[H], hr[H + 1]
        du, dh = u1 - u0, h1 - h0
        dpv = pr[P + 1] - pr[P]
        a = (cell[:, None] * N + so + np.arange(s.size)[None, :])  # 0-based (C, Ns)
        for ax in axes:
            num = ps.x_num if ax == "x" else ps.y_num
            if num == 1:
                continue
            if ax == "x":
                vel = vbar * (u1 * u1 - u0 * u0) / (2 * du)
                c_idx, trans = cx, (dy[cy] * dz[cz])
            else:
                sq = lambda u: u * np.sqrt(1 - u * u) + np.arcsin(u)  # noqa: E731
                vel = vbar * (sq(u1) - sq(u0)) / (2 * du) * (np.cos(h0) - np.cos(h1)) / dh
                c_idx, trans = cy, (dx[cx] * dz[cz])
            cp = c_idx + 1
            cm = c_idx - 1
            if periodic:
                cp = np.where(cp >= num, 0, cp)
                cm = np.where(cm < 0, num - 1, cm)
            else:
                cp = np.minimum(cp, num - 1)
                cm = np.maximum(cm, 0)
            if ax == "x":
                cellp = cp * ps.y_num * ps.z_num + cy * ps.z_num + cz
                cellm = cm * ps.y_num * ps.z_num + cy * ps.z_num + cz
            else:
                cellp = cx * ps.y_num * ps.z_num + cp * ps.z_num + cz
                cellm = cx * ps.y_num * ps.z_num + cm * ps.z_num + cz
            bp = cellp[:, None] * N + so + np.arange(s.size)[None, :]
            bm = cellm[:, None] * N + so + np.arange(s.size)[None, :]
            Bp = (dt0 * trans)[:, None] * vel[None, :]
            Bm = -Bp
            hpr, hpl, hmr, hml = _upwind(Bp, Bm)
            okp, okm = a != bp, a != bm
            for rows, cols, vals, ok in ((a, bp, Bp * hpr, okp), (a, a, Bp * hpl, okp),
                                         (a, a, Bm * hmr, okm), (a, bm, Bm * hml, okm)):
                XI.append(rows[ok]); XJ.append(cols[ok]); XV.append(vals[ok])
        if sync_K > 0.0:
            g = lambda u: 3 * u - u ** 3  # noqa: E731
            ang = g(u1) - g(u0)
            vol = (dt0 * dx[cx] * dy[cy] * dz[cz])[:, None]
            norm = dpv * du * dh
            if m != 0.0 and s.charge != 0.0:
                Ifn = lambda pp: -sync_K / (3 * m) * pp * np.sqrt(m * m + pp * pp) * ang * dh  # noqa
                Ip = vol * Ifn(pr[P + 1])[None, :]
                Im = -(vol * Ifn(pr[P])[None, :])
            else:
                Ip = np.zeros((C, s.size))
                Im = -np.zeros((C, s.size))
            hpr, hpl, hmr, hml = _upwind(Ip, Im)
            up = (P < s.p_num - 1)
            dn = (P > 0)
            bp = a + 1   # p+1 is the next state (p fastest)
            bm = a - 1
            norm_p = np.where(up, np.roll(norm, -1), 1.0)
            norm_m = np.where(dn, np.roll(norm, 1), 1.0)
            okp = np.broadcast_to(up[None, :], a.shape)
            okm = np.broadcast_to(dn[None, :], a.shape)
            for rows, cols, vals, ok in ((a, bp, Ip * hpr / norm_p, okp),
                                         (a, a, Ip * hpl / norm, okp),
                                         (a, a, Im * hmr / norm, okm),
                                         (a, bm, Im * hml / norm_m, okm)):
                PI.append(rows[ok]); PJ.append(cols[ok]); PV.append(vals[ok])
            # the J stencil along u (Fill_J_Flux!, src/Fluxes/BuildFluxMatrices.jl:420-557),
            # pushed after the I entries as BuildFluxMatrices pushes them (:117-120):
            # J(u) = K (1/2) m u (u^2 - 1) (h0 - h1) (2 asinh(p1/m) - 2 asinh(p0/m)) dt dx dy dz
            # (JFluxFunction, src/Fluxes/FluxFunctionsEmissionForces.jl:84-150, Ani / Axi),
            # closed u boundaries, zero-valued entries for massless or neutral species
            if m != 0.0 and s.charge != 0.0:
                ash = 2.0 * np.arcsinh(pr[P + 1] / m) - 2.0 * np.arcsinh(pr[P] / m)
                Jfn = lambda uu: sync_K * (0.5 * m * uu * (uu * uu - 1.0) * (h0 - h1) * ash)  # noqa
                Jp = vol * Jfn(u1)[None, :]
                Jm = -(vol * Jfn(u0)[None, :])
            else:
                Jp = np.zeros((C, s.size))
                Jm = -np.zeros((C, s.size))
            hpr, hpl, hmr, hml = _upwind(Jp, Jm)
            up = (U < s.u_num - 1)
            dn = (U > 0)
            bp = a + s.p_num   # u+1 is p_num states further (p fastest, then u)
            bm = a - s.p_num
            norm_p = np.where(up, np.roll(norm, -s.p_num), 1.0)
            norm_m = np.where(dn, np.roll(norm, s.p_num), 1.0)
            okp = np.broadcast_to(up[None, :], a.shape)
            okm = np.broadcast_to(dn[None, :], a.shape)
            for rows, cols, vals, ok in ((a, bp, Jp * hpr / norm_p, okp),
                                         (a, a, Jp * hpl / norm, okp),
                                         (a, a, Jm * hmr / norm, okm),
                                         (a, bm, Jm * hml / norm_m, okm)):
                PI.append(rows[ok]); PJ.append(cols[ok]); PV.append(vals[ok])

    def build(I, J, V):
        if not I:
            return coo_to_csc(n, [], [], [])
        # interleave per state in push order: concatenation keeps each list's internal order;
        # duplicates of one (row, col) are summed in the order pushed above.
        return coo_to_csc(n, np.concatenate(I) + 1, np.concatenate(J) + 1, np.concatenate(V))
    return build(XI, XJ, XV), build(PI, PJ, PV)


def sync_outflow_coefficient(ps: PhaseSpace, P_csc) -> float:
    """Largest diagonal outflow coefficient of P_Flux times invAp (rate per unit dt_scale)."""
    ap, _ = ap_vol(ps)
    I, J, V = csc_to_coo(P_csc)
    diag = I == J
    if not diag.any():
        return 0.0
    return float(np.max(V[diag] / ap[I[diag] - 1]))
