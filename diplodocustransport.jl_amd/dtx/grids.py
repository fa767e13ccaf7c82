"""Momentum/space grid helpers and the state-vector layout.

These restate the grid arithmetic that DiplodocusTransport.jl imports from DiplodocusCollisions
(bounds, deltaVector, meanVector, deltaEVector) as documented, with known-answer docstrings, in
src/ValuesOnTheGrid.jl:52-267 of the reference, and the state layout of
src/GlobalToStateIndices.jl:7-62.  They are setup-time helpers: the stepping library itself only
receives the per-state energy weights produced here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def p_range(pL: float, pH: float, num_p: int) -> np.ndarray:
    """10 .^ range(pL, pH, num_p+1)  (src/ValuesOnTheGrid.jl:128-131)."""
    return 10.0 ** np.linspace(pL, pH, num_p + 1)


def u_range(uL: float, uH: float, num_u: int) -> np.ndarray:
    """range(uL, uH, num_u+1)  (src/ValuesOnTheGrid.jl:153-155)."""
    return np.linspace(uL, uH, num_u + 1)


def bounds(low: float, up: float, num: int, grid: str) -> np.ndarray:
    """Bin bounds of one coordinate.  "l": log10-spaced between 10^low and 10^up;
    "u": uniform; "pi": uniform in units of pi (used for phi, radians)."""
    if grid == "l":
        return p_range(low, up, num)
    if grid == "u":
        return u_range(low, up, num)
    if grid == "pi":
        return u_range(low * math.pi, up * math.pi, num)
    raise ValueError(f"unknown grid type {grid!r}")


def deltaVector(valr: np.ndarray) -> np.ndarray:
    """|valr[i+1] - valr[i]|  (src/ValuesOnTheGrid.jl:67-76)."""
    valr = np.asarray(valr, dtype=np.float64)
    return np.abs(valr[1:] - valr[:-1])


def meanVector(valr: np.ndarray) -> np.ndarray:
    """(valr[i+1] + valr[i]) / 2  (src/ValuesOnTheGrid.jl:95-104)."""
    valr = np.asarray(valr, dtype=np.float64)
    return (valr[1:] + valr[:-1]) / 2


def deltaEVector(pr: np.ndarray, mu: float) -> np.ndarray:
    """Bin-integrated energy  (src/ValuesOnTheGrid.jl:199-223), in its cancellation-free form."""
    pr = np.asarray(pr, dtype=np.float64)
    if mu == 0.0:
        return (pr[1:] ** 2 - pr[:-1] ** 2) / 2
    E = mu + pr ** 2 / (np.sqrt(mu ** 2 + pr ** 2) + mu)
    dE = (pr[1:] - pr[:-1]) * mu
    dE = dE + (pr[1:] ** 3 / (E[1:] + mu) - pr[:-1] ** 3 / (E[:-1] + mu))
    dE = dE + mu ** 2 * (np.arcsinh(pr[1:] / mu) - np.arcsinh(pr[:-1] / mu))
    return dE / 2


def deltaEkinVector(pr: np.ndarray, mu: float) -> np.ndarray:
    """Bin-integrated kinetic energy  (src/ValuesOnTheGrid.jl:242-267)."""
    pr = np.asarray(pr, dtype=np.float64)
    if mu == 0.0:
        return (pr[1:] ** 2 - pr[:-1] ** 2) / 2
    E = mu + pr ** 2 / (np.sqrt(mu ** 2 + pr ** 2) + mu)
    dE = pr[1:] ** 3 / (E[1:] + mu) - pr[:-1] ** 3 / (E[:-1] + mu)
    dE = dE + mu ** 2 * (np.arcsinh(pr[1:] / mu) - np.arcsinh(pr[:-1] / mu))
    return dE / 2


def energy_weight_T00(pr: np.ndarray, m: float) -> np.ndarray:
    """Per-p-bin T^00 weight exactly as written in StressEnergyTensor
    (src/DistributionMoments.jl:145 massive, :162 massless).  Mathematically equal to
    deltaEVector(pr, m) / deltaVector(pr)  (Grids.dE_list, src/StructsAndDictionaries.jl:259)."""
    pr = np.asarray(pr, dtype=np.float64)
    p0, p1 = pr[:-1], pr[1:]
    if m != 0.0:
        return (p0 * np.sqrt(m * m + p0 * p0) - p1 * np.sqrt(m * m + p1 * p1)
                + m * m * (np.arcsinh(p0 / m) - np.arcsinh(p1 / m))) / (2 * (p0 - p1))
    return (p0 + p1) / 2


@dataclass
class Species:
    name: str
    mass: float          # normalised mass (units of m_e); 0 for photons
    charge: float        # charge number
    p_num: int
    u_num: int
    h_num: int
    p_low: float = -3.0  # log10 bounds of p
    p_up: float = 3.0
    u_low: float = -1.0
    u_up: float = 1.0
    h_low: float = 0.0   # phi bounds in units of pi
    h_up: float = 2.0

    @property
    def size(self) -> int:
        return self.p_num * self.u_num * self.h_num

    @property
    def pr(self) -> np.ndarray:
        return bounds(self.p_low, self.p_up, self.p_num, "l")

    @property
    def ur(self) -> np.ndarray:
        return bounds(self.u_low, self.u_up, self.u_num, "u")

    @property
    def hr(self) -> np.ndarray:
        return bounds(self.h_low, self.h_up, self.h_num, "pi")


@dataclass
class PhaseSpace:
    """The subset of PhaseSpaceStruct (src/StructsAndDictionaries.jl:290-308) the hot path needs:
    species momentum grids, the space grid and dt0."""
    species: list
    x_num: int = 1
    y_num: int = 1
    z_num: int = 1
    x_low: float = 0.0
    x_up: float = 1.0
    y_low: float = 0.0
    y_up: float = 1.0
    z_low: float = 0.0
    z_up: float = 1.0
    dt0: float = 1.0
    name_list: list = field(default_factory=list)
    # Characteristic.Bin_Norm (src/StructsAndDictionaries.jl:41): DiplodocusCollisions' rates in
    # the units of the step; multiplies every collision array on the way in
    Bin_Norm: float = 1.0

    def __post_init__(self):
        if not self.name_list:
            self.name_list = [s.name for s in self.species]

    @property
    def num_species(self) -> int:
        return len(self.species)

    @property
    def n_momentum(self) -> int:
        """N: momentum states per cell summed over species."""
        return sum(s.size for s in self.species)

    @property
    def n_space(self) -> int:
        return self.x_num * self.y_num * self.z_num

    @property
    def n(self) -> int:
        return self.n_momentum * self.n_space

    @property
    def species_offset(self) -> list:
        """momentum_species_offset (src/StructsAndDictionaries.jl:261-265)."""
        off, out = 0, []
        for s in self.species:
            out.append(off)
            off += s.size
        return out

    @property
    def xr(self):
        return bounds(self.x_low, self.x_up, self.x_num, "u")

    @property
    def yr(self):
        return bounds(self.y_low, self.y_up, self.y_num, "u")

    @property
    def zr(self):
        return bounds(self.z_low, self.z_up, self.z_num, "u")

    def state_index(self, x, y, z, px, py, pz, species_index) -> int:
        """GlobalIndices_To_StateIndex (src/GlobalToStateIndices.jl:7-25); 1-based in and out."""
        s = self.species[species_index - 1]
        off_space = (x - 1) * self.y_num * self.z_num + (y - 1) * self.z_num + z - 1
        off_mom = (pz - 1) * s.p_num * s.u_num + (py - 1) * s.p_num + px
        return off_space * self.n_momentum + self.species_offset[species_index - 1] + off_mom

    def view(self, state: np.ndarray, species_index: int, x_idx=1, y_idx=1, z_idx=1):
        """Location_Species_To_StateVector (src/GlobalToStateIndices.jl:35-62), 1-based species."""
        if species_index == 0:
            raise ValueError("Species not defined")
        s = self.species[species_index - 1]
        off_space = (x_idx - 1) * self.y_num * self.z_num + (y_idx - 1) * self.z_num + z_idx - 1
        start = self.n_momentum * off_space + self.species_offset[species_index - 1]
        return state[start:start + s.size]

    def energy_weights(self) -> np.ndarray:
        """Per-state T^00 weight E_bar(p) for one cell (length N), species blocks in order."""
        out = np.empty(self.n_momentum)
        for s, off in zip(self.species, self.species_offset):
            w = energy_weight_T00(s.pr, s.mass)
            out[off:off + s.size] = np.tile(w, s.u_num * s.h_num)
        return out
