"""Binary collision operators from DiplodocusCollisions' arrays: the interface of
BuildBinaryMatrices / LoadMatrices_Binary (src/Collisions/BuildBinaryMatrices.jl:6-73,
src/Collisions/LoadMatrices.jl:6-146), with the arrays handed over in memory.

The reference reads each interaction's GainMatrix3, GainMatrix4, LossMatrix1 and LossMatrix2 from
a JLD2 file written by DiplodocusCollisions (BinaryFileLoad_Matrix, LoadMatrices.jl:122-126) and
Fill_M_Bin! writes them into a dense or sparse N^2 x N M_Bin.  Here `arrays` supplies the same
four arrays (the file format needs DiplodocusCollisions, which is not available), and
dtx_add_binary_interaction adds each interaction's terms to the handle's operator directly, with no
N^2 x N array.  The name and scale rules of LoadMatrices_Binary (the Ele/Pos merging, Bin_Norm) are
applied here, on the host, as there.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional, Union

import numpy as np


class ModeType:
    """How the arrays are averaged over the angles (Bin_Mode)."""
    code = -1

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))

    def __repr__(self):
        return f"{type(self).__name__}()"


class Ani(ModeType):
    """Anisotropic: the arrays as they are."""
    code = 0


class Axi(ModeType):
    """Axisymmetric: averaged over phi (pz)."""
    code = 1


class Iso(ModeType):
    """Isotropic: averaged over u (py) and phi (pz)."""
    code = 2


@dataclass(frozen=True)
class BinaryStruct:
    """One binary interaction name1 + name2 -> name3 + name4 (DiplodocusCollisions' BinaryStruct,
    an element of Binary_list)."""
    name1: str
    name2: str
    name3: str
    name4: str

    @property
    def names(self):
        return (self.name1, self.name2, self.name3, self.name4)


@dataclass
class BinaryInteraction:
    """An interaction resolved against the phase space, ready for dtx_add_binary_interaction:
    species indices (0-based name_locs), the scales (GainScale, LossScale times Bin_Norm), and the
    four arrays in Julia's memory order (9-D gains (px3+2, py3, pz3, px1, py1, pz1, px2, py2, pz2),
    6-D losses (px1, py1, pz1, px2, py2, pz2))."""
    interaction: BinaryStruct
    file_names: tuple  # the names the arrays were looked up under (after the Pos -> Ele swap)
    name_locs: tuple
    gain_scale: float
    loss_scale: float
    GainMatrix3: np.ndarray
    GainMatrix4: np.ndarray
    LossMatrix1: np.ndarray
    LossMatrix2: np.ndarray


ArraySource = Union[Callable[[tuple, bool], tuple], dict]


def _lookup(arrays: ArraySource, names: tuple, corrected: bool):
    if callable(arrays):
        out = arrays(names, corrected)
    else:
        out = arrays[names]
    if len(out) != 4:
        raise ValueError("the collision arrays are (GainMatrix3, GainMatrix4, LossMatrix1, "
                         "LossMatrix2)")
    return tuple(np.asarray(a, dtype=np.float64) for a in out)


def _check_shapes(ps, locs, G3, G4, L1, L2):
    sp = ps.species

    def dims(i):
        s = sp[i]
        return (s.p_num, s.u_num, s.h_num)

    n1, n2, n3, n4 = locs
    want = {"GainMatrix3": (dims(n3)[0] + 2,) + dims(n3)[1:] + dims(n1) + dims(n2),
            "GainMatrix4": (dims(n4)[0] + 2,) + dims(n4)[1:] + dims(n1) + dims(n2),
            "LossMatrix1": dims(n1) + dims(n2), "LossMatrix2": dims(n2) + dims(n1)}
    for name, arr in zip(want, (G3, G4, L1, L2)):
        if tuple(arr.shape) != want[name]:
            raise ValueError(f"{name} has shape {tuple(arr.shape)}, expected {want[name]}")


def LoadMatrices_Binary(Binary_list, arrays: ArraySource, PhaseSpace, mode: ModeType = Ani(),
                        corrected: bool = True) -> list:
    """Resolve every interaction of Binary_list (src/Collisions/LoadMatrices.jl:6-146): species
    locations, the Ele/Pos rules when "Pos" is not a species of the phase space (:47-70), the
    Pos-Compton array swap (:73-76), and the Bin_Norm scaling (:129-133).  `arrays(names,
    corrected)` (or a dict keyed by the names) returns the four arrays BinaryFileLoad_Matrix
    would read.  As in the reference, a Pos Compton interaction met while Pos is merged into Ele
    ends the list (:51-52 return from the function)."""
    name_list = list(PhaseSpace.name_list)
    bin_norm = float(getattr(PhaseSpace, "Bin_Norm", 1.0))

    def loc(name):
        if name not in name_list:
            raise ValueError(f"species {name!r} of a binary interaction is not in the phase "
                             "space")
        return name_list.index(name)

    out = []
    for interaction in Binary_list:
        names = interaction.names
        name1, name2, name3, name4 = names
        if "Pos" in names and "Pos" not in name_list:
            # Ele stands for Ele + Pos with identical populations
            if names == ("Pos", "Pho", "Pos", "Pho"):
                return out  # Pos Compton is accounted for by the Ele population
            if names == ("Ele", "Pos", "Pho", "Pho"):  # annihilation: each is half of "Ele"
                gain_scale = loss_scale = 1.0 / 4.0
                locs = (loc(name1), loc("Ele"), loc(name3), loc(name4))
            elif names == ("Pho", "Pho", "Ele", "Pos"):  # pair production
                gain_scale = loss_scale = 1.0
                locs = (loc(name1), loc(name2), loc(name3), loc("Ele"))
            else:
                raise ValueError(f"interaction {names} involves Pos, which is not a species, "
                                 "and has no Ele/Pos rule")
        else:
            gain_scale = loss_scale = 1.0
            locs = (loc(name1), loc(name2), loc(name3), loc(name4))
        file_names = names
        if names == ("Pos", "Pho", "Pos", "Pho"):  # Pos Compton uses the Ele Compton arrays
            file_names = ("Ele", "Pho", "Ele", "Pho")
        G3, G4, L1, L2 = _lookup(arrays, file_names, corrected)
        _check_shapes(PhaseSpace, locs, G3, G4, L1, L2)
        out.append(BinaryInteraction(interaction, file_names, locs, gain_scale * bin_norm,
                                     loss_scale * bin_norm, G3, G4, L1, L2))
    return out


def BuildBinaryMatrices(PhaseSpace, Binary_list, Domain=None, DataDirectory: Optional[str] = None,
                        *, loading_check: bool = False, Bin_Mode: ModeType = Ani(),
                        Bin_corrected: bool = True, Bin_sparse: bool = False,
                        arrays: Optional[ArraySource] = None):
    """BinaryMatricesStruct of Binary_list (src/Collisions/BuildBinaryMatrices.jl:6-73) whose
    operator the stepping structs assemble from the collision arrays on the handle.  `arrays`
    takes the place of the JLD2 files in DataDirectory."""
    from .api import BinaryMatricesStruct
    if Bin_sparse and not Bin_corrected:
        raise ValueError("Bin_sparse=true requires Bin_corrected=true")  # :34-36
    if not isinstance(Bin_Mode, ModeType):
        raise TypeError("Bin_Mode must be Ani(), Axi() or Iso()")
    Binary_list = list(Binary_list)
    if not Binary_list:
        return BinaryMatricesStruct(Binary_list=[], Domain=Domain)
    if arrays is None:
        raise NotImplementedError(
            "reading DiplodocusCollisions' JLD2 files needs that package, which is not "
            f"available (DataDirectory={DataDirectory!r}); pass the arrays with arrays=")
    interactions = LoadMatrices_Binary(Binary_list, arrays, PhaseSpace, Bin_Mode, Bin_corrected)
    return BinaryMatricesStruct(Binary_list=Binary_list, Domain=Domain,
                                interactions=interactions, Bin_Mode=Bin_Mode,
                                Bin_sparse=Bin_sparse)


def angle_widths(PhaseSpace):
    """Grids.dpy_list and dpz_list concatenated over the species (deltaVector of the u and phi
    bounds, src/StructsAndDictionaries.jl:253-254)."""
    from .grids import deltaVector
    dpy = np.concatenate([deltaVector(s.ur) for s in PhaseSpace.species])
    dpz = np.concatenate([deltaVector(s.hr) for s in PhaseSpace.species])
    return dpy, dpz
