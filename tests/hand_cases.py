"""Small problems whose step results and diagnostics are worked out by hand.

Each case is one functor call with dt = dt0 (so dt_scale = 1), Ap = 1 (invAp = 1) and Vol = 1,
small binary fractions (every product and sum below is exact in binary64), and the expected
values derived in the comments from the reference's lines:
  ForwardEuler            src/SteppingMethods.jl:12-107
  ForwardSymplecticEuler  src/SteppingMethods.jl:120-243
They pin the oracle's and the brute force's diagnostics (finite flag, Cr and its parts) and the
HIP kernels' elementwise semantics independently of any second restatement.
"""
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from dtx import grids

INF, NAN = math.inf, math.nan
THR10 = float(np.float32(1e-10))   # Float64(1f-10) = 1.00000001335143196e-10
THR20 = float(np.float32(1e-20))   # Float64(1f-20) = 9.999999682655225e-21


def csc(dense):
    """1-based Int64 CSC (colptr, rowval, nzval) of a small dense matrix, keeping every entry
    given as non-zero in `dense` (explicit zeros are not needed by these cases)."""
    dense = np.asarray(dense, dtype=np.float64)
    n = dense.shape[1]
    colptr, rowval, nzval = [1], [], []
    for j in range(n):
        for i in range(dense.shape[0]):
            if dense[i, j] != 0.0:
                rowval.append(i + 1)
                nzval.append(dense[i, j])
        colptr.append(len(rowval) + 1)
    return (np.array(colptr, np.int64), np.array(rowval, np.int64), np.array(nzval))


@dataclass
class HandCase:
    name: str
    method: int                       # 0 ForwardEuler, 1 ForwardSymplecticEuler
    f: list
    F: Optional[tuple] = None         # ForwardEuler flux (CSC)
    X: Optional[tuple] = None         # symplectic space flux
    P: Optional[tuple] = None         # symplectic momentum flux
    M_Bin: Optional[np.ndarray] = None  # (N, N, N) Fortran: M[k, i, j]
    inj: Optional[list] = None
    expect_f: list = field(default_factory=list)   # after the step (NaN: any NaN)
    nan_sign_free: list = field(default_factory=list)  # indices clamped from a NaN: +-0 allowed
    expect: dict = field(default_factory=dict)     # stats at Verbose 3
    expect_v1: dict = field(default_factory=dict)  # stats at Verbose 1

    @property
    def n(self):
        return len(self.f)

    def phase_space(self):
        return grids.PhaseSpace([grids.Species("Ele", 1.0, -1.0, self.n, 1, 1)], dt0=1.0)

    @property
    def injection(self):
        return np.asarray(self.inj if self.inj is not None else [0.0] * self.n)


# F = [[0.5, 0, 0], [-0.5, 0.25, 0], [0, -0.25, 0]]: CSC column order gives
#   dF[0] = 0.5*f0,  dF[1] = -0.5*f0 + 0.25*f1,  dF[2] = -0.25*f1.
F3 = csc([[0.5, 0.0, 0.0], [-0.5, 0.25, 0.0], [0.0, -0.25, 0.0]])


def _m_bin_hand():
    """M[0,1,2] = 1 (dB[0] = f1 f2) and M[1,1,1] = -1 (dB[1] = -f1^2), everything else 0."""
    M = np.zeros((3, 3, 3), order="F")
    M[0, 1, 2] = 1.0
    M[1, 1, 1] = -1.0
    return M


EULER_BINARY = HandCase(
    name="euler_binary_flux", method=0, f=[1.0, 2.0, 4.0], F=F3, M_Bin=_m_bin_hand(),
    # dF = [0.5, -0.5 + 0.5 = +0.0, -0.5]; dB = [2*4, -(2*2), 0] = [8, -4, 0]
    # df = -dF + dB = [7.5, -0.0 + -4 = -4, 0.5]; sum(df) = 4 -> finite
    # Cr_Bin = -min(8/1, -4/2, 0/4) = 2;  Cr_Flux = -min(-0.5/1, -0.0/2, 0.5/4) = 0.5
    # Cr = -min(7.5, -2, 0.125) = 2
    # f + df = [8.5, -2, 4.5] -> clamp: -2 < 1f-10 -> -2 * false = -0.0
    expect_f=[8.5, -0.0, 4.5],
    expect=dict(finite=1, cr_evaluated=1, cr=2.0, cr_bin=2.0, cr_emi=0.0, cr_flux=0.5),
    expect_v1=dict(finite=1, cr_evaluated=1, cr=2.0, cr_bin=0.0, cr_emi=0.0, cr_flux=0.0),
)

EULER_NONFINITE = HandCase(
    name="euler_nonfinite", method=0, f=[1.0, INF, 4.0], F=F3,
    # CSC order: dF[1] = -0.5*1 + 0.25*Inf = Inf, dF[2] = -0.25*Inf = -Inf, dF[0] = 0.5
    # df = [-0.5, -Inf, Inf]; sum(df) = NaN -> finite = 0 ("non-finite value in df calculation")
    # sum(f) = Inf != 0: Cr = -min(-0.5/1, -Inf/Inf = NaN -> Inf, Inf/4) = 0.5; Cr_Flux the same
    # f + df = [0.5, Inf - Inf = NaN, 4 + Inf = Inf]
    #   NaN >= 1f-10 is false -> NaN * false = +-0 (the NaN's sign); Inf * (Inf != Inf) = +0.0
    expect_f=[0.5, 0.0, 0.0], nan_sign_free=[1],
    expect=dict(finite=0, cr_evaluated=1, cr=0.5, cr_bin=0.0, cr_emi=0.0, cr_flux=0.5),
    expect_v1=dict(finite=0, cr_evaluated=1, cr=0.5),
)

SYMPLECTIC = HandCase(
    name="symplectic_pflux_nan", method=1, f=[2.0, 0.0],
    P=csc([[-0.5, 0.0], [0.0, 0.0]]), X=csc([[0.5, 0.0], [-0.5, 0.0]]),
    # dP = [-1, +0.0 (no entries)]; df_Momentum = -dP = [1, -0.0]
    # f_Momentum = f + df_Momentum = [3, 0 + -0.0 = +0.0] -> clamp 1f-10: [3, +0.0]
    # dX = X f_Momentum = [1.5, -1.5]; df_Space = -dX = [-1.5, 1.5]; f_Space = [1.5, 1.5]
    # Cr_PFlux: -dP/f = [0.5, -0.0/0 = NaN] -> replace!(NaN => 0) -> [0.5, 0.0] -> -min = -0.0
    #   (the isnan -> Inf map alone would give -0.5)
    # Cr_Momentum = -min(1/2, -0.0/0 = NaN -> Inf) = -0.5
    # Cr_Space = -min(-1.5/3, 1.5/0 = Inf) = 0.5
    # Cr = -min((-1.5 + 1)/2, (1.5 - 0.0)/0 = Inf) = 0.25
    # f = f_Space * (f >= 1f-20) * sign(f) = [1.5, 1.5]
    expect_f=[1.5, 1.5],
    expect=dict(finite=1, cr_evaluated=1, cr=0.25, cr_bin=0.0, cr_emi=0.0, cr_flux=-0.0,
                cr_momentum=-0.5, cr_space=0.5),
    expect_v1=dict(cr_evaluated=1, cr=0.25, cr_flux=0.0, cr_momentum=0.0, cr_space=0.0),
)

# Final clamp of the symplectic method, f * (f >= 1f-20) * sign(f) then f * (f != Inf)
# (src/SteppingMethods.jl:233-237), on values brought in through the injection term: with f = 0
# and no operators, f_Momentum = f_Space = +0.0 and the final state is +0.0 + df_Inj.
SYMPLECTIC_CLAMP_INJ = [-1.0, 5e-21, THR20, float(np.nextafter(THR20, 0.0)), INF, 2.0, NAN,
                        -INF, 1e-15, -0.0, 1e-300]
SYMPLECTIC_CLAMP = HandCase(
    name="symplectic_final_clamp", method=1, f=[0.0] * len(SYMPLECTIC_CLAMP_INJ),
    inj=SYMPLECTIC_CLAMP_INJ,
    #  -1      -> -1 * false = -0.0, times sign(-1) = -1 -> +0.0
    #  5e-21   -> below 1f-20 -> +0.0 (times sign +1)
    #  1f-20   -> kept
    #  1f-20 - 1 ulp -> +0.0
    #  Inf     -> Inf * 1 = Inf, then Inf * (Inf != Inf) = +0.0
    #  2       -> 2
    #  NaN     -> NaN * false = +-0, times sign(NaN) = NaN -> NaN survives both clamps
    #  -Inf    -> -0.0 * -1 = +0.0
    #  1e-15   -> kept (the Forward Euler clamp at 1f-10 would remove it)
    #  -0.0    -> +0.0 + -0.0 = +0.0 -> +0.0 * false = +0.0, sign(+0.0) = +0.0 -> +0.0
    #  1e-300  -> +0.0
    expect_f=[0.0, 0.0, THR20, 0.0, 0.0, 2.0, NAN, 0.0, 1e-15, 0.0, 0.0],
    expect=dict(cr_evaluated=0, cr=0.0),
)

# Forward Euler clamps f * (f >= 1f-10) then f * (f != Inf) (src/SteppingMethods.jl:97-101) on the
# state itself: with no operators df = -(+0.0) = -0.0 and f + (df + 0.0) = f (except -0.0 + 0.0).
EULER_CLAMP_F = [-1.0, 5e-11, THR10, float(np.nextafter(THR10, 0.0)),
                 float(np.nextafter(THR10, 1.0)), INF, 2.0, NAN, -INF, -0.0, 1e-300]
EULER_CLAMP = HandCase(
    name="euler_clamp", method=0, f=EULER_CLAMP_F,
    #  -1 -> -1 * false = -0.0 (kept by the Inf clamp);  5e-11 -> +0.0;  1f-10 kept;
    #  1f-10 - 1 ulp -> +0.0;  1f-10 + 1 ulp kept;  Inf -> kept, then Inf * false = +0.0;
    #  2 kept;  NaN -> NaN * false = +-0 (the NaN's sign);  -Inf -> -0.0;
    #  -0.0 -> -0.0 + (-0.0 + 0.0) = +0.0 -> +0.0;  1e-300 -> +0.0
    # sum(f) = NaN != 0, so Cr is evaluated: df / f = -0.0 / f has no value below -0.0
    # (NaN -> Inf), so Cr = -(-0.0) = +0.0; sum(df) = -0.0 + ... is finite
    expect_f=[-0.0, 0.0, THR10, 0.0, float(np.nextafter(THR10, 1.0)), 0.0, 2.0, 0.0, -0.0, 0.0,
              0.0],
    nan_sign_free=[7],
    expect=dict(finite=1, cr_evaluated=1, cr=0.0, cr_bin=0.0, cr_emi=0.0, cr_flux=0.0),
    expect_v1=dict(finite=1, cr_evaluated=1, cr=0.0),
)

CASES = [EULER_BINARY, EULER_NONFINITE, EULER_CLAMP, SYMPLECTIC, SYMPLECTIC_CLAMP]


def same_bits(a: float, b: float) -> bool:
    """Equal including the sign of zero; NaN matches NaN."""
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b and math.copysign(1.0, a) == math.copysign(1.0, b)


def check_state(case: HandCase, f_out) -> None:
    for i, (got, want) in enumerate(zip(list(f_out), case.expect_f)):
        if i in case.nan_sign_free:
            assert got == 0.0, (case.name, i, got)
        else:
            assert same_bits(got, want), (case.name, i, got, want)


def check_stats(case: HandCase, st, verbose: int) -> None:
    """st: an object with the stats as attributes or a dict."""
    want = case.expect if verbose == 3 else case.expect_v1
    get = (lambda k: st[k]) if isinstance(st, dict) else (lambda k: getattr(st, k))
    for k, v in want.items():
        got = get(k)
        if isinstance(v, float):
            assert same_bits(float(got), v), (case.name, verbose, k, got, v)
        else:
            assert int(got) == v, (case.name, verbose, k, got, v)
