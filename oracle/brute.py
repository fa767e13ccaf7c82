"""Pure-Python brute-force restatement of the explicit-Euler step, for tiny inputs only.

TEST INFRASTRUCTURE ONLY.  An independent second implementation used to pin the C oracle:
plain Python floats (IEEE binary64, no fused multiply-add), explicit loops, written directly
from the reference's lines:
  - ForwardEuler functor        src/SteppingMethods.jl:12-107
  - update_Big_Bin!             src/SteppingMethods.jl:332-397
  - ForwardSymplecticEuler      src/SteppingMethods.jl:120-243
  - Solve                       src/Solver.jl:1-64
"""
from __future__ import annotations

import math
import struct


def _float32(x: float) -> float:
    """The Float32 literal x promoted exactly to Float64 (round to nearest binary32)."""
    return struct.unpack("<f", struct.pack("<f", x))[0]


THR_1F_10 = _float32(1e-10)  # Float64(1f-10) = 1.00000001335143196e-10
THR_1F_20 = _float32(1e-20)  # Float64(1f-20) = 9.999999682655225e-21


def bool_mul(x: float, cond: bool) -> float:
    """Julia's `x * cond` with a Bool: strong zero."""
    return x if cond else math.copysign(0.0, x)


def jl_sign(x: float) -> float:
    if x > 0:
        return 1.0
    if x < 0:
        return -1.0
    return x


def csc_matvec(csc, x):
    colptr, rowval, nzval = csc
    n = len(colptr) - 1
    y = [0.0] * len(x)
    for j in range(n):
        for k in range(int(colptr[j]) - 1, int(colptr[j + 1]) - 1):
            r = int(rowval[k]) - 1
            y[r] = y[r] + float(nzval[k]) * x[j]
    return y


def dense_matvec(M, x, rows, cols):
    """M indexed M[r + rows*c] (column-major flat list); per-row order ascending column."""
    y = [0.0] * rows
    for c in range(cols):
        xc = x[c]
        for r in range(rows):
            y[r] = y[r] + M[r + rows * c] * xc
    return y


def binary_term_two_gemv(Mflat, f, N, C, Vol, domain=None, prev=None):
    """update_Big_Bin!: per cell T = M_Bin * fc (N^2), dB = reshape(T, N, N) * fc, times Vol."""
    dB = list(prev) if prev is not None else [0.0] * (N * C)
    for off in range(C):
        fc = f[off * N:(off + 1) * N]
        all_zero = all(v == 0.0 for v in fc)
        if not (domain is None or off in domain or all_zero):
            continue
        T = dense_matvec(Mflat, fc, N * N, N)
        d = dense_matvec(T, fc, N, N)
        for k in range(N):
            dB[off * N + k] = d[k] * Vol[off]
    return dB


def binary_term_triple(M3, f, N, C, Vol):
    """Closed form dB[k] = Vol * sum_{i,j} M[k,i,j] f_i f_j (different summation order)."""
    dB = [0.0] * (N * C)
    for off in range(C):
        fc = f[off * N:(off + 1) * N]
        for k in range(N):
            s = 0.0
            for i in range(N):
                for j in range(N):
                    s += M3[k][i][j] * fc[i] * fc[j]
            dB[off * N + k] = s * Vol[off]
    return dB


def div(a: float, b: float) -> float:
    """IEEE a / b (Python raises on a zero divisor)."""
    try:
        return a / b
    except ZeroDivisionError:
        if a == 0.0 or math.isnan(a):
            return math.nan
        return math.copysign(math.inf, a) * math.copysign(1.0, b)


def seq_sum(xs) -> float:
    """Left-to-right sum from 0.0 (Python's sum() starts from the integer 0)."""
    t = 0.0
    for x in xs:
        t = t + x
    return t


def neg_min_ratio(num, den):
    """-mapreduce(x -> isnan(x) ? Inf : x, min, num ./ den), with Julia's min(-0.0, 0.0) = -0.0."""
    m = math.inf
    for a, b in zip(num, den):
        r = div(a, b)
        if math.isnan(r):
            r = math.inf
        if r < m or (r == m and math.copysign(1.0, r) < 0):
            m = r
    return -m


def forward_euler(state, t_start, t_stop, dt, verbose):
    """One ForwardEuler functor call on a dict state (f, df_Bin, operators); returns
    (dt, save, stats)."""
    if t_start + dt >= t_stop:
        dt = t_stop - t_start
        save = True
    else:
        save = False
    s = dt / state["dt0"]
    f = state["f"]
    n = len(f)
    dF = csc_matvec(state["F"], f) if state.get("F") is not None else [0.0] * n
    df = [-v for v in dF]
    dE = None
    if state.get("M_Emi") is not None:
        dE = dense_matvec(state["M_Emi"], f, n, n)
        df = [a + b for a, b in zip(df, dE)]
    if state.get("M_Bin") is not None:
        state["df_Bin"] = binary_term_two_gemv(state["M_Bin"], f, state["N"], state["C"],
                                               state["Vol"], state.get("domain"),
                                               state["df_Bin"])
        df = [a + b for a, b in zip(df, state["df_Bin"])]
    invAp = state["invAp"]
    df = [a * (ia * s) for a, ia in zip(df, invAp)]
    # src/SteppingMethods.jl:48-94: the stats start at 0.0 and are only set when sum(f) != 0
    stats = {"finite": math.isfinite(seq_sum(df)), "cr": 0.0, "cr_bin": 0.0, "cr_emi": 0.0,
             "cr_flux": 0.0, "cr_evaluated": False}
    if verbose in (1, 2, 3) and seq_sum(f) != 0.0:
        stats["cr_evaluated"] = True
        if verbose == 3:
            if state.get("M_Bin") is not None:
                stats["cr_bin"] = neg_min_ratio(state["df_Bin"], f)
            if dE is not None:
                stats["cr_emi"] = neg_min_ratio(dE, f)
            stats["cr_flux"] = neg_min_ratio([-v for v in dF], f)
        stats["cr"] = neg_min_ratio(df, f)
    inj = state["df_inj"]
    out, pre = [], []
    for i in range(n):
        x = f[i] + (df[i] + inj[i] * s)
        pre.append(x)
        x = bool_mul(x, x >= THR_1F_10)
        x = bool_mul(x, x != math.inf)
        out.append(x)
    state["f"] = out
    state["pre"] = pre
    return dt, save, stats


def forward_symplectic_euler(state, t_start, t_stop, dt, verbose):
    if t_start + dt >= t_stop:
        dt = t_stop - t_start
        save = True
    else:
        save = False
    s = dt / state["dt0"]
    f = state["f"]
    n = len(f)
    dP = csc_matvec(state["P"], f) if state.get("P") is not None else [0.0] * n
    dm = [-v for v in dP]
    dE = None
    if state.get("M_Emi") is not None:
        dE = dense_matvec(state["M_Emi"], f, n, n)
        dm = [a + b for a, b in zip(dm, dE)]
    if state.get("M_Bin") is not None:
        state["df_Bin"] = binary_term_two_gemv(state["M_Bin"], f, state["N"], state["C"],
                                               state["Vol"], state.get("domain"),
                                               state["df_Bin"])
        dm = [a + b for a, b in zip(dm, state["df_Bin"])]
    invAp = state["invAp"]
    dm = [a * (ia * s) for a, ia in zip(dm, invAp)]
    fm = []
    for i in range(n):
        x = f[i] + dm[i]
        fm.append(bool_mul(x, x >= THR_1F_10))
    dX = csc_matvec(state["X"], fm) if state.get("X") is not None else [0.0] * n
    dS = [((-a) * ia) * s for a, ia in zip(dX, invAp)]
    fs = []
    for i in range(n):
        x = fm[i] + dS[i]
        fs.append(bool_mul(x, x >= THR_1F_20))
    # src/SteppingMethods.jl:175-229
    stats = {"finite": True, "cr": 0.0, "cr_bin": 0.0, "cr_emi": 0.0, "cr_flux": 0.0,
             "cr_momentum": 0.0, "cr_space": 0.0, "cr_evaluated": False}
    if verbose in (1, 2, 3) and seq_sum(f) != 0.0:
        stats["cr_evaluated"] = True
        if verbose == 3:
            if state.get("M_Bin") is not None:
                stats["cr_bin"] = neg_min_ratio(
                    [(b * ia) * s for b, ia in zip(state["df_Bin"], invAp)], f)
            if dE is not None:
                stats["cr_emi"] = neg_min_ratio([(e * ia) * s for e, ia in zip(dE, invAp)], f)
            # Cr_PFlux: the ratio's NaNs are replaced by 0.0 before the reduction
            ratio = [div(((-p) * ia) * s, x) for p, ia, x in zip(dP, invAp, f)]
            ratio = [0.0 if math.isnan(r) else r for r in ratio]
            stats["cr_flux"] = neg_min_ratio(ratio, [1.0] * n)
            stats["cr_momentum"] = neg_min_ratio(dm, f)
            stats["cr_space"] = neg_min_ratio(dS, fm)
        stats["cr"] = neg_min_ratio([a + b for a, b in zip(dS, dm)], f)
    inj = state["df_inj"]
    out, pre = [], []
    for i in range(n):
        x = fs[i] + inj[i] * s
        pre.append(x)
        x = bool_mul(x, x >= THR_1F_20) * jl_sign(x)
        x = bool_mul(x, x != math.inf)
        out.append(x)
    state["f"] = out
    state["pre"] = pre
    return dt, save, stats


def solve(state, f_init, dt_initial, t_save, verbose=1, method=0):
    """Solve (src/Solver.jl:1-64): returns the list of saved states and the dt sequence."""
    step = forward_euler if method == 0 else forward_symplectic_euler
    out = [list(f_init)]
    state["f"] = list(f_init)
    dt = dt_initial
    dts = []
    for i in range(1, len(t_save)):
        t_start, t_stop, save = t_save[i - 1], t_save[i], False
        while not save:
            dt, save, _ = step(state, t_start, t_stop, dt, verbose)
            dts.append(dt)
            t_start += dt
        out.append(list(state["f"]))
    return out, dts


def make_state(w, mono_flat=None):
    """Brute-force state for a dtx.workloads.Workload (tiny sizes)."""
    N, C, n = w.ps.n_momentum, w.ps.n_space, w.ps.n
    st = {"dt0": w.ps.dt0, "N": N, "C": C, "Vol": [float(v) for v in w.Vol],
          "invAp": [1.0 / float(a) for a in w.Ap], "df_inj": [float(v) for v in w.inj],
          "df_Bin": [0.0] * n, "domain": set(w.domain) if w.domain is not None else None,
          "F": w.F, "X": w.X, "P": w.P}
    if mono_flat is None and w.M_Bin is not None:
        mono_flat = [float(v) for v in w.M_Bin.reshape(-1, order="F")]
    st["M_Bin"] = mono_flat
    from dtx.workloads import emission_dense
    E = emission_dense(w)
    st["M_Emi"] = [float(v) for v in E.reshape(-1, order="F")] if E is not None else None
    return st
