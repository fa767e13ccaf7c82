"""ctypes wrapper of the FP64 CPU oracle (libdtx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product (libdtx.so and the dtx package) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libdtx_oracle.so"

I32, I64, D = C.c_int32, C.c_int64, C.c_double
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)


class Csc(C.Structure):
    _fields_ = [("ncols", I64), ("colptr", PI64), ("rowval", PI64), ("nzval", PD)]


class Gain(C.Structure):
    _fields_ = [("s_out", I32), ("s1", I32), ("s2", I32), ("G", PD)]


class Loss(C.Structure):
    _fields_ = [("s1", I32), ("s2", I32), ("L", PD)]


class Problem(C.Structure):
    _fields_ = [("N", I64), ("C", I64), ("dt0", D), ("num_species", I32),
                ("species_offset", I64 * 16), ("species_size", I64 * 16),
                ("F", Csc), ("X", Csc), ("P", Csc), ("M_Emi", PD), ("M_Bin", PD),
                ("n_gain", I32), ("n_loss", I32), ("gain", C.POINTER(Gain)),
                ("loss", C.POINTER(Loss)), ("domain", PI64), ("n_domain", I64),
                ("invAp", PD), ("Vol", PD), ("df_inj", PD), ("batched", I32), ("Emi", Csc)]


class Stats(C.Structure):
    _fields_ = [("finite", C.c_int), ("cr_evaluated", C.c_int), ("cr", D), ("cr_bin", D),
                ("cr_emi", D), ("cr_flux", D), ("dt", D), ("t_start", D), ("save", C.c_int),
                ("cr_momentum", D), ("cr_space", D)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.dtxo_ws_create.restype = C.c_void_p
        lib.dtxo_ws_create.argtypes = [C.POINTER(Problem)]
        lib.dtxo_ws_destroy.argtypes = [C.c_void_p]
        lib.dtxo_ws_f.restype = PD
        lib.dtxo_ws_f.argtypes = [C.c_void_p]
        lib.dtxo_ws_pre_clamp.restype = PD
        lib.dtxo_ws_pre_clamp.argtypes = [C.c_void_p]
        lib.dtxo_ws_df_bin.restype = PD
        lib.dtxo_ws_df_bin.argtypes = [C.c_void_p]
        lib.dtxo_forward_euler.argtypes = [C.POINTER(Problem), C.c_void_p, D, D, D, C.c_int,
                                           PD, C.POINTER(C.c_int), C.POINTER(Stats)]
        lib.dtxo_forward_symplectic_euler.argtypes = lib.dtxo_forward_euler.argtypes
        lib.dtxo_solve.argtypes = [C.POINTER(Problem), C.c_int, PD, D, PD, I64, C.c_int, PD,
                                   PD, C.POINTER(Stats), I64, PI64]
        lib.dtxo_binary_term.argtypes = [C.POINTER(Problem), PD, PD]
        lib.dtxo_moments.argtypes = [PD, I64, I64, I64, PD, D, PD, PD]
        lib.dtxo_num_threads.restype = C.c_int
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(PD) if a is not None else None


def _csc(keep, m):
    if m is None:
        return Csc(0, None, None, None)
    colptr, rowval, nzval = (np.ascontiguousarray(m[0], np.int64),
                             np.ascontiguousarray(m[1], np.int64),
                             np.ascontiguousarray(m[2], np.float64))
    keep += [colptr, rowval, nzval]
    return Csc(colptr.size - 1, colptr.ctypes.data_as(PI64), rowval.ctypes.data_as(PI64),
               nzval.ctypes.data_as(PD))


class OracleProblem:
    """Inputs of the oracle; mirrors ForwardEulerStruct's fields (src/SteppingStructs.jl:2-31).
    M_Bin: array whose Fortran-order memory is the (N^2 x N) tensor, or gains/losses (channel
    form).  F / X / P: (colptr, rowval, nzval) 1-based CSC.  M_Emi: dense n x n."""

    def __init__(self, ps, Ap, Vol, df_inj, F=None, X=None, P=None, M_Emi=None, M_Bin=None,
                 gains=(), losses=(), domain=None, batched=False, M_Emi_csc=None):
        self.ps = ps
        self.keep = []
        p = Problem()
        p.N, p.C, p.dt0 = ps.n_momentum, ps.n_space, ps.dt0
        p.num_species = ps.num_species
        for i, (off, s) in enumerate(zip(ps.species_offset, ps.species)):
            p.species_offset[i] = off
            p.species_size[i] = s.size
        p.F, p.X, p.P = _csc(self.keep, F), _csc(self.keep, X), _csc(self.keep, P)
        if M_Emi is not None:
            me = np.asfortranarray(M_Emi, dtype=np.float64)
            self.keep.append(me)
            p.M_Emi = _dp(me)
        if M_Bin is not None:
            mb = np.asfortranarray(M_Bin, dtype=np.float64)
            self.keep.append(mb)
            p.M_Bin = _dp(mb)
        garr = (Gain * max(1, len(gains)))()
        for k, g in enumerate(gains):
            G = np.asfortranarray(g.G, dtype=np.float64)
            self.keep.append(G)
            garr[k] = Gain(g.s_out, g.s1, g.s2, _dp(G))
        larr = (Loss * max(1, len(losses)))()
        for k, l in enumerate(losses):
            L = np.asfortranarray(l.L, dtype=np.float64)
            self.keep.append(L)
            larr[k] = Loss(l.s1, l.s2, _dp(L))
        self.keep += [garr, larr]
        p.n_gain, p.n_loss = len(gains), len(losses)
        p.gain, p.loss = garr, larr
        if domain is None:
            p.n_domain = -1
        else:
            dom = np.ascontiguousarray(domain, dtype=np.int64)
            self.keep.append(dom)
            p.domain = dom.ctypes.data_as(PI64)
            p.n_domain = dom.size
        invAp = 1.0 / np.asarray(Ap, dtype=np.float64)  # 1 ./ Ap_Flux (src/SteppingStructs.jl:55)
        vol = np.ascontiguousarray(Vol, dtype=np.float64)
        inj = np.ascontiguousarray(df_inj, dtype=np.float64)
        self.keep += [invAp, vol, inj]
        p.invAp, p.Vol, p.df_inj = _dp(invAp), _dp(vol), _dp(inj)
        p.batched = 1 if batched else 0
        p.Emi = _csc(self.keep, M_Emi_csc)
        self.p = p

    def binary_term(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.empty_like(f)
        load().dtxo_binary_term(C.byref(self.p), _dp(f), _dp(out))
        return out

    def solve(self, f_init, dt_initial, t_save, verbose=1, method=0, max_stats=100000):
        lib = load()
        n = self.ps.n
        t_save = np.ascontiguousarray(t_save, dtype=np.float64)
        f_init = np.ascontiguousarray(f_init, dtype=np.float64)
        fout = np.empty((t_save.size, n))
        tout = np.empty(t_save.size)
        stats = (Stats * max_stats)()
        nsteps = C.c_int64()
        rc = lib.dtxo_solve(C.byref(self.p), method, _dp(f_init), float(dt_initial),
                            _dp(t_save), t_save.size, verbose, _dp(fout), _dp(tout), stats,
                            max_stats, C.byref(nsteps))
        if rc != 0:
            raise ValueError("oracle Solve: the carried dt would need more than 1e8 steps")
        return fout, tout, list(stats)[:min(nsteps.value, max_stats)], nsteps.value


class Stepper:
    """Functor-style use of the oracle (one call = one reference functor call)."""

    def __init__(self, problem: OracleProblem, f_init, method=0):
        self.problem = problem
        self.lib = load()
        self.ws = self.lib.dtxo_ws_create(C.byref(problem.p))
        self.method = method
        self.set_f(f_init)

    def set_f(self, f):
        n = self.problem.ps.n
        ptr = self.lib.dtxo_ws_f(self.ws)
        np.ctypeslib.as_array(ptr, shape=(n,))[:] = f

    @property
    def f(self):
        n = self.problem.ps.n
        return np.ctypeslib.as_array(self.lib.dtxo_ws_f(self.ws), shape=(n,)).copy()

    @property
    def pre_clamp(self):
        """The last step's state before its final clamps."""
        n = self.problem.ps.n
        return np.ctypeslib.as_array(self.lib.dtxo_ws_pre_clamp(self.ws), shape=(n,)).copy()

    @property
    def df_bin(self):
        n = self.problem.ps.n
        return np.ctypeslib.as_array(self.lib.dtxo_ws_df_bin(self.ws), shape=(n,)).copy()

    def __call__(self, t_start, t_stop, dt, verbose):
        dt_out = C.c_double()
        save = C.c_int()
        st = Stats()
        fn = (self.lib.dtxo_forward_euler if self.method == 0
              else self.lib.dtxo_forward_symplectic_euler)
        fn(C.byref(self.problem.p), self.ws, float(t_start), float(t_stop), float(dt),
           int(verbose), C.byref(dt_out), C.byref(save), C.byref(st))
        return dt_out.value, bool(save.value), st

    def __del__(self):
        try:
            self.lib.dtxo_ws_destroy(self.ws)
        except Exception:
            pass


def moments(f_block, s, m=None):
    """N^0 and T^00 of one (cell, species) block, as StressEnergyTensor sums them."""
    lib = load()
    f_block = np.ascontiguousarray(f_block, dtype=np.float64)
    pr = np.ascontiguousarray(s.pr, dtype=np.float64)
    num, en = C.c_double(), C.c_double()
    lib.dtxo_moments(_dp(f_block), s.p_num, s.u_num, s.h_num, _dp(pr),
                     float(s.mass if m is None else m), C.byref(num), C.byref(en))
    return num.value, en.value


def num_threads():
    return int(load().dtxo_num_threads())


def set_threads(n: int):
    os.environ["OMP_NUM_THREADS"] = str(n)


def problem_from_workload(w, use_monolithic=None, batched=False, sparse_emission=False):
    """The oracle problem for a dtx.workloads.Workload (same arrays the HIP path receives).
    batched: channel form evaluated for many cells at once (bitwise the same result).
    sparse_emission: M_Emi as the CSC of its per-cell blocks instead of the dense n x n."""
    from dtx.workloads import emission_csc, emission_dense
    mono = (w.M_Bin is not None) if use_monolithic is None else use_monolithic
    kw = {"batched": batched}
    if sparse_emission:
        kw["M_Emi_csc"] = emission_csc(w)
    if mono:
        kw["M_Bin"] = w.M_Bin
    else:
        kw["gains"], kw["losses"] = w.gains, w.losses
    return OracleProblem(w.ps, w.Ap, w.Vol, w.inj, F=w.F, X=w.X, P=w.P,
                         M_Emi=None if sparse_emission else emission_dense(w), domain=w.domain,
                         **kw)
