"""Serving: many simulations stepped together on one GPU (continuous batching).

Thin wrapper of the dtx_server_* C ABI.  The scheduler runs inside libdtx.so: requests are
admitted FIFO when a slot is free and their HBM footprint fits the budget, all admitted requests
advance together through grouped contraction/update launches, and finished requests are retired
at once so waiting ones take their place.  Each request's result does not depend on what it ran
with.

`request_mix` draws the serving workload: p bins in {16, 32, 64}, u bins in {4, 8}, one phi bin,
x cells in {1, 32, 128}, 1-2 species, steps uniform in [200, 2000], with Poisson arrival times.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .synth import config_seed, stream_key, uniform

STATUS = {0: "queued", 1: "running", 2: "done"}


@dataclass
class RequestSpec:
    p_num: int
    u_num: int
    h_num: int
    x_num: int
    species: int
    steps: int
    seed: int
    theta: float = 1.0
    with_flux: bool = True
    arrival_s: float = 0.0

    @property
    def states(self) -> int:
        return self.species * self.p_num * self.u_num * self.h_num

    @property
    def n(self) -> int:
        return self.states * self.x_num

    def to_c(self) -> _lib.Request:
        r = _lib.Request()
        cfg = r.cfg
        cfg.num_species = self.species
        for s in range(self.species):
            cfg.px_num[s], cfg.py_num[s], cfg.pz_num[s] = self.p_num, self.u_num, self.h_num
        cfg.x_num, cfg.y_num, cfg.z_num = self.x_num, 1, 1
        cfg.dt0 = 1.0
        r.seed = self.seed
        r.steps = self.steps
        r.theta = self.theta
        r.with_flux = int(self.with_flux)
        return r


def request_mix(count: int = 256, rate_per_s: float = 1.0, seed: int | None = None,
                steps_range=(200, 2000)):
    """Seeded serving workload with Poisson arrivals at `rate_per_s`."""
    seed = config_seed(3) if seed is None else seed
    key = stream_key(seed, 0)
    u = uniform(np.arange(8 * count, dtype=np.uint64), key).reshape(count, 8)
    out, t = [], 0.0
    for i in range(count):
        t += -np.log(u[i, 0]) / rate_per_s
        out.append(RequestSpec(
            p_num=[16, 32, 64][int(u[i, 1] * 3)], u_num=[4, 8][int(u[i, 2] * 2)], h_num=1,
            x_num=[1, 32, 128][int(u[i, 3] * 3)], species=1 + int(u[i, 4] * 2),
            steps=steps_range[0] + int(u[i, 5] * (steps_range[1] - steps_range[0] + 1)),
            seed=config_seed(3, i), arrival_s=t))
    return out


def server_config(precision="fp32", hbm_budget_gb: float = 200.0, max_active: int = 256,
                  device: int = 0, moment_correction: bool = True) -> _lib.ServerConfig:
    cfg = _lib.ServerConfig()
    cfg.device = device
    cfg.precision = _lib.PRECISIONS[precision]
    cfg.hbm_budget = int(hbm_budget_gb * 1e9)
    cfg.max_active = max_active
    cfg.no_moment_correction = 0 if moment_correction else 1
    return cfg


class RequestHandle:
    """The simulation the server builds for a synthetic request (dtx_request_create), as a
    stand-alone handle: step it with `step` (dtx_step, t_stop never reached) to get the served
    result; or give it to Server.submit_handle."""

    def __init__(self, spec: "RequestSpec", precision="fp32", moment_correction=True, device=0):
        self.lib = _lib.load()
        self.spec = spec
        cfg = server_config(precision, device=device, moment_correction=moment_correction)
        req = spec.to_c()
        self._h = C.c_void_p()
        dt = C.c_double()
        rc = self.lib.dtx_request_create(C.byref(req), C.byref(cfg), C.byref(self._h),
                                         C.byref(dt))
        if rc != 0:
            raise _lib.DtxError(f"dtx_request_create failed ({rc}): "
                                f"{self.lib.dtx_last_error(None).decode()}")
        self.dt = dt.value

    def step(self, steps: int):
        out_dt, save = C.c_double(), C.c_int32()
        t = 0.0
        for _ in range(steps):
            _lib.check(self.lib.dtx_step(self._h, t, 1e300, self.dt, 0, C.byref(out_dt),
                                         C.byref(save), None), self._h, "dtx_step")
            t += self.dt

    @property
    def f(self) -> np.ndarray:
        out = np.empty(self.spec.n)
        _lib.check(self.lib.dtx_get_f(self._h, _lib.dptr(out)), self._h, "dtx_get_f")
        return out

    def release(self):
        """Give up ownership (after Server.submit_handle)."""
        h, self._h = self._h, C.c_void_p()
        return h

    def close(self):
        if self._h and self._h.value:
            self.lib.dtx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    def __init__(self, precision="fp32", hbm_budget_gb: float = 200.0, max_active: int = 256,
                 device: int = 0, moment_correction: bool = True):
        self.lib = _lib.load()
        cfg = server_config(precision, hbm_budget_gb, max_active, device, moment_correction)
        self._s = C.c_void_p()
        rc = self.lib.dtx_server_create(C.byref(cfg), C.byref(self._s))
        if rc != 0:
            raise _lib.DtxError(f"dtx_server_create failed ({rc})")

    def submit(self, spec: RequestSpec) -> int:
        rid = C.c_int64()
        req = spec.to_c()
        rc = self.lib.dtx_server_submit(self._s, C.byref(req), C.byref(rid))
        if rc != 0:
            raise _lib.DtxError(f"dtx_server_submit failed ({rc})")
        return rid.value

    def submit_handle(self, handle, steps: int, dt: float) -> int:
        """A caller-built ForwardEuler handle (the server takes ownership): a RequestHandle or
        a ForwardEulerStruct built with this server's precision."""
        h = handle.release() if hasattr(handle, "release") else handle
        rid = C.c_int64()
        rc = self.lib.dtx_server_submit_handle(self._s, h, int(steps), float(dt), C.byref(rid))
        if rc != 0:
            raise _lib.DtxError(f"dtx_server_submit_handle failed ({rc})")
        return rid.value

    def poll(self, rid: int) -> _lib.RequestResult:
        res = _lib.RequestResult()
        rc = self.lib.dtx_server_poll(self._s, rid, C.byref(res))
        if rc != 0:
            raise _lib.DtxError(f"dtx_server_poll failed ({rc})")
        return res

    def fetch(self, rid: int, n: int) -> np.ndarray:
        out = np.empty(n)
        rc = self.lib.dtx_server_fetch(self._s, rid, _lib.dptr(out), n)
        if rc != 0:
            raise _lib.DtxError(f"dtx_server_fetch failed ({rc})")
        return out

    def drain(self):
        rc = self.lib.dtx_server_drain(self._s)
        if rc != 0:
            raise _lib.DtxError("dtx_server_drain failed")

    def last_error(self) -> str:
        return self.lib.dtx_server_last_error(self._s).decode()

    def close(self):
        if self._s and self._s.value:
            self.lib.dtx_server_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(specs, precision="fp32", hbm_budget_gb=200.0, max_active=256, realtime=True,
        fetch=False, moment_correction=True):
    """Submit the requests at their arrival times (or all at once) and wait for all of them.
    Returns (per-request results, wall seconds, final states if fetched)."""
    srv = Server(precision, hbm_budget_gb, max_active, moment_correction=moment_correction)
    t0 = time.perf_counter()
    ids = []
    for s in specs:
        if realtime:
            delay = s.arrival_s - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        ids.append(srv.submit(s))
    srv.drain()
    wall = time.perf_counter() - t0
    results = [srv.poll(i) for i in ids]
    states = [srv.fetch(i, s.n) for i, s in zip(ids, specs)] if fetch else None
    err = srv.last_error()
    srv.close()
    bad = [r for r in results if r.status != 2]
    if bad:
        raise _lib.DtxError(f"{len(bad)} requests failed: {err}")
    return results, wall, states


def latency_stats(results):
    tts = np.array([r.finish_ms - r.submit_ms for r in results]) / 1e3
    return dict(p50_s=float(np.percentile(tts, 50)), p99_s=float(np.percentile(tts, 99)),
                mean_s=float(tts.mean()), max_s=float(tts.max()))
