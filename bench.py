#!/usr/bin/env python3
"""Benchmark of the explicit-Euler step on one MI355X.

Workload (default): 1D transport -- 2 species (electron, photon) x (32 log p bins on
[1e-3, 1e3] x 8 u bins x 1 phi bin) = 512 states per cell, 128 periodic x cells (65,536
unknowns), upwind x advection + synchrotron cooling, per-cell binary collisions with lognormal
gain tensors, plain FP64.  A step is one ForwardEuler functor call over the whole state; inputs
are resident in HBM before the timed region.

Prints ONE JSON line: steps/s of the FP64 path (value) over a timed window with no event
records in it; the kernel durations of the roofline come from HIP events the library records on
its own stream around every step's kernels in a separate, untimed pass.  Beside it: the FP32-MFMA
and split-BF16 paths (speed, and their relative error and number / energy drift against FP64
over the configuration's full step count), the same comparison at the largest configuration
(transport_2d, where the split-BF16 contraction is MFMA-bound), the emission GEMV's HBM rate
(transport_1d with dense per-cell emission blocks), the serving loop, the collision-matrix
ingest, and the FP64 CPU oracle timed on a bounded sample on the host cores.
"""
import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "diplodocustransport.jl_amd"))

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
PEAK_HBM_GBS = 8000.0
TIMING_STEPS = 40   # untimed steps with HIP events around every step's kernels

DESCRIPTIONS = {
    "transport_1d": "2 species x 256 states (32 p x 8 u x 1 phi), 128 periodic x cells, 65536 "
                    "unknowns, upwind x advection + synchrotron cooling, per-cell binary "
                    "collisions, plain FP64",
    "precision_box": "homogeneous box, 3 species x 1024 states (64 p x 8 u x 2 phi), 9 gain "
                     "channels of up to 1024^3 generated on the device, plain FP64",
    "transport_2d": "2 species x 512 states (32 p x 8 u x 2 phi), 32 x 32 periodic x-y cells, "
                    "1048576 unknowns, upwind x/y advection + synchrotron cooling, per-cell "
                    "binary collisions with a shared operator, plain FP64",
    "homogeneous_box": "homogeneous box, 2 species x 256 states (32 p x 8 u x 1 phi), "
                       "monolithic 512^2 x 512 FP64 binary tensor, plain FP64",
}

# HBM bytes per launch of the FP64 contraction from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (separate runs, each scaled by the same counter over a 1 GiB device copy):
# profiles/transport_1d_hbm_traffic.txt, profiles/precision_box_hbm_traffic.txt.
TRAFFIC = {("transport_1d", "bin_f64_wide_kernel"): 470_180_000,
           ("precision_box", "bin_gemv_kernel<0>"): 64_916_380_000}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def t_save_marks(w, marks):
    """Save times after the given step counts (exact multiples of the workload's dt)."""
    out, t, k = [0.0], 0.0, 0
    for m in marks[1:]:
        while k < m:
            t += w.dt
            k += 1
        out.append(t)
    return out


def species_totals(w, f, weights=None):
    import numpy as np
    N, C = w.ps.n_momentum, w.ps.n_space
    fc = np.asarray(f).reshape(C, N)
    if weights is not None:
        fc = fc * weights[None, :]
    return np.array([fc[:, o:o + s.size].sum() for o, s in zip(w.ps.species_offset,
                                                              w.ps.species)])


def run_gpu(w, precision, steps, warmup, verbose=1, moment_correction=True, accuracy_steps=0,
            device_flux=False):
    """When accuracy_steps > 0, first a Solve of that many steps from f0 whose final state the
    accuracy comparison uses (it also brings the GPU out of its idle clocks: after 0.2 s idle
    the first 5 ms of steps run about 20% slower); then from f0 the warmup steps, the timed
    window (no event records), and an untimed pass with HIP events around every step's kernels
    (their average durations)."""
    from dtx import workloads
    m = workloads.make_struct(w, precision=precision, use_monolithic=False,
                              moment_correction=moment_correction,
                              device_flux=device_flux and w.flux_spec is not None)
    work = m.work()
    acc = {}
    if accuracy_steps:
        out, n = m.solve_native(w.dt, t_save_marks(w, [0, accuracy_steps]), Verbose=0)
        assert n == accuracy_steps
        f = out.f[-1]
        acc = dict(f=f, number=species_totals(w, f),
                   energy=species_totals(w, f, w.ps.energy_weights()))
    m.reset()
    t, big = 0.0, 1e300
    for _ in range(warmup):
        dt, _ = m(t, big, w.dt, verbose)
        t += dt
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt, _ = m(t, big, w.dt, verbose)
        t += dt
    m.synchronize()
    t1 = time.perf_counter()
    m.kernel_timing(enable=True, reset=True, every=1)
    for _ in range(TIMING_STEPS):
        dt, _ = m(t, big, w.dt, verbose)
        t += dt
    kt = m.kernel_timing(enable=False, reset=True)
    res = dict(ms_per_step=(t1 - t0) * 1e3 / steps, steps_per_s=steps / (t1 - t0),
               bin_ms=kt.bin_ms / max(1, kt.bin_launches),
               update_ms=kt.update_ms / max(1, kt.update_launches),
               emission_ms=kt.emission_ms / max(1, kt.emission_launches),
               emission_bytes=work.emission_bytes,
               bin_flops=work.bin_flops, gain_bytes=work.gain_bytes, bin_bytes=work.bin_bytes,
               update_bytes=work.update_bytes, items=work.bin_launch_items,
               device_gb=m.device_bytes() / 1e9, **acc)
    m.close()
    return res


def accuracy(r, base):
    """Relative error in f (states above 1e-8 of the largest) and the largest relative number
    and energy drift per species, against the FP64 run of the same steps."""
    import numpy as np
    fp = base["f"]
    mask = np.abs(fp) > 1e-8 * np.max(np.abs(fp))
    return dict(rel_linf_f=float(np.max(np.abs(r["f"][mask] - fp[mask]) / np.abs(fp[mask]))),
                number_drift=float(np.max(np.abs(r["number"] - base["number"]) /
                                          np.abs(base["number"]))),
                energy_drift=float(np.max(np.abs(r["energy"] - base["energy"]) /
                                          np.abs(base["energy"]))))


def run_largest(steps, warmup):
    """transport_2d (1,048,576 unknowns, 500 steps): FP64 against FP32 and split-BF16, speed over
    a timed window and accuracy over the full 500 steps (flux operators built on the device,
    which the tests show equal to the host CSC)."""
    from dtx import workloads
    w = workloads.transport_2d()
    out = {"workload": f"transport_2d: {DESCRIPTIONS['transport_2d']}",
           "accuracy_steps": w.steps, "timed_steps": steps}
    res = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        log(f"[bench] transport_2d {prec}: {warmup} warmup + {steps} timed steps, "
            f"{w.steps}-step accuracy run")
        res[prec] = run_gpu(w, prec, steps, warmup, accuracy_steps=w.steps, device_flux=True)
        log(f"[bench] transport_2d {prec}: {res[prec]['steps_per_s']:.1f} steps/s")
    base = res["fp64"]
    out["fp64_steps_per_s"] = round(base["steps_per_s"], 2)
    out["fp64_binary_kernel_us"] = round(base["bin_ms"] * 1e3, 1)
    for prec in ("fp32", "bf16x3"):
        r = res[prec]
        out[prec] = dict(steps_per_s=round(r["steps_per_s"], 2),
                         speedup_vs_fp64=round(r["steps_per_s"] / base["steps_per_s"], 3),
                         binary_kernel_us=round(r["bin_ms"] * 1e3, 1), **accuracy(r, base))
    return out


def run_emission(steps, warmup):
    """transport_1d with dense per-cell emission blocks (128 x 512^2 FP64, 268 MB streamed per
    step): the FP64 step rate and the emission GEMV's HBM rate against the 8 TB/s peak."""
    from dtx import workloads
    w = workloads.transport_1d(emission=True)
    log(f"[bench] transport_1d + emission: {warmup} warmup + {steps} timed steps")
    r = run_gpu(w, "fp64", steps, warmup)
    gbs = r["emission_bytes"] / (r["emission_ms"] * 1e-3) / 1e9
    log(f"[bench] emission GEMV {r['emission_ms'] * 1e3:.1f} us, {gbs:.0f} GB/s")
    return dict(workload="transport_1d with dense per-cell emission blocks (128 cells x "
                         "512 x 512 FP64)",
                steps_per_s=round(r["steps_per_s"], 2),
                kernel="emission_gemv_kernel<2, 2>",
                kernel_avg_us=round(r["emission_ms"] * 1e3, 2),
                bytes_per_launch=r["emission_bytes"],
                achieved_gbs=round(gbs, 1), peak_gbs=PEAK_HBM_GBS,
                frac=round(gbs / PEAK_HBM_GBS, 4),
                update_kernel_us=round(r["update_ms"] * 1e3, 2))


def fp64_kernels_per_step(w):
    """Launches of one plain FP64 ForwardEuler step: the contraction and the update, whose
    launch also finalises the previous step when the block partials fold in one workgroup (at
    most 16 groups of 256); larger grids add the two finalisation launches."""
    C, N, S = w.ps.n_space, w.ps.n_momentum, len(w.ps.species)
    nseg = (N + 63) // 64
    groups = (max(C * nseg, C * S) + 255) // 256
    return 2 if groups <= 16 else 4


def run_cpu_baseline(w, sample_steps):
    """The oracle in reference structure (per cell two GEMVs on the monolithic (N^2 x N) FP64
    tensor, CSC SpMV, elementwise passes) on the host cores, over a bounded number of steps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import oracle as O
    from dtx import synth
    M = synth.assemble_M_Bin(w.ps, w.gains, w.losses)
    prob = O.OracleProblem(w.ps, w.Ap, w.Vol, w.inj, F=w.F, M_Bin=M)
    st = O.Stepper(prob, w.f0)
    times, t = [], 0.0
    for _ in range(sample_steps):
        t0 = time.perf_counter()
        dt, _, _ = st(t, 1e300, w.dt, 1)
        times.append(time.perf_counter() - t0)
        t += dt
    med = float(np.median(times))
    return dict(value=1.0 / med, unit="steps/s", cores=O.num_threads(), kind="port",
                sample=f"{sample_steps} ForwardEuler steps of the same workload, reference "
                       f"structure (per-cell dgemv pair on the monolithic {w.ps.n_momentum}^2 x "
                       f"{w.ps.n_momentum} FP64 tensor), median step {med:.3f} s on "
                       f"{O.num_threads()} threads of {cpu_model()}")


def run_cpu_optimized(w, sample_steps):
    """A second, optimised CPU column: the same ForwardEuler step with the binary term batched
    over cells (one multithreaded DGEMM of the monolithic (N^2 x N) FP64 tensor against the f of
    all cells, then one batched contraction), the flux as a scipy CSC SpMV and vectorised
    elementwise passes.  Checked against one oracle step from the same state."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import scipy.sparse as sp
    import oracle as O
    from dtx import synth
    N, C, n = w.ps.n_momentum, w.ps.n_space, w.ps.n
    M = synth.assemble_M_Bin(w.ps, w.gains, w.losses)
    Mf = np.asfortranarray(M, dtype=np.float64).reshape(N * N, N, order="F")
    F = None
    if w.F is not None:
        colptr, rowval, nzval = w.F
        F = sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))
    inv_ap = 1.0 / w.Ap
    vol = np.repeat(np.asarray(w.Vol, dtype=np.float64), N)
    dt0 = w.ps.dt0

    def step(f, dt):
        s = dt / dt0
        df = -(F @ f) if F is not None else np.zeros(n)
        fm = f.reshape(C, N).T                                  # (N, C), cells as columns
        T = Mf @ fm                                              # (N^2, C): one DGEMM
        dB = np.einsum("kic,ic->kc", T.reshape(N, N, C, order="F"), fm)
        df = df + dB.T.reshape(-1) * vol
        df = df * (inv_ap * s)
        nz = f != 0.0
        cr = -np.min(df[nz] / f[nz]) if nz.any() else 0.0        # Cr at Verbose 1
        x = f + (df + w.inj * s)
        x = np.where(x >= np.float32(1e-10), x, 0.0)
        x = np.where(x != np.inf, x, 0.0)
        return x, cr

    # one step against the oracle from the same state
    prob = O.OracleProblem(w.ps, w.Ap, w.Vol, w.inj, F=w.F, M_Bin=M)
    st = O.Stepper(prob, w.f0)
    st(0.0, 1e300, w.dt, 0)
    f1, _ = step(np.array(w.f0, dtype=np.float64), w.dt)
    ref = st.f
    rel = float(np.max(np.abs(f1 - ref)) / max(np.max(np.abs(ref)), 1e-300))
    del st, prob
    f, times = np.array(w.f0, dtype=np.float64), []
    for _ in range(sample_steps):
        t0 = time.perf_counter()
        f, _ = step(f, w.dt)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=round(1.0 / med, 3), unit="steps/s", cores=threads, kind="port",
                rel_linf_vs_oracle_step=rel,
                sample=f"{sample_steps} ForwardEuler steps, binary term as one DGEMM over all "
                       f"{C} cells (numpy BLAS) + scipy CSC SpMV, median step {med * 1e3:.1f} ms "
                       f"on {threads} threads of {cpu_model()}")


def request_work(spec):
    """Math FLOPs and packed FP32 gain-operator bytes of one grouped step of a request: per
    interaction a + b -> a + b (a <= b) the gain channels of its output species, 2 No pairs C
    FLOPs and No pairs 4 bytes each, and per species the FP64 loss GEMV, 2 Ns N C FLOPs."""
    Ns = spec.p_num * spec.u_num * spec.h_num
    S, C = spec.species, spec.x_num
    N = S * Ns
    flops = bytes_ = 0.0
    for a in range(S):
        for b in range(a, S):
            pairs = Ns * (Ns + 1) // 2 if a == b else Ns * Ns
            outs = 1 if a == b else 2
            flops += outs * 2.0 * Ns * pairs * C
            bytes_ += outs * 4.0 * Ns * pairs
    flops += S * 2.0 * Ns * N * C
    bytes_ += S * 8.0 * Ns * N
    return flops, bytes_


def run_serving(count):
    """Serving workload: `count` seeded requests (p in {16,32,64}, u in {4,8}, x in {1,32,128},
    1-2 species, 200-2000 steps).  Capacity = solves/s with every request queued at t = 0 (FP32
    grouped kernels with the FP64 moment-rate correction); then Poisson arrivals at 80% of that
    capacity give the time-to-solution percentiles; then every request one at a time in plain
    FP64 (a server with one slot), whose results give the served FP32 number / energy drift;
    then the capacity with split-BF16 grouped kernels (corrected) and its drift."""
    import numpy as np
    from dtx import serving
    specs = serving.request_mix(count=count, rate_per_s=1e9)
    log(f"[bench] serving capacity: {count} requests queued at t=0 (FP32 grouped kernels)")
    res_cap, wall, _ = serving.run(specs, precision="fp32", realtime=False)
    capacity = count / wall
    flops = sum(request_work(sp)[0] * sp.steps for sp in specs)
    gbytes = sum(request_work(sp)[1] * sp.steps for sp in specs)
    rate = 0.8 * capacity
    specs_p = serving.request_mix(count=count, rate_per_s=rate)
    log(f"[bench] serving: Poisson arrivals at {rate:.2f}/s")
    res, wall_p, _ = serving.run(specs_p, precision="fp32", realtime=True)
    lat = serving.latency_stats(res)
    log(f"[bench] serial plain FP64: all {count} requests one at a time")
    res64, wall64, _ = serving.run(specs, precision="fp64", realtime=False, max_active=1)
    serial = count / wall64
    dn = max(abs(a.number_total - b.number_total) / abs(b.number_total)
             for a, b in zip(res_cap, res64))
    de = max(abs(a.energy_total - b.energy_total) / abs(b.energy_total)
             for a, b in zip(res_cap, res64))
    log(f"[bench] serving capacity with split-BF16 grouped kernels: {count} requests at t=0")
    res_bf, wall_bf, _ = serving.run(specs, precision="bf16x3", realtime=False)
    dn_bf = max(abs(a.number_total - b.number_total) / abs(b.number_total)
                for a, b in zip(res_bf, res64))
    de_bf = max(abs(a.energy_total - b.energy_total) / abs(b.energy_total)
                for a, b in zip(res_bf, res64))
    tf = flops / wall / 1e12
    gbs = gbytes / wall / 1e9
    return dict(requests=count, precision="fp32", moment_correction=True, hbm_budget_gb=200,
                capacity_solves_per_s=round(capacity, 3), arrival_rate_per_s=round(rate, 3),
                solves_per_s=round(count / wall_p, 3),
                p50_time_to_solution_s=round(lat["p50_s"], 3),
                p99_time_to_solution_s=round(lat["p99_s"], 3),
                serial_fp64_solves_per_s=round(serial, 3), serial_fp64_requests=count,
                capacity_vs_serial_fp64=round(capacity / serial, 2),
                number_drift_vs_fp64=dn, energy_drift_vs_fp64=de,
                bf16x3=dict(capacity_solves_per_s=round(count / wall_bf, 3),
                            capacity_vs_serial_fp64=round(count / wall_bf / serial, 2),
                            number_drift_vs_fp64=dn_bf, energy_drift_vs_fp64=de_bf),
                roofline=dict(achieved_tflops=round(tf, 2),
                              mfma_frac_fp32=round(tf / PEAK_TFLOPS["fp32"], 4),
                              operator_gbs=round(gbs, 1),
                              hbm_frac=round(gbs / PEAK_HBM_GBS, 4),
                              note="math FLOPs and packed FP32 operator bytes of every grouped "
                                   "step of the capacity run over its wall time"))


def run_collision_ingest():
    """Collision-matrix ingest (dtx_add_binary_interaction) at transport_1d's momentum grid
    (2 species x 256 states, one cell): three interactions, 0.86 GB of DiplodocusCollisions-
    shaped arrays, Ani.  Wall time of the three ingest calls (host C++, threads over the output
    px bins) and of the first step after them (packing the FP64 channels into the device
    operator, upload, the step itself); GB/s of arrays read by the ingest.  Its parity is the
    GPU tests' (oracle at small sizes, the vectorised collision term at this size)."""
    import os
    import numpy as np
    from dtx import collisions, workloads
    from dtx.api import (BinaryMatricesStruct, EmissionMatricesStruct, FluxMatricesStruct,
                         ForwardEulerStruct)
    ps, blist, arrays, f0, Ap, Vol = workloads.collision_ingest()
    BinM = collisions.BuildBinaryMatrices(ps, blist, None, arrays=arrays)
    nbytes = sum(a.nbytes for it in BinM.interactions
                 for a in (it.GainMatrix3, it.GainMatrix4, it.LossMatrix1, it.LossMatrix2))
    # a handle without collisions; the interactions are added to it alone, timed
    m = ForwardEulerStruct(ps, f0, np.zeros(ps.n), BinaryMatricesStruct(),
                           EmissionMatricesStruct(), FluxMatricesStruct(Ap, None, None, Vol))
    t0 = time.perf_counter()
    m._add_interactions(m._h, ps, BinM)
    t1 = time.perf_counter()
    m(0.0, 1e300, 1e-6 * ps.dt0, 0)
    m.synchronize()
    t2 = time.perf_counter()
    m.close()
    return dict(workload="collision_ingest: 2 species x 256 states (p 32 x u 8), one cell, "
                         "3 interactions, Ani", array_bytes=nbytes,
                ingest_s=round(t1 - t0, 4), ingest_gb_per_s=round(nbytes / (t1 - t0) / 1e9, 2),
                first_step_pack_upload_s=round(t2 - t1, 4),
                host_threads=min(16, os.cpu_count() or 1))


def run_flux_setup(name):
    """Set-up time of the flux operators, Ap and Vol: the host path (flux_matrices + X + P as
    CSC, ap_vol, then dtx_set_flux_csc's CSC -> ELL conversion and upload) against
    dtx_synth_flux, which builds the same ELL rows on the device.  Wall-clock, one call each
    after one untimed call of both (first-touch allocations)."""
    import ctypes as C
    import numpy as np
    from dtx import _lib, synth, workloads
    w = getattr(workloads, name)()
    ps, spec, n = w.ps, w.flux_spec, w.ps.n
    m = workloads.make_struct(w, device_flux=True)
    lib, h = m.lib, m._h
    fs, keep = _lib.make_flux_synth(ps, **spec)
    times = {}
    for rep in range(2):
        t0 = time.perf_counter()
        X, P = synth.flux_matrices(ps, **spec)
        F = synth.csc_add(n, X, P)
        Ap, Vol = synth.ap_vol(ps)
        t1 = time.perf_counter()
        cp, rv, nz = (np.ascontiguousarray(a) for a in F)
        _lib.check(lib.dtx_set_flux_csc(h, _lib.FLUX_F, n, _lib.i64ptr(cp), _lib.i64ptr(rv),
                                        _lib.dptr(nz)), h, "dtx_set_flux_csc")
        t2 = time.perf_counter()
        _lib.check(lib.dtx_synth_flux(h, C.byref(fs)), h, "dtx_synth_flux")
        t3 = time.perf_counter()
        times = {"host_assembly_s": round(t1 - t0, 4), "host_csc_to_ell_upload_s":
                 round(t2 - t1, 4), "device_synth_s": round(t3 - t2, 5)}
    host = times["host_assembly_s"] + times["host_csc_to_ell_upload_s"]
    # the two operators at this size: the same device footprint (ELL width) and the same state
    # after a step from f0 (the handle's last operator is the device-built one)
    ref = workloads.make_struct(w)
    bytes_dev, bytes_host = m.device_bytes(), ref.device_bytes()
    m.reset()
    m(0.0, 1e300, w.dt, 0)
    ref(0.0, 1e300, w.dt, 0)
    fd, fh = m.f, ref.f
    rel = float(np.max(np.abs(fd - fh)) / np.max(np.abs(fh)))
    ref.close()
    m.close()
    print(json.dumps({"flux_setup": dict(workload=name, n=n, nnz=int(F[0][-1] - 1), **times,
                                         speedup=round(host / times["device_synth_s"], 1),
                                         device_bytes_equal=bytes_dev == bytes_host,
                                         step_rel_linf=rel,
                                         step_bitwise=bool(np.array_equal(fd, fh)))}))
    del keep
    if bytes_dev != bytes_host or rel > 1e-13:
        sys.exit("flux set-up: the device-built operator differs from the host CSC")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="transport_1d",
                    choices=["transport_1d", "precision_box", "transport_2d", "homogeneous_box"],
                    help="transport_1d is the headline; the others are the precision, largest-"
                         "solve and homogeneous-box configurations, measured the same way")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 500 for transport_1d, else the workload's own)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-opt-steps", type=int, default=5,
                    help="timed steps of the optimised CPU column (batched DGEMM)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precisions", default="fp64,fp32,bf16x3",
                    help="contraction precisions to time (fp64 is always run)")
    ap.add_argument("--no-uncorrected", action="store_true",
                    help="skip the FP32 / BF16x3 runs without the FP64 moment-rate correction")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-ingest", action="store_true",
                    help="skip the collision-matrix ingest timing")
    ap.add_argument("--serving-requests", type=int, default=256)
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the full-length accuracy runs of the precision modes")
    ap.add_argument("--no-largest", action="store_true",
                    help="skip the transport_2d FP64 / FP32 / split-BF16 comparison")
    ap.add_argument("--no-emission", action="store_true",
                    help="skip transport_1d with emission blocks (the emission GEMV's HBM rate)")
    ap.add_argument("--largest-steps", type=int, default=50,
                    help="timed steps per precision of the transport_2d comparison")
    ap.add_argument("--copy-kernel-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (5 times), the "
                         "known-size kernel that scales the FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving capacity run (--serving-requests requests queued at "
                         "t = 0, grouped kernels of --serving-precision): one JSON line with "
                         "solves/s")
    ap.add_argument("--serving-precision", default="fp32", choices=["fp64", "fp32", "bf16x3"],
                    help="contraction precision of the --serving-only run")
    ap.add_argument("--flux-setup", action="store_true",
                    help="only compare the flux set-up of the workload: host assembly + CSC "
                         "upload against the stencils built on the device (dtx_synth_flux)")
    args = ap.parse_args()
    if args.flux_setup:
        run_flux_setup(args.workload)
        return
    if args.serving_only:
        from dtx import serving
        specs = serving.request_mix(count=args.serving_requests, rate_per_s=1e9)
        _, wall, _ = serving.run(specs, precision=args.serving_precision, realtime=False)
        print(json.dumps({"serving_requests": args.serving_requests,
                          "precision": args.serving_precision,
                          "capacity_solves_per_s": round(args.serving_requests / wall, 3),
                          "request_steps": sum(sp.steps for sp in specs)}))
        return
    if args.copy_kernel_mib:
        import torch
        n = args.copy_kernel_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the path shards by cells across GPUs only by design, see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from dtx import workloads
    t0 = time.perf_counter()
    w = getattr(workloads, args.workload)()
    steps = args.steps if args.steps is not None else (500 if args.workload == "transport_1d"
                                                        else w.steps)
    log(f"workload {w.name} built in {time.perf_counter() - t0:.1f}s: N={w.ps.n_momentum} "
        f"C={w.ps.n_space} n={w.ps.n} dt={w.dt:.4e}")
    res = {}
    runs = []
    precisions = [p for p in args.precisions.split(",") if p]
    if "fp64" not in precisions:  # the line's value and every comparison are the FP64 run's
        log("[bench] fp64 added to --precisions: the reported value is the FP64 rate")
        precisions.insert(0, "fp64")
    for prec in precisions:
        runs.append((prec, True))
        if prec != "fp64" and not args.no_uncorrected:
            runs.append((prec, False))
    acc_steps = 0 if args.no_accuracy else w.steps
    for prec, corr in runs:
        key = prec if corr or prec == "fp64" else prec + "_uncorrected"
        log(f"[bench] {key}: {args.warmup} warmup + {steps} timed steps, {TIMING_STEPS} event-"
            f"timed steps, {acc_steps}-step accuracy run")
        res[key] = run_gpu(w, prec, steps, args.warmup, moment_correction=corr,
                           accuracy_steps=acc_steps)
        r = res[key]
        log(f"[bench] {key}: {r['steps_per_s']:.1f} steps/s, binary kernel "
            f"{r['bin_ms']*1e3:.1f} us, update kernel {r['update_ms']*1e3:.1f} us, "
            f"{r['items']} work items")
    base = res["fp64"]
    prec_report = {}
    for key, r in res.items():
        if key == "fp64":
            continue
        prec_report[key] = dict(
            steps_per_s=round(r["steps_per_s"], 2),
            speedup_vs_fp64=round(r["steps_per_s"] / base["steps_per_s"], 3),
            binary_kernel_us=round(r["bin_ms"] * 1e3, 2),
            binary_tflops=round(r["bin_flops"] / (r["bin_ms"] * 1e-3) / 1e12, 2),
            binary_hbm_gbs=round(r["bin_bytes"] / (r["bin_ms"] * 1e-3) / 1e9, 1))
        if acc_steps:
            prec_report[key].update(accuracy_steps=acc_steps, **accuracy(r, base))
    # The contraction is MFMA-bound when its arithmetic intensity exceeds the FP64 ridge
    # (78.6 TFLOP/s / 8 TB/s), else HBM-bound (single-cell boxes stream the whole operator).
    ai = base["bin_flops"] / base["bin_bytes"]
    C = w.ps.n_space
    kernel = ("bin_gemv_kernel<0>" if C == 1 else "bin_f64_kernel<16>" if C <= 16 else
              "bin_f64_wide_kernel" if 64 < C <= 256 else "bin_f64_kernel<64>")
    if ai >= PEAK_TFLOPS["fp64"] * 1e12 / (PEAK_HBM_GBS * 1e9):
        achieved = base["bin_flops"] / (base["bin_ms"] * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=round(achieved, 3), peak=PEAK_TFLOPS["fp64"],
                    unit="TFLOP/s", frac=round(achieved / PEAK_TFLOPS["fp64"], 4),
                    traffic=TRAFFIC.get((w.name, kernel)), kernel=kernel,
                    flops_per_launch=base["bin_flops"], bytes_per_launch=base["bin_bytes"],
                    kernel_avg_us=round(base["bin_ms"] * 1e3, 2))
    else:
        achieved = base["bin_bytes"] / (base["bin_ms"] * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBS, 4),
                    traffic=TRAFFIC.get((w.name, kernel)), kernel=kernel,
                    flops_per_launch=base["bin_flops"], bytes_per_launch=base["bin_bytes"],
                    kernel_avg_us=round(base["bin_ms"] * 1e3, 2))
    emission = None
    if not args.no_emission and args.workload == "transport_1d":
        emission = run_emission(200, 10)
    largest = None
    if not args.no_largest and args.workload == "transport_1d":
        largest = run_largest(args.largest_steps, 5)
    ingest = None
    if not args.no_ingest and args.workload == "transport_1d":
        log("[bench] collision-matrix ingest (3 interactions, transport_1d momentum grid)")
        ingest = run_collision_ingest()
        log(f"[bench] ingest {ingest['ingest_s']} s ({ingest['ingest_gb_per_s']} GB/s), first "
            f"step with pack + upload {ingest['first_step_pack_upload_s']} s")
    serving_line = None
    if not args.no_serving and args.workload == "transport_1d":
        serving_line = run_serving(args.serving_requests)
    cpu = cpu_opt = None
    if not args.no_cpu and args.workload in ("transport_1d", "homogeneous_box"):
        log(f"[bench] cpu baseline: {args.cpu_steps} oracle steps")
        cpu = run_cpu_baseline(w, args.cpu_steps)
        log("[bench] optimised CPU column: binary term batched over cells (one DGEMM)")
        cpu_opt = run_cpu_optimized(w, args.cpu_opt_steps)
    line = {
        "metric": "steps/s + solves/s, p50/p99 latency; FP32-MFMA vs FP64 at <=1e-6 rel err",
        "value": round(base["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(base["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal gain tensors, Maxwell-Juttner f0)",
        "config": {"workload": f"{w.name}: {DESCRIPTIONS[w.name]}",
                   "solve_steps": w.steps,
                   "solves_per_s": round(base["steps_per_s"] / w.steps, 4),
                   "kernels_per_step": fp64_kernels_per_step(w)},
        "precision_modes": prec_report,
        "update_kernel_us": round(base["update_ms"] * 1e3, 2),
        "roofline": roof,
        "largest_solve": largest,
        "emission": emission,
        "serving": serving_line,
        "collision_ingest": ingest,
        "cpu_baseline": cpu,
        "cpu_optimized": cpu_opt,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
