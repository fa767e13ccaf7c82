#!/usr/bin/env python3
"""Benchmark of the explicit-Euler step on one MI355X.

Workload (default): 1D transport -- 2 species (electron, photon) x (32 log p bins on
[1e-3, 1e3] x 8 u bins x 1 phi bin) = 512 states per cell, 128 periodic x cells (65,536
unknowns), upwind x advection + synchrotron cooling, per-cell binary collisions with lognormal
gain tensors, plain FP64.  A step is one ForwardEuler functor call over the whole state; inputs
are resident in HBM before the timed region.

Prints ONE JSON line: steps/s of the FP64 path (value), with the FP32-MFMA and split-BF16 paths
and their relative error against FP64 beside it, the live roofline of the dominant kernel (the
binary contraction, HIP events on the library's stream around the kernels of every 8th timed
step: events around every step would add about 14 us to each) and the FP64 CPU oracle timed on
a bounded sample on the host cores.
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
TIME_EVERY = 8   # kernel durations from HIP events on every 8th timed step

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
TRAFFIC = {("transport_1d", "bin_f64_wide_kernel"): 497_210_000,
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


def run_gpu(w, precision, steps, warmup, verbose=1, moment_correction=True):
    from dtx import workloads
    m = workloads.make_struct(w, precision=precision, use_monolithic=False,
                              moment_correction=moment_correction)
    work = m.work()
    m.reset()
    t, big = 0.0, 1e300
    for _ in range(warmup):
        dt, _ = m(t, big, w.dt, verbose)
        t += dt
    m.synchronize()
    m.kernel_timing(enable=True, reset=True, every=TIME_EVERY if steps >= 10 * TIME_EVERY else 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        dt, _ = m(t, big, w.dt, verbose)
        t += dt
    m.synchronize()
    t1 = time.perf_counter()
    kt = m.kernel_timing(enable=False, reset=True)
    f_end = m.f
    num, en = m.moments()
    res = dict(ms_per_step=(t1 - t0) * 1e3 / steps, steps_per_s=steps / (t1 - t0),
               bin_ms=kt.bin_ms / max(1, kt.bin_launches),
               update_ms=kt.update_ms / max(1, kt.update_launches),
               bin_flops=work.bin_flops, gain_bytes=work.gain_bytes, bin_bytes=work.bin_bytes,
               update_bytes=work.update_bytes, items=work.bin_launch_items,
               device_gb=m.device_bytes() / 1e9, f=f_end, number=float(num.sum()),
               energy=float(en.sum()))
    m.close()
    return res


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


def run_serving(count, serial_sample):
    """Serving workload: `count` seeded requests (p in {16,32,64}, u in {4,8}, x in {1,32,128},
    1-2 species, 200-2000 steps).  Capacity = solves/s with every request queued at t = 0; then
    Poisson arrivals at 80% of that capacity give the time-to-solution percentiles.  The serial
    plain-FP64 rate runs the first `serial_sample` requests one at a time."""
    from dtx import serving
    specs = serving.request_mix(count=count, rate_per_s=1e9)
    log(f"[bench] serving capacity: {count} requests queued at t=0 (FP32 grouped kernels)")
    res, wall, _ = serving.run(specs, precision="fp32", realtime=False)
    capacity = count / wall
    rate = 0.8 * capacity
    specs = serving.request_mix(count=count, rate_per_s=rate)
    log(f"[bench] serving: Poisson arrivals at {rate:.2f}/s")
    res, wall, _ = serving.run(specs, precision="fp32", realtime=True)
    lat = serving.latency_stats(res)
    serial = None
    if serial_sample:
        log(f"[bench] serial plain FP64: first {serial_sample} requests one at a time")
        t0 = time.perf_counter()
        for sp in specs[:serial_sample]:
            serving.run([sp], precision="fp64", realtime=False)
        serial = serial_sample / (time.perf_counter() - t0)
    return dict(requests=count, precision="fp32", hbm_budget_gb=200,
                capacity_solves_per_s=round(capacity, 3), arrival_rate_per_s=round(rate, 3),
                solves_per_s=round(count / wall, 3), p50_time_to_solution_s=round(lat["p50_s"], 3),
                p99_time_to_solution_s=round(lat["p99_s"], 3),
                serial_fp64_solves_per_s=round(serial, 3) if serial else None,
                serial_fp64_sample=serial_sample)


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
    print(json.dumps({"flux_setup": dict(workload=name, n=n, nnz=int(F[0][-1] - 1), **times,
                                         speedup=round(host / times["device_synth_s"], 1))}))
    del keep


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
    ap.add_argument("--precisions", default="fp64,fp32,bf16x3")
    ap.add_argument("--no-uncorrected", action="store_true",
                    help="skip the FP32 / BF16x3 runs without the FP64 moment-rate correction")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=256)
    ap.add_argument("--serving-serial-sample", type=int, default=16)
    ap.add_argument("--copy-kernel-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (5 times), the "
                         "known-size kernel that scales the FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--flux-setup", action="store_true",
                    help="only compare the flux set-up of the workload: host assembly + CSC "
                         "upload against the stencils built on the device (dtx_synth_flux)")
    args = ap.parse_args()
    if args.flux_setup:
        run_flux_setup(args.workload)
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
    for prec in args.precisions.split(","):
        runs.append((prec, True))
        if prec != "fp64" and not args.no_uncorrected:
            runs.append((prec, False))
    for prec, corr in runs:
        key = prec if corr or prec == "fp64" else prec + "_uncorrected"
        log(f"[bench] {key}: {args.warmup} warmup + {steps} timed steps")
        res[key] = run_gpu(w, prec, steps, args.warmup, moment_correction=corr)
        r = res[key]
        log(f"[bench] {key}: {r['steps_per_s']:.1f} steps/s, binary kernel "
            f"{r['bin_ms']*1e3:.1f} us, update kernel {r['update_ms']*1e3:.1f} us, "
            f"{r['items']} work items")
    base = res["fp64"]
    fp = base["f"]
    prec_report = {}
    for key, r in res.items():
        if key == "fp64":
            continue
        mask = np.abs(fp) > 1e-8 * np.max(np.abs(fp))
        prec_report[key] = dict(
            steps_per_s=round(r["steps_per_s"], 2),
            speedup_vs_fp64=round(r["steps_per_s"] / base["steps_per_s"], 3),
            rel_linf_f=float(np.max(np.abs(r["f"][mask] - fp[mask]) / np.abs(fp[mask]))),
            number_drift=abs(r["number"] - base["number"]) / abs(base["number"]),
            energy_drift=abs(r["energy"] - base["energy"]) / abs(base["energy"]),
            binary_kernel_us=round(r["bin_ms"] * 1e3, 2),
            binary_tflops=round(r["bin_flops"] / (r["bin_ms"] * 1e-3) / 1e12, 2),
            binary_hbm_gbs=round(r["bin_bytes"] / (r["bin_ms"] * 1e-3) / 1e9, 1))
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
    serving_line = None
    if not args.no_serving and args.workload == "transport_1d":
        serving_line = run_serving(args.serving_requests, args.serving_serial_sample)
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
        "serving": serving_line,
        "cpu_baseline": cpu,
        "cpu_optimized": cpu_opt,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
