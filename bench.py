"""Benchmark: cell-updates/s of the KH time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--calibrate-copy MB]

Workload (the headline config): plain FP64 single run on a 2048x1024 grid,
x walled, z periodic, dx = dz = 0.1, alpha = 1, V0 = 0.05, a = 1.0, seeded
kz mode (amplitude 1e-6) plus 1e-9 white noise (seed 1), CFL 0.4, no mode
filter.  With Lx = 204.8 the frozen background Az1 = -V0 log cosh(xi) reaches
-5 at the walls and the tachyonic outer branch blows the run up by step ~110
(the reference's own halt rule E/E0 > 5 fires there); the workload therefore
adds the reference's xi_cut key at |xi| = 12.8 (a Dirichlet wall for the
colour-2/3 sector at C1's wall position), which keeps E/E0 < 5 for ~980 steps.
A "step" is one full time step of the reference loop (two FCT sweeps per
species, Maxwell, sources, the xi_cut, and the energy check at the reference
cadence every (int)(0.5/DT) steps).  All state is resident in HBM before the
timed region.  Timed regions longer than SEGMENT steps are split into
segments that restart from the initial state (untimed) so every timed step is
physically finite.  The FP32 precision form of the same workload is timed too,
and so is the precision-form workload (4096x2048, V0 = 0.1, a = 0.5, keep
|kz| <= 8 on MFMA) in FP32 and FP64, with the kz = 1 growth rate of both
(max-R^2 fit over the same window) and its relative error, and the batched
sweep (64 uneven FP32 runs through captured graphs: runs/hour), and the
EMHD 1024x1024 FP32 wave run with its MFMA space-time dispersion check.

Prints one JSON line (one GPU: N = 1 only).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "kh-simulations_amd"))

METRIC = "cell-updates/s/GPU FP32 vs FP64 at ≤1e-3 growth-rate err; sweep runs/hour"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(nx_override=2048, nz_override=1024, lx_override=204.8, lz_override=102.4, alpha_YM=1,
                V0=0.05, eps_override=1.0, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1,
                perturb_amp=2e-5, noise_amp=1e-9, noise_seed=1, n_steps=5000, xi_cut=12.8)
SEGMENT = 700
# Precision-form workload: 4096x2048, V0 = 0.1, a = 0.5, keep |kz| <= 8 (the
# column filter runs as MFMA contractions), xi_cut at |xi| = 12.8 for the same
# reason as above; E/E0 stays below 5 for ~650 steps, so segments are 500.
WORKLOAD_C3 = dict(nx_override=4096, nz_override=2048, lx_override=409.6, lz_override=204.8, alpha_YM=1,
                   V0=0.1, eps_override=0.5, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1,
                   perturb_amp=1e-5, noise_amp=1e-9, noise_seed=1, n_steps=10000, kz_keep_max=8, xi_cut=12.8)
SEGMENT_C3 = 500

# Compulsory (algorithmic) HBM traffic of each step kernel, in arrays per cell:
# every input array read once and every output written once, no halo re-reads.
#   k_fct_x : R fluid 12 + Ex 3 + By 3,            W stage 12
#   k_fct_z : R stage 12 + step-start fluid 12 + Ez 3 + By 3 + Az 3, W fluid 12
#   k_maxwell: R new fluid 12 + fields 12,         W E 6 + By 3 + Az2,3 2 (Az1 frozen)
#   k_fct_xz: R fluid 12 + Ex, Ez, By, Az 12,       W fluid 12
KERNEL_ARRAYS = {"k_fct_x": 18 + 12, "k_fct_z": 33 + 12, "k_fct_xz": 24 + 12, "k_maxwell": 24 + 11,
                 "k_fct_zm": 36 + 23}
# The whole step's algorithmic minimum: R 24 + W 23 arrays (frozen Az1 read only).
STEP_ARRAYS = 47
ZM_ROWS = {1: 16, 2: 32, 3: 64}       # chunk height of the fused z sweep + Maxwell variants
# What limits the step kernels other than HBM bytes, from rocprofv3 SQ counters
# on this workload (profiles/sq_counters_2048x1024_fp64.md), keyed by slot.
LIMITER = {
    "k_fct_xz": "FP64 VALU issue and latency at 2 waves per SIMD: SIMD VALU busy 0.43 of the kernel's "
                "cycles, waves waiting on memory counters 0.32 of their life (profiles/sq_counters_2048x1024_fp64.md)",
    "k_maxwell": "HBM: 702 MB per launch (1.20x algorithmic) at 5.6 TB/s; waves wait on memory 0.47 of their life",
}
# HBM bytes per launch of the FP64 step kernels on this workload, measured with
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (separate runs), each scaled
# by what the counter reads for a plain 256 MiB device copy (fetch x2.0,
# write x1.0); summary under profiles/.  Keyed by kernel instance.
PMC_TRAFFIC_MB = {
    "k_fct_x<double, 128, YMModel, 0>": 507.6, "k_fct_x<double, 128, YMModel, 1>": 512.2,
    "k_fct_x<double, 256, YMModel, 0>": 507.6, "k_fct_x<double, 256, YMModel, 1>": 512.3,
    "k_fct_x<double, 64, YMModel, 0>": 507.4, "k_fct_x<double, 64, YMModel, 1>": 512.8,
    "k_fct_xr<double, 4>": 507.7, "k_fct_xr<double, 8>": 528.5, "k_fct_xz2<double, 16>": 706.8,
    "k_fct_xz2<double, 32>": 665.9, "k_fct_xz2<double, 64>": 642.0, "k_fct_xz<double, 32>": 680.7,
    "k_fct_xzg<double, 16, 256>": 677.1, "k_fct_xzg<double, 32, 128>": 725.2, "k_fct_xzg<double, 32, 256>": 704.2,
    "k_fct_xzg<double, 64, 256>": 699.4, "k_fct_xzg<double, 8, 128>": 657.4, "k_fct_xzr<double, 128>": 648.8,
    "k_fct_xzr<double, 32>": 680.6, "k_fct_xzr<double, 64>": 661.0, "k_fct_z2<double, 16, YMModel>": 970.3,
    "k_fct_z2<double, 32, YMModel>": 865.2, "k_fct_z2<double, 64, YMModel>": 811.3,
    "k_fct_z<double, 16, 0, YMModel>": 978.6, "k_fct_z<double, 16, 1, YMModel>": 971.1,
    "k_fct_z<double, 32, 0, YMModel>": 870.4, "k_fct_z<double, 32, 1, YMModel>": 869.1,
    "k_fct_z<double, 64, 0, YMModel>": 806.5, "k_fct_z<double, 64, 1, YMModel>": 815.4,
    "k_fct_zg<double, 16, 0, YMModel>": 1051.1, "k_fct_zg<double, 16, 1, YMModel>": 993.1,
    "k_fct_zg<double, 32, 0, YMModel>": 982.8, "k_fct_zg<double, 32, 1, YMModel>": 927.8,
    "k_fct_zg<double, 64, 0, YMModel>": 936.6, "k_fct_zg<double, 64, 1, YMModel>": 940.0,
    "k_maxwell<double, 16>": 846.5, "k_maxwell<double, 4>": 718.5, "k_maxwell<double, 8>": 859.2,
    "k_maxwell_cell<double>": 597.7, "k_maxwell_zm<double, 16, 0>": 646.3, "k_maxwell_zm<double, 16, 1>": 640.0,
    "k_maxwell_zm<double, 32, 1>": 614.0, "k_maxwell_zm<double, 4, 1>": 702.4, "k_maxwell_zm<double, 8, 0>": 678.3,
    "k_maxwell_zm<double, 8, 1>": 662.8,
}


def kernel_instance(slot, t):
    """Kernel instance (template arguments) that the tuned slot launches."""
    if slot == "k_fct_x":
        if t.get("x_lane_cells", 0) > 0:
            return f"k_fct_xr<double, {t['x_lane_cells']}>"
        return f"k_fct_x<double, {t['x_segment']}, YMModel, {t['x_serial']}>"
    if slot == "k_fct_z":
        name = "k_fct_zg" if t["z_groups"] else "k_fct_z"
        return f"{name}<double, {t['z_chunk']}, {t['z_prefetch']}, YMModel>"
    if slot == "k_fct_zm":
        return f"k_fct_zmx<double, {ZM_ROWS.get(t.get('zm_fused', 0), 0)}>"
    if slot == "k_fct_xz":
        v = t["xz_variant"]
        name, tz = {0: ("k_fct_xz", 32), 1: ("k_fct_xz2", 16), 2: ("k_fct_xz2", 32), 3: ("k_fct_xz2", 64),
                    4: ("k_fct_xzr", 32), 5: ("k_fct_xzr", 64), 6: ("k_fct_xzr", 128), 7: ("k_fct_xzg", 16),
                    8: ("k_fct_xzg", 32), 9: ("k_fct_xzg", 64), 10: ("k_fct_xzg", 32),
                    11: ("k_fct_xzg", 8)}.get(v, (None, 0))
        if name is None:
            return None
        if name == "k_fct_xzg":
            return f"{name}<double, {tz}, {128 if v >= 10 else 256}>"
        return f"{name}<double, {tz}>"
    if slot == "k_maxwell" and t["maxwell"] >= 4:
        tzpf = {4: (16, 1), 5: (32, 1), 6: (16, 0), 7: (8, 1), 8: (8, 0), 9: (4, 1)}.get(t["maxwell"])
        return f"k_maxwell_zm<double, {tzpf[0]}, {tzpf[1]}>" if tzpf else None
    if slot == "k_maxwell":
        return {0: "k_maxwell<double, 16>", 1: "k_maxwell_cell<double>", 2: "k_maxwell<double, 4>",
                3: "k_maxwell<double, 8>"}.get(t["maxwell"])
    return None


def run_gpu(precision, steps, warmup, fused=-1, workload=WORKLOAD, segment=SEGMENT, kernel_steps=120):
    """Timed khs_step over `steps` steps (segments restarting from init; whole
    energy strides replay as captured graphs), then `kernel_steps` eager steps
    with HIP events around every launch for the per-kernel durations."""
    import khs

    keys = dict(workload, fused_sweep=fused)
    if precision == "fp32":
        keys["precision"] = "fp32"
    cfg = khs.Config(**keys)
    sim = khs.Simulation(cfg)
    total_s, done, nseg = 0.0, 0, 0
    times = {}
    e_ratio = None
    sim.init()
    if sim.step(2 * cfg.energy_stride + warmup):      # autotune and graph capture, untimed
        raise RuntimeError("warm-up halted")
    while done < steps:
        seg = min(segment, steps - done)
        sim.init()
        if sim.step(warmup):
            raise RuntimeError("warm-up halted")
        sim.sync()
        t0 = time.perf_counter()
        halted = sim.step(seg)
        sim.sync()
        total_s += time.perf_counter() - t0
        if halted:
            raise RuntimeError("timed region halted")
        e_ratio = sim.energy() / sim.energy0
        done += seg
        nseg += 1
    sim.init()
    if sim.step(warmup):
        raise RuntimeError("warm-up halted")
    sim.sync()
    sim.timing(True)
    sim.timing_reset()
    t0 = time.perf_counter()
    if sim.step(kernel_steps):
        raise RuntimeError("kernel-timing steps halted")
    sim.sync()
    eager_s = time.perf_counter() - t0
    for k, (ms, n) in sim.kernel_times().items():
        times[k] = (ms, n)
    sim.timing(False)
    cells = cfg.nx * cfg.nz
    tuning = sim.tuning()
    sim.close()
    return dict(seconds=total_s, cells=cells, times=times, e_ratio=e_ratio, nx=cfg.nx, nz=cfg.nz,
                segments=nseg, tuning=tuning, eager_ms_per_step=eager_s / kernel_steps * 1e3)


def growth_c3(precision, n_steps=500, stride=10):
    """kz = 1 growth rate of the precision-form workload (max-R^2 fit of the
    circular Az2/Az3 amplitude, sampled every `stride` steps from init)."""
    import khs

    cfg = khs.Config(**WORKLOAD_C3, precision=precision)
    sim = khs.Simulation(cfg)
    sim.init()
    ts = sim.timeseries(1, n_steps, stride)
    ts["energy_ratio"] = sim.energy() / sim.energy0
    ts["gauss"] = sim.gauss()
    sim.close()
    if ts["halted"]:
        raise RuntimeError("growth-rate run halted")
    return ts


def sweep(n=64, steps=3000, lanes=8, segment=50, threads=16):
    """Batched-sweep workload: n uneven-grid FP32 runs (khs.sweep) stepped
    three ways: all runs' segments of `segment` steps as one captured graph
    over `lanes` streams ("graph"); `threads` host threads each stepping its
    share of the runs on their own streams with graph-replayed energy strides
    ("concurrent"); the runs one after another ("eager").  A run ends at
    `steps` or when it halts (E/E0 rule)."""
    import khs
    from khs.sweep import sweep_configs

    cfgs = sweep_configs(n, steps)
    cells = [c["nx_override"] * c["nz_override"] for c in cfgs]
    out = {"workload": f"{n} runs, NX in 256..2048 x NZ in 128..1024, V0~U[0.01,0.2], a~U[0.5,2], kz 1..3, "
                       f"half with suppress_kz0, FP32, {steps} steps each (a run that halts on the E/E0 rule "
                       f"is counted in 'halted'), CFL 0.1, xi_cut=4",
           "lanes": lanes, "segment": segment, "threads": threads,
           "cell_steps": sum(cells) * steps}
    for mode in ("graph", "concurrent", "eager"):
        b = khs.Batch(cfgs, lanes=lanes, segment=segment)
        b.init()
        run = {"graph": b.run, "eager": b.run_eager,
               "concurrent": lambda k: b.run_concurrent(k, threads=threads)}[mode]
        t0 = time.perf_counter()
        done, halted = run(steps)
        el = time.perf_counter() - t0
        b.close()
        out[mode] = {"seconds": el, "runs_per_hour": n / el * 3600.0, "halted": int(sum(halted)),
                     "completed": int(sum(d >= steps for d in done)),
                     "cell_updates_per_s": sum(c * d for c, d in zip(cells, done)) / el}
    out["graph_over_eager"] = out["eager"]["seconds"] / out["graph"]["seconds"]
    out["concurrent_over_eager"] = out["eager"]["seconds"] / out["concurrent"]["seconds"]
    best = min(("graph", "concurrent"), key=lambda m: out[m]["seconds"])
    out["runs_per_hour"] = out[best]["runs_per_hour"]
    out["halted"] = out[best]["halted"]
    out["best_mode"] = best
    return out


def emhd_waves(n=1024, steps=20000, stride=20, kmax=32, timed_steps=2000):
    """Two-fluid EMHD wave run: n x n periodic, FP32, random-phase By spectrum
    (kx, kz = 1..32, amplitude 1e-5, seed 2).  The step is timed on its own;
    then the whole run records the By spectrum (MFMA z and x contractions)
    every `stride` steps and every mode's frequency is compared with the
    dispersion relation of the discretised model."""
    import numpy as np
    import khs
    from khs.dispersion import emhd_wave_omega, peak_frequency, record_spectrum

    keys = dict(model="emhd", nx_override=n, nz_override=n, periodic=1, species=2, emhd_init="spectrum",
                spectrum_kmax=kmax, spectrum_amp=1e-5, spectrum_seed=2, smooth_init=0, precision="fp32",
                n_steps=steps)
    sim = khs.Simulation(khs.Config(**keys))
    sim.init()
    sim.step(20)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(timed_steps)
    sim.sync()
    el = time.perf_counter() - t0
    sim.init()
    t1 = time.perf_counter()
    t, F = record_spectrum(sim, "By", kmax, steps, stride)
    el_run = time.perf_counter() - t1
    cfg = sim.config
    sim.close()
    kx, kz = np.meshgrid(np.arange(1, kmax + 1), np.arange(1, kmax + 1))
    want = emhd_wave_omega(kx, kz, cfg.nx, cfg.nz, cfg.dx, cfg.dz, charges=(-1.0, 1.0), densities=(1.0, 1.0))
    got = np.array([[peak_frequency(t, F[:, i, j]) for j in range(kmax)] for i in range(kmax)])
    err = np.abs(got / want - 1.0)
    return {"workload": f"EMHD two-fluid {n}x{n} periodic FP32, random-phase By kx,kz=1..{kmax}, amp 1e-5, seed 2",
            "cell_updates_per_s": n * n * timed_steps / el, "ms_per_step": el / timed_steps * 1e3,
            "run_steps": steps, "spectrum_stride": stride, "run_seconds_with_spectra": el_run,
            "modes": kmax * kmax, "omega_rel_err_max": float(err.max()), "omega_rel_err_median": float(np.median(err)),
            "relation": "omega^2 = 2 sum q^2 n + k_Yee^2 (both FCT sweeps apply the Lorentz impulse)"}


def profile_leg(leg, steps, warmup, fused=-1):
    """One short leg for a profiler: the kernels of the precision form (C3,
    both precisions) or of the EMHD wave run with its spectrum (C5)."""
    import khs

    if leg in ("sweep", "sweep-graph"):
        from khs.sweep import sweep_configs

        b = khs.Batch(sweep_configs(64, steps), lanes=8, segment=50)
        b.init()
        t0 = time.perf_counter()
        done, halted = b.run_concurrent(steps, threads=16) if leg == "sweep" else b.run(steps)
        el = time.perf_counter() - t0
        b.close()
        return {"leg": leg, "runs": 64, "steps": steps, "seconds": el, "halted": int(sum(halted)),
                "run_steps": int(sum(done))}
    if leg == "c3":
        out = {}
        for prec in ("fp32", "fp64"):
            r = run_gpu(prec, steps, warmup, fused, WORKLOAD_C3, SEGMENT_C3, kernel_steps=steps)
            out[prec] = dict(ms_per_step=r["seconds"] / steps * 1e3, tiles=r["tuning"],
                             kernels_avg_us={k: v[0] / v[1] * 1e3 for k, v in r["times"].items() if v[1]})
        return {"leg": "c3", "workload": "4096x2048 keep |kz|<=8 (MFMA filter)", "steps": steps, **out}
    from khs.dispersion import record_spectrum

    keys = dict(model="emhd", nx_override=1024, nz_override=1024, periodic=1, species=2, emhd_init="spectrum",
                spectrum_kmax=32, spectrum_amp=1e-5, spectrum_seed=2, smooth_init=0, precision="fp32",
                n_steps=steps)
    sim = khs.Simulation(khs.Config(**keys))
    sim.init()
    sim.step(warmup)
    sim.sync()
    t0 = time.perf_counter()
    record_spectrum(sim, "By", 32, steps, 20)
    el = time.perf_counter() - t0
    sim.close()
    return {"leg": "c5", "workload": "EMHD two-fluid 1024x1024 FP32, By spectrum kx,kz=1..32 every 20 steps",
            "steps": steps, "seconds": el}


def cpu_baseline(budget_s=15.0):
    """FP64 OpenMP oracle on the host cores, same workload, bounded sample."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import khs
    import oracle_ref as orc

    cfg = khs.Config(**WORKLOAD)
    o = orc.Oracle(orc.params_from_config(cfg))
    o.init_from_config(cfg)
    o.sources()
    o.step(1)                          # warm-up (page-in, thread pool)
    n, t0 = 0, time.perf_counter()
    while True:
        o.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 200:
            break
    cells = cfg.nx * cfg.nz
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
    except OSError:
        pass
    triad = orc.stream_triad()
    value = cells * n / el
    return dict(value=value, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"{n} FP64 steps of the 2048x1024 workload ({el:.1f} s wall, "
                       f"OMP_NUM_THREADS={threads}, energy check excluded)",
                cpu_model=model, stream_triad_GBs=triad,
                roofline_frac=(value * STEP_ARRAYS * 8 / 1e9 / triad) if triad else None)


def calibrate_copy(mb):
    """Plain device-to-device copy of known size (the HIP runtime's blit kernel),
    for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE of the step kernels."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mb * (1 << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    for _ in range(5):
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)      # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 4096x2048 precision-form leg")
    ap.add_argument("--c3-steps", type=int, default=200)
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep (runs/hour) leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the EMHD wave-dispersion leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB")
    ap.add_argument("--profile-leg", choices=["c3", "c5", "sweep", "sweep-graph"], default=None,
                    help="run one short leg for a profiler and print its kernel times: c3 = the 4096x2048 "
                         "precision form (FP32 and FP64, MFMA keep filter), c5 = the EMHD 1024^2 run with "
                         "its MFMA space-time spectrum every 20 steps, sweep / sweep-graph = the 64-run "
                         "sweep in the concurrent / single-thread graph mode (--steps steps per run)")
    ap.add_argument("--split", action="store_true", help="force the separate x and z sweep kernels")
    ap.add_argument("--fused", action="store_true", help="force the fused x+z sweep kernel")
    a = ap.parse_args()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; the multi-GPU path (independent runs per GPU or x-slabs) "
              "is a design, not a measured configuration. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    fused = 1 if a.fused else (0 if a.split else -1)     # -1: the autotuner's choice
    if a.profile_leg:
        print(json.dumps(profile_leg(a.profile_leg, a.steps, a.warmup, fused)))
        return
    r64 = run_gpu("fp64", a.steps, a.warmup, fused)
    ms_step = r64["seconds"] / a.steps * 1e3
    value = r64["cells"] * a.steps / r64["seconds"]

    # per-kernel live roofline (HIP events on the library's stream); the xi_cut
    # pass writes Q2/Q3 of both species on the cut columns only
    import numpy as np
    x = np.arange(r64["nx"]) * 0.1
    cut_frac = float(np.mean(np.abs((x - r64["nx"] * 0.1 / 2) / 1.0) > 12.8))
    arrays_of = dict(KERNEL_ARRAYS, filters=4 * cut_frac)
    kern = {}
    for name, arrays in arrays_of.items():
        ms, cnt = r64["times"].get(name, (0.0, 0))
        if cnt:
            avg_s = ms / cnt * 1e-3
            bytes_alg = arrays * 8 * r64["cells"]
            kern[name] = dict(avg_us=avg_s * 1e6, launches=cnt, alg_bytes=bytes_alg,
                              alg_GBs=bytes_alg / avg_s / 1e9)
    dom = max(kern, key=lambda k: kern[k]["avg_us"])
    d = kern[dom]
    inst = kernel_instance(dom, r64["tuning"])
    traffic = PMC_TRAFFIC_MB.get(inst)
    traffic = traffic * 1e6 if (traffic and r64["nx"] == 2048 and r64["nz"] == 1024) else None
    for name, kd in kern.items():
        ki = kernel_instance(name, r64["tuning"])
        t_mb = PMC_TRAFFIC_MB.get(ki)
        kd["instance"] = ki
        if t_mb:
            kd["pmc_traffic_bytes"] = t_mb * 1e6
            kd["pmc_traffic_GBs"] = t_mb * 1e6 / (kd["avg_us"] * 1e-6) / 1e9
    step_alg_GBs = STEP_ARRAYS * 8 * r64["cells"] / (ms_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "eager_ms_per_step": r64["eager_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded KH equilibrium + 1e-9 white noise, generated on the host)",
        "config": {"workload": "plain FP64 single run, 2048x1024, x walled / z periodic, dx=dz=0.1, "
                               "V0=0.05, a=1.0, seed 1, CFL 0.4, no mode filter, xi_cut=12.8",
                   "segments": r64["segments"],
                   "tiles": r64["tuning"],
                   "nx": r64["nx"], "nz": r64["nz"], "precision": "fp64",
                   "sweep": "fused x+z" if r64["tuning"]["fused_sweep"] else "split x / z"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": d["alg_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": d["alg_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_instance": inst, "alg_bytes_per_launch": d["alg_bytes"], "avg_launch_us": d["avg_us"],
                     "traffic_GBs": traffic / (d["avg_us"] * 1e-6) / 1e9 if traffic else None,
                     "limiter": LIMITER.get(dom)},
        "step_roofline": {"alg_bytes_per_cell": STEP_ARRAYS * 8, "achieved_GBs": step_alg_GBs,
                          "frac": step_alg_GBs / HBM_PEAK_GBS},
        "kernels": kern,
        "energy_ratio_end": r64["e_ratio"],
    }
    if not a.no_fp32:
        r32 = run_gpu("fp32", a.steps, a.warmup, fused)
        v32 = r32["cells"] * a.steps / r32["seconds"]
        out["fp32_form"] = {"value": v32, "ms_per_step": r32["seconds"] / a.steps * 1e3,
                            "speedup_vs_fp64": v32 / value,
                            "kernels_avg_us": {k: (v[0] / v[1] * 1e3 if v[1] else None)
                                               for k, v in r32["times"].items()}}
    if not a.no_c3:
        import khs
        c3 = {}
        for prec in ("fp32", "fp64"):
            r = run_gpu(prec, a.c3_steps, 20, fused, WORKLOAD_C3, SEGMENT_C3)
            c3[prec] = dict(value=r["cells"] * a.c3_steps / r["seconds"],
                            ms_per_step=r["seconds"] / a.c3_steps * 1e3,
                            kernels_avg_us={k: (v[0] / v[1] * 1e3 if v[1] else None) for k, v in r["times"].items()})
        t32, t64 = growth_c3("fp32"), growth_c3("fp64")
        g64, t0, t1, r2 = khs.fit_growth_max_r2(t64["t"], t64["amp"])
        g32 = khs.fit_growth_max_r2(t32["t"], t32["amp"], t_min=t0, t_max=t1)[0]
        g64w = khs.fit_growth_max_r2(t64["t"], t64["amp"], t_min=t0, t_max=t1)[0]
        out["precision_form"] = {
            "workload": "4096x2048, V0=0.1, a=0.5, keep |kz|<=8 (MFMA filter), xi_cut=12.8, walled x",
            "fp32_cell_updates_per_s": c3["fp32"]["value"], "fp64_cell_updates_per_s": c3["fp64"]["value"],
            "fp32_over_fp64": c3["fp32"]["value"] / c3["fp64"]["value"],
            "gamma_fp64": g64w, "gamma_fp32": g32, "gamma_rel_err": abs(g32 - g64w) / g64w,
            "energy_ratio_fp64": t64["energy_ratio"], "energy_ratio_fp32": t32["energy_ratio"],
            "energy_drift_err": abs(t32["energy_ratio"] - t64["energy_ratio"]),
            "gauss_rms_fp64": t64["gauss"]["rms_ref"], "gauss_rms_fp32": t32["gauss"]["rms_ref"],
            "gauss_rel_scale": t64["gauss"]["Etot_scale"],
            "fit_window_t": [t0, t1], "steps": a.c3_steps, "segments_of": SEGMENT_C3,
            "kernels_avg_us": {p: c3[p]["kernels_avg_us"] for p in c3}}
    if not a.no_sweep:
        out["sweep"] = sweep()
    if not a.no_c5:
        out["emhd_waves"] = emhd_waves()
    if not a.no_cpu:
        cb = cpu_baseline(a.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
