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
physically finite.

Legs beside the headline (each reported in the same JSON line):
  * fp32_form: the headline workload in the FP32 precision form;
  * c1: 256x128 FP64, walls, keep |kz| <= 4, 2000 steps, with the reference's
    production window (xi_cut = 5, suppress_kz0 = 1): throughput over the
    2000 steps; the kz = 1 growth rate (plateau of the local rate over a
    4000-step run: in the first 2000 steps the amplitude still beats between
    the top eigenmodes) against the eigensolver restatement, energy drift;
  * precision_form (C3): 4096x2048, V0 = 0.1, a = 0.5, keep |kz| <= 8 on MFMA,
    same production window, FP32 and FP64 throughput, growth-rate, energy and
    Gauss-residual errors of FP32 against FP64 over the stable part of the run;
  * sweep (C4): 64 uneven FP32 runs of 3000 steps (runs/hour);
  * emhd_waves (C5): EMHD 1024x1024 FP32 waves and their MFMA dispersion;
  * cpu_baseline: the FP64 OpenMP oracle on the host cores.

Prints one JSON line (one GPU: N = 1 only).  --growth-refinement instead
prints the growth-rate refinement study of C1 and of C3's physics (dx = 0.1,
0.05, 0.025; about a minute).
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
# a plain 256 MiB device-to-device copy on MI355X (bench.py --calibrate-copy 256)
COPY_GBS = 5300.0

WORKLOAD = dict(nx_override=2048, nz_override=1024, lx_override=204.8, lz_override=102.4, alpha_YM=1,
                V0=0.05, eps_override=1.0, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1,
                perturb_amp=2e-5, noise_amp=1e-9, noise_seed=1, n_steps=5000, xi_cut=12.8)
SEGMENT = 700
# The reference's production window for the colour-2/3 sector: xi_cut = 5 and
# the kz = 0 removal (ymgpu2d/OUTER_REGION.md: the cut inside which the frozen
# background's tachyonic branch stays out, suppress_kz0 "mandatory in
# production").  Inside it, at the C1 and C3 geometries, the kz = 1 mode is
# the fastest linear eigenmode (eigensolver restatement with the same Dirichlet
# rows), so its growth rate is what the runs measure.
PRODUCTION_WINDOW = dict(xi_cut=5.0, suppress_kz0=1)
# C1: 256x128 FP64, walls, keep |kz| <= 4, 2000 steps (reference-scale run)
WORKLOAD_C1 = dict(nx_override=256, nz_override=128, lx_override=25.6, lz_override=12.8, alpha_YM=1, V0=0.05,
                   eps_override=1.0, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1,
                   perturb_amp=2e-5, noise_amp=1e-9, noise_seed=0, n_steps=2000, kz_keep_max=4,
                   **PRODUCTION_WINDOW)
# C3 precision-form workload: 4096x2048, V0 = 0.1, a = 0.5, keep |kz| <= 8
# (the column filter runs as MFMA contractions).  The run is stable to step
# ~4170 (t ~ 167): at V0 >= 0.08 the saturated shear-layer mode cavitates
# the beams and the run ends on the energy rule, the reference's documented
# "V0 >= 0.08 hard wall" (OUTER_REGION.md), so the configured 10000 steps are
# not reachable.  Throughput is timed over segments of SEGMENT_C3 steps from
# the initial state; the diagnostics compare FP32 with FP64 at C3_DIAG_STEPS.
WORKLOAD_C3 = dict(nx_override=4096, nz_override=2048, lx_override=409.6, lz_override=204.8, alpha_YM=1,
                   V0=0.1, eps_override=0.5, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1,
                   perturb_amp=1e-5, noise_amp=1e-9, noise_seed=1, n_steps=10000, kz_keep_max=8,
                   **PRODUCTION_WINDOW)
C1_DIAG_STEPS = 4000
SEGMENT_C3 = 2000
C3_DIAG_STEPS = 3500

# Compulsory (algorithmic) HBM traffic of each step kernel, in arrays per cell:
# every input array read once and every output written once, no halo re-reads.
#   k_fct_x : R fluid 12 + Ex 3 + By 3,            W stage 12
#   k_fct_z : R stage 12 + step-start fluid 12 + Ez 3 + By 3 + Az 3, W fluid 12
#   k_maxwell: R new fluid 12 + fields 12,         W E 6 + By 3 + Az2,3 2 (Az1 frozen)
#   k_fct_xz: R fluid 12 + Ex, Ez, By, Az 12,       W fluid 12
#   k_mxz   : R fluid 12 + fields 12,               W fields 11 + fluid 12 (Maxwell + the next sweeps)
KERNEL_ARRAYS = {"k_fct_x": 18 + 12, "k_fct_z": 33 + 12, "k_fct_xz": 24 + 12, "k_maxwell": 24 + 11,
                 "k_fct_zm": 36 + 23, "k_mxz": 24 + 23}
# The whole step's algorithmic minimum: R 24 + W 23 arrays (frozen Az1 read only).
STEP_ARRAYS = 47
ZM_ROWS = {1: 16, 2: 32, 3: 64}       # chunk height of the fused z sweep + Maxwell variants
# What limits a kernel other than its HBM bytes, from rocprofv3 SQ counters and
# the decomposition of the fused sweep (profiles/sq_counters_2048x1024_fp64.md,
# profiles/fused_sweep_decomposition_2048x1024_fp64.md), keyed by (slot,
# precision, nx, nz).
LIMITER = {
    ("k_mxz", "fp64", 2048, 1024):
        "issue and dependency latency of the sweep wavefronts' row march: two FP64 waves per SIMD (one sweep, one "
        "Maxwell wave, 252 VGPRs), SIMD VALU busy 0.50 of the launch, each wave waiting on memory and LDS counters "
        "for 0.34 of its lifetime (profiles/sq_counters_2048x1024_fp64.md); HBM traffic 1.07x algorithmic at "
        "~4.1 TB/s, under the ~5.3 TB/s of a plain device copy",
    ("k_fct_xz", "fp64", 2048, 1024):
        "memory-level parallelism: ~13,700 HBM reads (~1.5 MB) in flight at a 0.63 us read latency, a third of "
        "what 8 TB/s needs (profiles/latency_icache_2048x1024_fp64.md), because each wave holds one row of loads "
        "ahead in registers at 2 waves per SIMD (~190 VGPRs); its access pattern alone (the same loads, lane "
        "shifts and stores with a trivial sum) takes 174-181 us, ~80% of the kernel's time, and the rest is FP64 "
        "arithmetic that 2 waves per SIMD do not hide (VALU busy 0.43)",
    ("k_maxwell", "fp64", 2048, 1024): "HBM: 620 MB per launch (1.06x algorithmic) at 5.0 TB/s, ~21,500 reads "
                                       "in flight (profiles/latency_icache_2048x1024_fp64.md)",
    ("k_fct_xz", "fp32", 4096, 2048):
        "the ring sweep's FP32 arithmetic: the same instruction count as the FP64 form on half the bytes, its "
        "sweep wavefronts' row march issue-bound as in FP64",
    ("k_fct_xz", "fp64", 4096, 2048):
        "issue latency of the ring sweep's sweep wavefronts (k_mxz<double, 4, false>: four row-loading wavefronts "
        "fill the LDS ring, so the sweep wavefronts wait on memory for 0.14 of their lifetime against 0.48 in "
        "k_fct_xzg; VALU busy 0.48 at 2048x1024, profiles/sq_counters_2048x1024_fp64.md)",
    ("k_fct_xz", "fp64", 256, 128):
        "latency: 640 waves of 10 rows each on a 256x128 grid (under a third of the wave slots at 2 per SIMD), each row "
        "waiting on the loads issued one row earlier",
}


# launch shapes of k_mxz (one kernel name per Maxwell wave count: the chunk
# height is a run-time argument); the committed traffic tables label them so
MXZ_SHAPES = {0: "v0, 16-row chunks", 1: "v1, 32-row chunks", 2: "v2, 64-row chunks", 3: "v3, 16-row chunks",
              4: "v4, 32-row chunks", 5: "v5, 64-row chunks", 6: "v6, one round", 7: "v7, two rounds",
              8: "v8, one round"}


def kernel_instance(slot, t, T="double"):
    """Kernel instance (template arguments) that the tuned slot launches."""
    if slot == "k_fct_x":
        if t.get("x_lane_cells", 0) > 0:
            return f"k_fct_xr<{T}, {t['x_lane_cells']}>"
        return f"k_fct_x<{T}, {t['x_segment']}, YMModel, {t['x_serial']}>"
    if slot == "k_fct_z":
        name = "k_fct_zg" if t["z_groups"] else "k_fct_z"
        return f"{name}<{T}, {t['z_chunk']}, {t['z_prefetch']}, YMModel>"
    if slot == "k_fct_zm":
        return f"k_fct_zmx<{T}, {ZM_ROWS.get(t.get('zm_fused', 0), 0)}>"
    if slot == "k_fct_xz":
        v = t["xz_variant"]
        name, tz = {0: ("k_fct_xz", 32), 1: ("k_fct_xz2", 16), 2: ("k_fct_xz2", 32), 3: ("k_fct_xz2", 64),
                    4: ("k_fct_xzr", 32), 5: ("k_fct_xzr", 64), 6: ("k_fct_xzr", 128), 7: ("k_fct_xzg", 16),
                    8: ("k_fct_xzg", 32), 9: ("k_fct_xzg", 64), 10: ("k_fct_xzg", 32),
                    11: ("k_fct_xzg", 8), 12: ("k_fct_xzg", 4)}.get(v, (None, 0))
        if v in (13, 14):
            return f"k_mxz<{T}, 4, false> ring sweep, {('one', 'two')[v - 13]} round{'s' * (v - 13)}"
        if name is None:
            return None
        if name == "k_fct_xzg":
            return f"{name}<{T}, {tz}, {128 if v >= 10 else 256}>"
        return f"{name}<{T}, {tz}>"
    if slot == "k_maxwell" and t["maxwell"] >= 4:
        tzpf = {4: (16, 1), 5: (32, 1), 6: (16, 0), 7: (8, 1), 8: (8, 0), 9: (4, 1), 10: (2, 1)}.get(t["maxwell"])
        return f"k_maxwell_zm<{T}, {tzpf[0]}, {tzpf[1]}>" if tzpf else None
    if slot == "k_mxz":
        v = t.get("mxz_variant", 1)
        name = f"k_mxz<{T}, {(2, 2, 2, 4, 4, 4, 4, 4, 2)[v]}>"
        return name if v == 6 else f"{name} {MXZ_SHAPES[v]}"
    if slot == "k_maxwell":
        return {0: f"k_maxwell<{T}, 16>", 1: f"k_maxwell_cell<{T}>", 2: f"k_maxwell<{T}, 4>",
                3: f"k_maxwell<{T}, 8>"}.get(t["maxwell"])
    return None


def pmc_traffic(nx, nz, precision):
    """HBM bytes per launch of each kernel instance, read from the committed
    rocprofv3 summary of this grid and precision (FETCH_SIZE x 2 + WRITE_SIZE
    x 1, the device-copy calibration): {instance: bytes}, empty when no
    summary covers the configuration."""
    path = os.path.join(REPO, "profiles", f"hbm_traffic_{nx}x{nz}_{precision}.md")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 4 and cells[0].startswith("`"):
                try:
                    out[cells[0].replace("`", "")] = float(cells[3]) * 1e6
                except ValueError:
                    pass
    return out


def window_fractions(cfg):
    """Fractions of columns that the xi_cut zeroes and that the sponge damps
    (YM_Physics.cu:161-210: the cut wins where both apply)."""
    import numpy as np
    x = np.arange(cfg.nx) * cfg.dx
    xi = np.abs((x - cfg.nx * cfg.dx / 2) / cfg.eps)
    cut = (xi > float(cfg.xi_cut)) if cfg.xi_cut > 0 else np.zeros(cfg.nx, bool)
    sponge = ((xi > float(cfg.xi_sponge)) & ~cut) if cfg.xi_sponge > 0 else np.zeros(cfg.nx, bool)
    return float(np.mean(cut)), float(np.mean(sponge))


def filter_arrays(cfg):
    """Algorithmic arrays per cell of the filter stage (each array read once,
    written once): the charges Q2/Q3 of both species zeroed on cut columns
    (written) or damped on sponge columns (read and written), the z-mean of
    the 12 colour-2/3 arrays and colour-1 zeroing (suppress_kz0), the keep
    projection of the 21 filtered arrays (18 when colour 1 is zeroed anyway)."""
    cut, sponge = window_fractions(cfg)
    a = 4.0 * cut + 8.0 * sponge
    if cfg.suppress_kz0:
        a += 24 + 3
    if cfg.kz_keep_max > 0:
        a += 2 * (18 if cfg.suppress_kz0 else 21)
    return a


def kernel_roofline(r, cfg, precision):
    """Per-slot algorithmic bytes and rates of a run_gpu() result, and the
    roofline object of the slot with the longest launches."""
    esz = 4 if precision == "fp32" else 8
    T = "float" if precision == "fp32" else "double"
    arrays_of = dict(KERNEL_ARRAYS, filters=filter_arrays(cfg))
    traffic_tab = pmc_traffic(cfg.nx, cfg.nz, precision)
    kern = {}
    for name, arrays in arrays_of.items():
        ms, cnt = r["times"].get(name, (0.0, 0))
        if cnt and arrays > 0:
            avg_s = ms / cnt * 1e-3
            bytes_alg = arrays * esz * r["cells"]
            inst = kernel_instance(name, r["tuning"], T)
            kd = dict(avg_us=avg_s * 1e6, launches=cnt, alg_bytes=bytes_alg, alg_GBs=bytes_alg / avg_s / 1e9,
                      instance=inst)
            if inst in traffic_tab:
                kd["pmc_traffic_bytes"] = traffic_tab[inst]
                kd["pmc_traffic_GBs"] = traffic_tab[inst] / avg_s / 1e9
            kern[name] = kd
    # the dominant single kernel: "filters" times a stage of several launches
    singles = [k for k in kern if k != "filters"] or list(kern)
    dom = max(singles, key=lambda k: kern[k]["avg_us"])
    d = kern[dom]
    roof = {"bound": "hbm", "kernel": dom, "achieved": d["alg_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": d["alg_GBs"] / HBM_PEAK_GBS, "traffic": d.get("pmc_traffic_bytes"),
            "kernel_instance": d["instance"], "alg_bytes_per_launch": d["alg_bytes"], "avg_launch_us": d["avg_us"],
            "traffic_GBs": d.get("pmc_traffic_GBs"),
            "traffic_source": (f"profiles/hbm_traffic_{cfg.nx}x{cfg.nz}_{precision}.md"
                               if d.get("pmc_traffic_bytes") else None),
            "limiter": LIMITER.get((dom, precision, cfg.nx, cfg.nz))}
    # what holds the kernel below the roofline: its measured HBM traffic rate
    # against what a plain copy reaches ("hbm" within 15% of it), else the
    # issue / memory latency its counters show
    tg = d.get("pmc_traffic_GBs")
    roof["limited_by"] = None if tg is None else ("hbm" if tg >= 0.85 * COPY_GBS else "latency")
    return kern, roof


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


def growth_run(workload, precision, n_steps, stride):
    """kz = 1 circular Az2/Az3 amplitude sampled every `stride` steps from the
    initial state, with E/E0 and the Gauss residual at the end."""
    import khs

    cfg = khs.Config(**workload, precision=precision)
    sim = khs.Simulation(cfg)
    sim.init()
    ts = sim.timeseries(1, n_steps, stride)
    ts["energy_ratio"] = sim.energy() / sim.energy0
    ts["gauss"] = sim.gauss()
    sim.close()
    if ts["halted"]:
        raise RuntimeError("growth-rate run halted")
    return ts


def eigen_rates(workload, n=3):
    """The n fastest linear growth rates of the kz = 1 colour-2/3 mode at the
    workload's geometry (khs.eigenmode, Dirichlet rows outside xi_cut)."""
    import numpy as np
    from khs import eigenmode

    w = workload
    kz = 2 * np.pi * w["k_mode"] / w["lz_override"]
    vals, _ = eigenmode.solve(kz, float(w["alpha_YM"]), float(w["V0"]), float(w["eps_override"]),
                              int(w["nx_override"]), Lx=float(w["lx_override"]), n_eigs=12, sigma=0.4,
                              xi_cut=w.get("xi_cut") or None)
    return [float(v.real) for v in vals[:n]]


def growth_compare(workload, n_steps, stride, t_max=None):
    """FP64 and FP32 runs of the workload: max-R^2 kz = 1 growth rate (FP32
    fitted over the FP64 window), energy drift and Gauss residuals."""
    import khs

    import numpy as np

    t64, t32 = growth_run(workload, "fp64", n_steps, stride), growth_run(workload, "fp32", n_steps, stride)
    g64max, t0, t1, r2 = khs.fit_growth_max_r2(t64["t"][1:], t64["amp"][1:], t_max=t_max)

    def slope(ts):                   # least-squares rate of log(amp) over exactly [t0, t1]
        m = (ts["t"] >= t0) & (ts["t"] <= t1)
        return float(np.polyfit(ts["t"][m], np.log(ts["amp"][m]), 1)[0])

    g64, g32 = slope(t64), slope(t32)
    p64 = khs.fit_growth_plateau(t64["t"][1:], t64["amp"][1:])
    p32 = khs.fit_growth_plateau(t32["t"][1:], t32["amp"][1:])
    eig = eigen_rates(workload)
    return {"gamma_fp64": g64, "gamma_fp32": g32, "gamma_rel_err": abs(g32 - g64) / g64, "fit_window_t": [t0, t1],
            "gamma_fit": "max-R^2 window of the FP64 run (batch_analyze.py:44-69), then the least-squares rate of "
                         "log(amp) of each run over exactly that window",
            "gamma_fp64_max_r2": g64max,
            "fit_r2": r2, "gamma_eigensolver": eig[0], "gamma_eigensolver_next": eig[1:],
            "gamma_fp64_over_eigensolver": g64 / eig[0], "diag_steps": n_steps,
            "gamma_plateau_fp64": p64["gamma"], "gamma_plateau_fp32": p32["gamma"],
            "plateau_window_t": [p64["t0"], p64["t1"]],
            "gamma_plateau_fit": "the flat stretch of the local growth rate (batch_analyze.py:90-152)",
            "gamma_plateau_fp64_over_eigensolver": p64["gamma"] / eig[0],
            "energy_ratio_fp64": t64["energy_ratio"], "energy_ratio_fp32": t32["energy_ratio"],
            "energy_drift_fp64": t64["energy_ratio"] - 1.0,
            "energy_drift_err": abs(t32["energy_ratio"] - t64["energy_ratio"]),
            "gauss_rms_fp64": t64["gauss"]["rms_ref"], "gauss_rms_fp32": t32["gauss"]["rms_ref"],
            "gauss_rel_err": [abs(x - y) / max(y, 1e-300) for x, y in zip(t32["gauss"]["rms_ref"],
                                                                          t64["gauss"]["rms_ref"])],
            "gauss_rel_scale": t64["gauss"]["Etot_scale"]}


def c1_leg(fused=-1):
    """C1: 256x128 FP64, walls, keep |kz| <= 4, production window: the whole
    2000-step run timed (graph-replayed strides), per-kernel times, the kz = 1
    growth rate against the eigensolver and the energy drift."""
    import khs

    cfg = khs.Config(**WORKLOAD_C1)
    r = run_gpu("fp64", cfg.total_steps, 20, fused, WORKLOAD_C1, cfg.total_steps, kernel_steps=200)
    kern, roof = kernel_roofline(r, cfg, "fp64")
    # growth rate: over the timed 2000 steps (t = 80) the kz = 1 amplitude
    # still beats between the top eigenmodes (0.089, 0.064, 0.054), so the
    # diagnostic run continues to C1_DIAG_STEPS (t = 160, amplitude ~3e-2) and
    # the rate is the plateau of the local rate (batch_analyze.py:90-152)
    # the energy drift and the Gauss residual are read at the configured 2000
    # steps; the same run then continues for the plateau
    import numpy as np

    sim = khs.Simulation(khs.Config(**WORKLOAD_C1))
    sim.init()
    first = sim.timeseries(1, cfg.total_steps, 10)
    drift, gauss = sim.energy() / sim.energy0 - 1.0, sim.gauss()["rms_ref"]
    rest = sim.timeseries(1, C1_DIAG_STEPS - cfg.total_steps, 10)
    sim.close()
    if first["halted"] or rest["halted"]:
        raise RuntimeError("C1 growth-rate run halted")
    t = np.concatenate([first["t"], rest["t"][1:]])
    amp = np.concatenate([first["amp"], rest["amp"][1:]])
    g, t0, t1, r2 = khs.fit_growth_max_r2(first["t"][1:], first["amp"][1:])
    pl = khs.fit_growth_plateau(t[1:], amp[1:])
    eig = eigen_rates(WORKLOAD_C1)
    return {"workload": "256x128 FP64, x walled, dx=dz=0.1, V0=0.05, a=1.0, kz=1 seed 1e-6 + 1e-9 noise (seed 0), "
                        "CFL 0.4, keep |kz|<=4, xi_cut=5, suppress_kz0=1, 2000 steps",
            "cell_updates_per_s": r["cells"] * cfg.total_steps / r["seconds"],
            "ms_per_step": r["seconds"] / cfg.total_steps * 1e3, "steps": cfg.total_steps,
            "roofline": roof, "kernels": kern, "tiles": r["tuning"],
            "gamma_plateau": pl["gamma"], "plateau_window_t": [pl["t0"], pl["t1"]],
            "gamma_plateau_over_eigensolver": pl["gamma"] / eig[0],
            "gamma_first_2000_steps": g, "fit_window_t": [t0, t1], "fit_r2": r2,
            "gamma_first_2000_over_eigensolver": g / eig[0],
            "gamma_eigensolver": eig[0], "gamma_eigensolver_next": eig[1:], "diag_steps": C1_DIAG_STEPS,
            "energy_drift": drift, "gauss_rms": gauss, "diagnostics_at_step": cfg.total_steps}


def growth_refinement():
    """The kz = 1 growth rate of C1 and of C3's physics under grid refinement
    (dx = dz = 0.1, 0.05, 0.025, steps scaled to the same time), each against
    the eigensolver at the same NX: C1 as configured over t = 160; C3's
    parameters on Lx = 25.6 instead of 409.6 (the xi_cut window |x| < 2.5
    holds every colour-2/3 field; at dx = 0.1 the plateau is C3's to 1e-7)
    over t = 140.  Rates are plateau fits (batch_analyze.py:90-152)."""
    import khs
    import numpy as np

    out = {}
    for name, base, extra, steps, stride in (
            ("c1", WORKLOAD_C1, {}, C1_DIAG_STEPS, 10),
            ("c3_physics_lx25.6", WORKLOAD_C3, dict(nx_override=256, lx_override=25.6), C3_DIAG_STEPS, 25)):
        rows = []
        for f in (1, 2, 4):
            w = dict(base, **extra)
            w.update(nx_override=w["nx_override"] * f, nz_override=w["nz_override"] * f)
            sim = khs.Simulation(khs.Config(**w))
            sim.init()
            ts = sim.timeseries(1, steps * f, stride * f)
            sim.close()
            t, a = np.asarray(ts["t"]), np.asarray(ts["amp"])
            pl = khs.fit_growth_plateau(t[1:], a[1:])
            eig = eigen_rates(w)
            rows.append({"dx": 0.1 / f, "nx": w["nx_override"], "nz": w["nz_override"], "steps": steps * f,
                         "halted": bool(ts["halted"]), "gamma_plateau": pl["gamma"],
                         "plateau_window_t": [pl["t0"], pl["t1"]], "gamma_eigensolver": eig[0],
                         "ratio": pl["gamma"] / eig[0]})
            print(json.dumps({"refinement": name, **rows[-1]}), file=sys.stderr, flush=True)
        out[name] = rows
    return out


def sweep(n=64, steps=3000, lanes=8, segment=50, threads=16):
    """Batched-sweep workload: n uneven-grid FP32 runs (khs.sweep) stepped
    four ways: every step of all runs as three shared launches, segments of
    `segment` steps captured as one graph ("packed"); all runs' segments as
    one captured graph of per-run launches over `lanes` streams ("graph");
    `threads` host threads each stepping its share of the runs on their own
    streams with graph-replayed energy strides ("concurrent"); the runs one
    after another ("eager").  A run ends at `steps` or when it halts (E/E0
    rule)."""
    import khs
    from khs.sweep import sweep_configs

    cfgs = sweep_configs(n, steps)
    cells = [c["nx_override"] * c["nz_override"] for c in cfgs]
    out = {"workload": f"{n} runs, NX in 256..2048 x NZ in 128..1024, V0~U[0.01,0.2], a~U[0.5,2], kz 1..3, "
                       f"half with suppress_kz0, FP32, {steps} steps each (a run that halts on the E/E0 rule "
                       f"is counted in 'halted'), CFL 0.1, xi_cut=4",
           "lanes": lanes, "segment": segment, "threads": threads,
           "cell_steps": sum(cells) * steps}
    for mode in ("packed", "graph", "concurrent", "eager"):
        b = khs.Batch(cfgs, lanes=lanes, segment=segment)
        b.init()
        run = {"packed": b.run_packed, "graph": b.run, "eager": b.run_eager,
               "concurrent": lambda k: b.run_concurrent(k, threads=threads)}[mode]
        t0 = time.perf_counter()
        done, halted = run(steps)
        el = time.perf_counter() - t0
        b.close()
        out[mode] = {"seconds": el, "runs_per_hour": n / el * 3600.0, "halted": int(sum(halted)),
                     "completed": int(sum(d >= steps for d in done)),
                     "cell_updates_per_s": sum(c * d for c, d in zip(cells, done)) / el}
    out["graph_over_eager"] = out["eager"]["seconds"] / out["graph"]["seconds"]
    out["packed_over_eager"] = out["eager"]["seconds"] / out["packed"]["seconds"]
    out["concurrent_over_eager"] = out["eager"]["seconds"] / out["concurrent"]["seconds"]
    best = min(("packed", "graph", "concurrent"), key=lambda m: out[m]["seconds"])
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

    if leg in ("sweep", "sweep-graph", "sweep-packed"):
        from khs.sweep import sweep_configs

        b = khs.Batch(sweep_configs(64, steps), lanes=8, segment=50)
        b.init()
        t0 = time.perf_counter()
        run = {"sweep": lambda k: b.run_concurrent(k, threads=16), "sweep-graph": b.run,
               "sweep-packed": b.run_packed}[leg]
        done, halted = run(steps)
        el = time.perf_counter() - t0
        b.close()
        return {"leg": leg, "runs": 64, "steps": steps, "seconds": el, "halted": int(sum(halted)),
                "run_steps": int(sum(done))}
    if leg == "c1":
        r = run_gpu("fp64", steps, warmup, fused, WORKLOAD_C1, steps, kernel_steps=steps)
        return {"leg": "c1", "workload": "256x128 FP64 keep |kz|<=4, production window", "steps": steps,
                "ms_per_step": r["seconds"] / steps * 1e3, "tiles": r["tuning"],
                "kernels_avg_us": {k: v[0] / v[1] * 1e3 for k, v in r["times"].items() if v[1]}}
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
    """The FP64 OpenMP oracle behind the library's C ABI (oracle/libkhs_cpu.so)
    on the host cores, headline workload, bounded sample.  Runs in a child
    process so that the OpenMP placement (OMP_PROC_BIND=close,
    OMP_PLACES=cores) is set before the OpenMP runtime starts."""
    import subprocess

    allowed = len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", KHS_BASELINE_CPUS=str(allowed))
    env.setdefault("OMP_NUM_THREADS", str(allowed))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", str(budget_s)], env=env,
                       capture_output=True, text=True, timeout=budget_s * 10 + 300)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline failed: " + r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def _host_cpu():
    """lscpu's model name and physical cores (core, socket pairs)."""
    import subprocess

    model, cores = None, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), None)
        pairs = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True).stdout
        cores = len({l for l in pairs.splitlines() if l and not l.startswith("#")})
    except (OSError, StopIteration):
        pass
    return model, cores


def _cpu_baseline_child(budget_s):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import khs
    import oracle_ref as orc

    khs._lib = khs.load_library(os.path.join(REPO, "oracle", "libkhs_cpu.so"))
    cfg = khs.Config(**WORKLOAD)
    sim = khs.Simulation(cfg)
    sim.init()
    sim.step(1)                        # warm-up (page-in, thread pool)
    n, t0 = 0, time.perf_counter()
    while True:
        sim.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 400:
            break
    sim.close()
    cells = cfg.nx * cfg.nz
    threads = int(os.environ["OMP_NUM_THREADS"])
    model, phys = _host_cpu()
    allowed = int(os.environ.get("KHS_BASELINE_CPUS", "0")) or None
    triad = orc.stream_triad()
    value = cells * n / el
    print(json.dumps(dict(
        value=value, unit="cell-updates/s", cores=threads, kind="port",
        sample=f"{n} FP64 steps of the 2048x1024 workload through libkhs_cpu.so's khs_step (energy checks at the "
               f"reference cadence included), {el:.1f} s wall, OMP_NUM_THREADS={threads}, OMP_PROC_BIND=close, "
               f"OMP_PLACES=cores",
        cpu_model=model, physical_cores=phys, cpus_allowed=allowed, compiler="g++ -O2 -fopenmp -ffp-contract=off",
        stream_triad_GBs=triad, roofline_frac=(value * STEP_ARRAYS * 8 / 1e9 / triad) if triad else None)))


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
    ap.add_argument("--c3-steps", type=int, default=2000)
    ap.add_argument("--no-c1", action="store_true", help="skip the 256x128 C1 leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep (runs/hour) leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the EMHD wave-dispersion leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--growth-refinement", action="store_true",
                    help="only the growth-rate refinement study of C1 and C3's physics (one JSON object)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB")
    ap.add_argument("--profile-leg", choices=["c1", "c3", "c5", "sweep", "sweep-graph", "sweep-packed"], default=None,
                    help="run one short leg for a profiler and print its kernel times: c1 = the 256x128 run, c3 = the 4096x2048 "
                         "precision form (FP32 and FP64, MFMA keep filter), c5 = the EMHD 1024^2 run with "
                         "its MFMA space-time spectrum every 20 steps, sweep / sweep-graph = the 64-run "
                         "sweep in the concurrent / single-thread graph / packed mode (--steps steps per run)")
    ap.add_argument("--split", action="store_true", help="force the separate x and z sweep kernels")
    ap.add_argument("--fused", action="store_true", help="force the fused x+z sweep kernel")
    ap.add_argument("--cpu-baseline-child", type=float, default=0.0, help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.cpu_baseline_child:
        _cpu_baseline_child(a.cpu_baseline_child)
        return
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
    if a.growth_refinement:
        print(json.dumps(growth_refinement()))
        return
    import khs

    r64 = run_gpu("fp64", a.steps, a.warmup, fused)
    ms_step = r64["seconds"] / a.steps * 1e3
    value = r64["cells"] * a.steps / r64["seconds"]
    # per-kernel live roofline (HIP events on the library's stream)
    cfg64 = khs.Config(**WORKLOAD)
    kern, roof = kernel_roofline(r64, cfg64, "fp64")
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
        "roofline": roof,
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
    if not a.no_c1:
        out["c1"] = c1_leg(fused)
    if not a.no_c3:
        c3 = {}
        cfg3 = khs.Config(**WORKLOAD_C3)
        for prec in ("fp32", "fp64"):
            r = run_gpu(prec, a.c3_steps, 20, fused, WORKLOAD_C3, SEGMENT_C3)
            kern3, roof3 = kernel_roofline(r, cfg3, prec)
            c3[prec] = dict(value=r["cells"] * a.c3_steps / r["seconds"], ms_per_step=r["seconds"] / a.c3_steps * 1e3,
                            roofline=roof3, kernels=kern3, tiles=r["tuning"])
        out["precision_form"] = {
            "workload": "4096x2048, V0=0.1, a=0.5, keep |kz|<=8 (MFMA filter), xi_cut=5, suppress_kz0=1, walled x; "
                        f"FP32 bulk + FP64 reductions vs plain FP64; {a.c3_steps} steps timed per precision "
                        f"(segments of {SEGMENT_C3} from init); diagnostics at step {C3_DIAG_STEPS} "
                        "(stable to ~4170 steps: the V0 >= 0.08 cavitation of the reference's OUTER_REGION.md)",
            "fp32_cell_updates_per_s": c3["fp32"]["value"], "fp64_cell_updates_per_s": c3["fp64"]["value"],
            "fp32_over_fp64": c3["fp32"]["value"] / c3["fp64"]["value"],
            "fp32_ms_per_step": c3["fp32"]["ms_per_step"], "fp64_ms_per_step": c3["fp64"]["ms_per_step"],
            "roofline": c3["fp32"]["roofline"], "roofline_fp64": c3["fp64"]["roofline"],
            "kernels": {p: c3[p]["kernels"] for p in c3}, "tiles": {p: c3[p]["tiles"] for p in c3},
            "steps": a.c3_steps, "segments_of": SEGMENT_C3,
            **growth_compare(WORKLOAD_C3, C3_DIAG_STEPS, 25)}
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
