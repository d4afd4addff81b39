"""Campaigns: sets of runs scheduled over GPUs and executed as captured-graph
batches, each run's KH growth rate measured from in-memory mode amplitudes.

The reference generates one shell script per wave with a here-doc .ini per
run, deals the runs round-robin over 3 GPUs and runs ym_coupled once per
.ini (ymgpu2d/analysis/gen_multigpu_campaign.py:30-235); growth rates are
then fitted from the 6-digit CSV snapshots (analysis/remote_timeseries.py,
batch_analyze.py).  Here:

  * int_kz_keys / half_kz_keys restate the two .ini templates (:89-116,
    :119-168) and xi_sponge_for their sponge radius (:33-37);
  * a manifest (JSON: {"defaults": {...}, "runs": [{...}, ...],
    "sample_stride": S, "lanes": L, "segment": G, "mode": "concurrent" |
    "graph" | "packed"}, or a directory / glob of .ini files) names the runs;
  * lpt_assign deals them to GPUs by longest processing time first (cost =
    cells x steps), instead of round-robin;
  * run_part steps one GPU's share as one khs.Batch (each run stops at its own
    n_steps / target_tu or when it halts; the runs advance concurrently on
    host threads and their own streams, whole energy strides as captured
    graphs; with "mode": "graph" as one graph per segment of all runs; with
    "mode": "packed" every step of all runs as one launch per kernel, for
    runs of one precision whose filters are xi_cut / sponge / suppress_kz0),
    sampling the circular Az2 + i Az3
    amplitude at each run's k_mode every S steps, and fits the growth rate
    (khs.fit_growth_max_r2 and khs.fit_growth_plateau);
  * main() is the command line; with --gpus N > 1 it starts one process per
    GPU (HIP_VISIBLE_DEVICES) and merges their result tables.
"""
import csv
import glob
import json
import math
import os
import subprocess
import sys

import numpy as np

LZ_4PI = 4.0 * math.pi
NZ_4PI = 128


def xi_sponge_for(alpha, V0, kz_phys):
    """Sponge radius just beyond xi_crit of kz_peak ~ 2 alpha, clamped to 5..55."""
    kz_peak_est = max(0.5, round(2 * alpha) * 1.0)
    xi_crit = kz_peak_est / (alpha * V0)
    return max(5, min(55, int(xi_crit * 1.3)))


def int_kz_keys(alpha, V0, k, sponge_kz=None):
    """Integer-kz run (Lz = 2 pi, NZ = 64): the .ini of the reference template;
    the sponge radius is taken for sponge_kz (the largest kz of the combo)."""
    xi = xi_sponge_for(alpha, V0, sponge_kz if sponge_kz is not None else k)
    return dict(k_mode=k, alpha_YM=alpha, V0=V0, perturb_amp=0.001, run_mode=6, xi_sponge=float(xi),
                sigma_sponge=5.0, suppress_kz0=1, hyp_diff=5e-5, kz_suppress_max=k - 1, eps_override=0.15,
                kz_suppress_hi=14, target_tu=100)


def half_kz_keys(alpha, V0, kz_phys):
    """Half-integer kz run (Lz = 4 pi, NZ = 128, k_mode = 2 kz)."""
    k_mode = int(round(2 * kz_phys))
    return dict(k_mode=k_mode, alpha_YM=alpha, V0=V0, perturb_amp=0.001, run_mode=6,
                xi_sponge=float(xi_sponge_for(alpha, V0, kz_phys)), sigma_sponge=5.0, suppress_kz0=1,
                hyp_diff=5e-5, kz_suppress_max=k_mode - 1, eps_override=0.15, kz_suppress_hi=28, target_tu=100,
                lz_override=LZ_4PI, nz_override=NZ_4PI)


def load_manifest(path):
    """-> (list of (name, keys dict or .ini path), options dict)."""
    opts = dict(sample_stride=24, lanes=4, segment=24, mode="concurrent")
    if os.path.isdir(path) or any(ch in path for ch in "*?["):
        files = sorted(glob.glob(os.path.join(path, "*.ini") if os.path.isdir(path) else path))
        return [(os.path.splitext(os.path.basename(f))[0], f) for f in files], opts
    with open(path) as f:
        m = json.load(f)
    defaults = m.get("defaults", {})
    runs = []
    for i, r in enumerate(m["runs"]):
        r = dict(r)
        name = r.pop("name", f"run{i:03d}")
        keys = dict(defaults)
        keys.update(r)
        runs.append((name, keys))
    for k in opts:
        if k in m:
            opts[k] = m[k] if k == "mode" else int(m[k])
    if opts["mode"] not in ("concurrent", "graph", "packed"):
        raise ValueError(f"{path}: mode must be 'concurrent', 'graph' or 'packed', not {opts['mode']!r}")
    return runs, opts


def make_config(spec):
    import khs

    return khs.Config.from_ini(spec) if isinstance(spec, str) else khs.Config(**spec)


def run_cost(cfg):
    return float(cfg.nx) * float(cfg.nz) * float(cfg.total_steps)


def lpt_assign(costs, n_parts):
    """Longest processing time first: each run (largest cost first, index
    order on ties) goes to the part with the least total so far (lowest
    part index on ties).  -> list of n_parts lists of run indices."""
    parts = [[] for _ in range(n_parts)]
    load = [0.0] * n_parts
    for i in sorted(range(len(costs)), key=lambda j: (-costs[j], j)):
        p = min(range(n_parts), key=lambda q: (load[q], q))
        parts[p].append(i)
        load[p] += costs[i]
    return parts


def fit_run(t, amp):
    """Max-R^2 and plateau growth rates of one amplitude series."""
    import khs

    out = dict(gamma_r2=float("nan"), r2=float("nan"), t0=float("nan"), t1=float("nan"),
               gamma_plateau=float("nan"))
    ok = np.isfinite(amp) & (amp > 0)
    if ok.sum() < 8:
        return out
    g, t0, t1, r2 = khs.fit_growth_max_r2(t[ok], amp[ok])
    out.update(gamma_r2=g, r2=r2, t0=t0, t1=t1)
    p = khs.fit_growth_plateau(t[ok], amp[ok])
    if p["has_plateau"]:
        out["gamma_plateau"] = p["gamma"]
    return out


def run_part(runs, opts, out_root):
    """Step `runs` [(name, spec), ...] as one batch; write each run's
    amplitude series to out_root/<name>/timeseries.csv and return result rows."""
    import khs

    cfgs = [make_config(spec) for _, spec in runs]
    if not cfgs:
        return []
    stride = max(2, int(opts["sample_stride"]))
    b = khs.Batch(cfgs, lanes=int(opts["lanes"]), segment=int(opts["segment"]))
    try:
        b.init()
        for i, c in enumerate(cfgs):
            b.set_step_limit(i, c.total_steps)
        n = len(cfgs)
        ts = [[0.0] for _ in range(n)]
        amps = [[b.mode_amplitude(i, cfgs[i].k_mode)[0]] for i in range(n)]
        steps = [0] * n
        halted = [False] * n
        while True:
            live = [i for i in range(n) if not halted[i] and steps[i] < cfgs[i].total_steps]
            if not live:
                break
            mode = opts.get("mode", "concurrent")
            if mode == "graph":
                done, hl = b.run(stride)
            elif mode == "packed":
                done, hl = b.run_packed(stride)
            else:
                done, hl = b.run_concurrent(stride, threads=int(opts["lanes"]))
            for i in live:
                if done[i] != steps[i]:
                    steps[i] = done[i]
                    ts[i].append(done[i] * cfgs[i].dt)
                    amps[i].append(b.mode_amplitude(i, cfgs[i].k_mode)[0])
                halted[i] = hl[i]
        rows = []
        for i, (name, _) in enumerate(runs):
            c = cfgs[i]
            t, a = np.array(ts[i]), np.array(amps[i])
            d = os.path.join(out_root, name)
            os.makedirs(d, exist_ok=True)
            with open(os.path.join(d, "timeseries.csv"), "w", newline="") as f:
                w = csv.writer(f)
                w.writerow(["step", "t", "amp_kz"])
                for tt, aa in zip(t, a):
                    w.writerow([int(round(tt / c.dt)), f"{tt:.6f}", f"{aa:.9e}"])
            E, E0 = b.energy(i)
            row = dict(name=name, k_mode=c.k_mode, alpha=float(c.alpha_YM), V0=float(c.V0), nx=c.nx, nz=c.nz,
                       steps=steps[i], halted=int(halted[i]), t_end=steps[i] * c.dt, E_ratio=E / E0)
            row.update(fit_run(t, a))
            rows.append(row)
        return rows
    finally:
        b.close()


FIELDS = ["name", "k_mode", "alpha", "V0", "nx", "nz", "steps", "halted", "t_end", "E_ratio", "gamma_r2", "r2",
          "t0", "t1", "gamma_plateau"]


def write_results(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow({k: r.get(k) for k in FIELDS})


def read_results(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description="Run a campaign manifest as batched graph runs")
    ap.add_argument("manifest", help="JSON manifest, a directory of .ini files or a glob")
    ap.add_argument("--out", default="campaign_out")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--part", type=int, default=-1, help="(internal) run part i of --gpus")
    a = ap.parse_args(argv)
    runs, opts = load_manifest(a.manifest)
    os.makedirs(a.out, exist_ok=True)
    costs = [run_cost(make_config(spec)) for _, spec in runs]
    parts = lpt_assign(costs, max(1, a.gpus))
    if a.part >= 0:
        rows = run_part([runs[i] for i in parts[a.part]], opts, a.out)
        write_results(os.path.join(a.out, f"results_part{a.part}.csv"), rows)
        return 0
    if a.gpus <= 1:
        rows = run_part(runs, opts, a.out)
    else:
        procs = []
        for g in range(a.gpus):
            env = dict(os.environ, HIP_VISIBLE_DEVICES=str(g))
            procs.append(subprocess.Popen([sys.executable, "-m", "khs.campaign", a.manifest, "--out", a.out,
                                           "--gpus", str(a.gpus), "--part", str(g)], env=env))
        rc = [p.wait() for p in procs]
        if any(rc):
            print(f"campaign parts failed: {rc}", file=sys.stderr)
            return 1
        rows = []
        for g in range(a.gpus):
            rows += read_results(os.path.join(a.out, f"results_part{g}.csv"))
    write_results(os.path.join(a.out, "results.csv"), rows)
    for r in rows:
        print(f"{r['name']:>24} kz={r['k_mode']} gamma={float(r['gamma_r2']):.5f} steps={r['steps']} "
              f"halted={r['halted']}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
