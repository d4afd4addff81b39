"""A campaign part run as one captured-graph batch (khs.campaign.run_part):
every run stops at its own n_steps, and its sampled mode-amplitude series is
the one of the same run stepped alone (khs.Simulation.timeseries)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import khs
from khs import campaign as C

pytestmark = pytest.mark.gpu

BASE = dict(lz_override=6.4, nz_override=64, alpha_YM=1.0, V0=0.05, eps_override=1.0, courant_override=0.4,
            run_mode=0, freeze_override=1, perturb_amp=0.01, noise_amp=1e-9, xi_sponge=4.0, kz_keep_max=4)
RUNS = [dict(name="w128_k1", nx_override=128, lx_override=12.8, k_mode=1, n_steps=44, noise_seed=1),
        dict(name="w96_k2", nx_override=96, lx_override=9.6, k_mode=2, n_steps=30, noise_seed=2),
        dict(name="w160_k1", nx_override=160, lx_override=16.0, k_mode=1, n_steps=52, noise_seed=3,
             precision="fp32")]


def manifest(tmp_path, mode="concurrent"):
    m = {"defaults": BASE, "runs": RUNS, "sample_stride": 10, "lanes": 2, "segment": 10, "mode": mode}
    p = tmp_path / "campaign.json"
    p.write_text(json.dumps(m))
    return str(p)


@pytest.mark.parametrize("mode", ["concurrent", "graph"])
def test_campaign_part_matches_single_runs(tmp_path, mode):
    runs, opts = C.load_manifest(manifest(tmp_path, mode))
    assert opts["mode"] == mode
    out = str(tmp_path / "out")
    rows = C.run_part(runs, opts, out)
    assert [r["name"] for r in rows] == [r["name"] for r in RUNS]
    for spec, row in zip(RUNS, rows):
        keys = dict(BASE, **{k: v for k, v in spec.items() if k != "name"})
        print("campaign run", spec["name"], flush=True)
        assert row["steps"] == spec["n_steps"] and row["halted"] == 0
        sim = khs.Simulation(khs.Config(**keys))
        sim.init()
        ts = sim.timeseries(spec["k_mode"], spec["n_steps"], 10)
        sim.close()
        data = np.genfromtxt(os.path.join(out, spec["name"], "timeseries.csv"), delimiter=",", names=True)
        assert list(data["step"].astype(int)) == [round(t / khs.Config(**keys).dt) for t in ts["t"]]
        np.testing.assert_array_equal(data["amp_kz"], np.array([float(f"{a:.9e}") for a in ts["amp"]]))


# packed launches take runs of one precision filtered only by xi_cut /
# sponge / suppress_kz0
PACKED_RUNS = [dict(name="p128_k1", nx_override=128, lx_override=12.8, k_mode=1, n_steps=44, noise_seed=1,
                    suppress_kz0=1),
               dict(name="p96_k2", nx_override=96, lx_override=9.6, k_mode=2, n_steps=30, noise_seed=2),
               dict(name="p160_k1", nx_override=160, lx_override=16.0, k_mode=1, n_steps=52, noise_seed=3,
                    suppress_kz0=1, xi_cut=6.0)]


def test_campaign_packed_matches_single_runs(tmp_path):
    base = {k: v for k, v in BASE.items() if k != "kz_keep_max"}
    m = {"defaults": base, "runs": PACKED_RUNS, "sample_stride": 10, "lanes": 2, "segment": 10, "mode": "packed"}
    p = tmp_path / "packed.json"
    p.write_text(json.dumps(m))
    runs, opts = C.load_manifest(str(p))
    out = str(tmp_path / "out")
    rows = C.run_part(runs, opts, out)
    for spec, row in zip(PACKED_RUNS, rows):
        keys = dict(base, **{k: v for k, v in spec.items() if k != "name"})
        print("packed campaign run", spec["name"], flush=True)
        assert row["name"] == spec["name"] and row["steps"] == spec["n_steps"] and row["halted"] == 0
        sim = khs.Simulation(khs.Config(**keys))
        sim.init()
        ts = sim.timeseries(spec["k_mode"], spec["n_steps"], 10)
        sim.close()
        data = np.genfromtxt(os.path.join(out, spec["name"], "timeseries.csv"), delimiter=",", names=True)
        np.testing.assert_array_equal(data["amp_kz"], np.array([float(f"{a:.9e}") for a in ts["amp"]]))


def test_campaign_cli(tmp_path):
    out = str(tmp_path / "cli")
    env = dict(os.environ)
    r = subprocess.run([sys.executable, "-m", "khs.campaign", manifest(tmp_path), "--out", out], env=env,
                       capture_output=True, text=True, timeout=300,
                       cwd=os.path.dirname(os.path.dirname(khs.__file__)))
    assert r.returncode == 0, r.stdout + r.stderr
    rows = C.read_results(os.path.join(out, "results.csv"))
    assert sorted(x["name"] for x in rows) == sorted(x["name"] for x in RUNS)
    assert all(int(x["halted"]) == 0 for x in rows)
