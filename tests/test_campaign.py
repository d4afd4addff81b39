"""Campaign manifests and scheduling (khs.campaign), CPU side: the .ini
templates of ymgpu2d/analysis/gen_multigpu_campaign.py, manifest loading and
the longest-processing-time assignment of runs to GPUs."""
import itertools
import json

import pytest

import khs
from khs import campaign as C


def test_sponge_radius_and_templates():
    # kz_peak ~ 2 alpha (rounded), xi_crit = kz_peak / (alpha V0), x1.3, clamped to 5..55
    assert C.xi_sponge_for(2.0, 0.05, 3) == 52            # 4 / 0.1 = 40 -> 52
    assert C.xi_sponge_for(1.0, 0.05, 1) == 52            # 2 / 0.05 = 40 -> 52
    assert C.xi_sponge_for(0.2, 0.2, 1) == 16             # max(0.5, 0) / 0.04 = 12.5 -> 16
    assert C.xi_sponge_for(0.5, 0.01, 1) == 55            # clamped above
    assert C.xi_sponge_for(5.0, 0.2, 1) == 13             # 10 / 1 -> 13
    k = C.int_kz_keys(2.0, 0.05, 3)
    assert k == dict(k_mode=3, alpha_YM=2.0, V0=0.05, perturb_amp=0.001, run_mode=6, xi_sponge=52.0,
                     sigma_sponge=5.0, suppress_kz0=1, hyp_diff=5e-5, kz_suppress_max=2, eps_override=0.15,
                     kz_suppress_hi=14, target_tu=100)
    h = C.half_kz_keys(2.0, 0.05, 2.5)
    assert (h["k_mode"], h["kz_suppress_max"], h["kz_suppress_hi"], h["nz_override"]) == (5, 4, 28, 128)
    cfg = khs.Config(**h)
    assert cfg.nz == 128 and abs(cfg.lz - C.LZ_4PI) < 1e-6
    # every template is a valid configuration of the library
    khs.Config(**k)


def makespan(costs, parts):
    return max(sum(costs[i] for i in p) for p in parts)


@pytest.mark.parametrize("costs,n", [([5, 4, 3, 3, 3], 2), ([7, 7, 6, 6, 5, 4, 4, 3], 3),
                                     ([1, 1, 1, 1], 3), ([10], 4), ([3, 1, 4, 1, 5, 9, 2, 6], 2)])
def test_lpt_assignment(costs, n):
    parts = C.lpt_assign(costs, n)
    assert sorted(i for p in parts for i in p) == list(range(len(costs)))
    # optimum by brute force; LPT is within 4/3 - 1/(3n) of it (Graham)
    best = min(makespan(costs, [[i for i in range(len(costs)) if a[i] == q] for q in range(n)])
               for a in itertools.product(range(n), repeat=len(costs)))
    assert makespan(costs, parts) <= (4 / 3 - 1 / (3 * n)) * best + 1e-9
    assert C.lpt_assign(costs, n) == parts           # deterministic
    # the largest run goes first to part 0
    assert max(range(len(costs)), key=lambda j: (costs[j], -j)) in parts[0]


def test_load_manifest_json_and_ini(tmp_path):
    m = {"defaults": {"nx_override": 128, "nz_override": 64, "n_steps": 30}, "sample_stride": 10,
         "runs": [{"name": "a", "k_mode": 1}, {"k_mode": 2, "nx_override": 96}]}
    p = tmp_path / "m.json"
    p.write_text(json.dumps(m))
    runs, opts = C.load_manifest(str(p))
    assert [r[0] for r in runs] == ["a", "run001"]
    assert runs[1][1] == {"nx_override": 96, "nz_override": 64, "n_steps": 30, "k_mode": 2}
    assert opts["sample_stride"] == 10 and opts["lanes"] == 4 and opts["mode"] == "concurrent"
    d = tmp_path / "inis"
    d.mkdir()
    (d / "r1.ini").write_text("k_mode = 2\nnz_override = 64\n")
    (d / "r0.ini").write_text("k_mode = 1\n")
    runs, _ = C.load_manifest(str(d))
    assert [r[0] for r in runs] == ["r0", "r1"]
    assert C.make_config(runs[1][1]).k_mode == 2
    assert C.run_cost(C.make_config(runs[1][1])) > 0


def test_load_manifest_modes(tmp_path):
    p = tmp_path / "m.json"
    for mode in ("concurrent", "graph", "packed"):
        p.write_text(json.dumps({"runs": [{"k_mode": 1}], "mode": mode}))
        assert C.load_manifest(str(p))[1]["mode"] == mode
    p.write_text(json.dumps({"runs": [{"k_mode": 1}], "mode": "threads"}))
    with pytest.raises(ValueError, match="mode must be"):
        C.load_manifest(str(p))
