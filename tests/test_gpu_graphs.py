"""khs_step replays whole energy strides as one captured HIP graph: the
states, step counts, energy rows and halting must be those of the eager
launches, over several replays, from strides that start unaligned, across a
re-capture after the ping-pong parity changes, with odd strides (a graph
from each ping-pong buffer) and with every filter kind (none of which may
leave a memset node in the graph)."""
import numpy as np
import pytest

import khs

pytestmark = pytest.mark.gpu

BASE = dict(nx_override=160, nz_override=64, lx_override=16.0, lz_override=6.4, alpha_YM=1.0, V0=0.05,
            eps_override=1.0, courant_override=0.4, run_mode=0, freeze_override=1, perturb_amp=0.01,
            noise_amp=1e-9, noise_seed=5, n_steps=1000)
CASES = {
    "keep_mfma_sponge": dict(BASE, kz_keep_max=4, xi_sponge=4.0),
    "bands_zmean_hyp": dict(BASE, run_mode=6, suppress_kz0=1, kz_suppress_max=2, kz_suppress_hi=9, hyp_diff=1e-3,
                            xi_cut=6.0),
    "fp32_keep": dict(BASE, precision="fp32", kz_keep_max=4),
    "odd_stride": dict(BASE, kz_keep_max=4, energy_stride=7),     # one graph per starting buffer
    "emhd": dict(model="emhd", nx_override=96, nz_override=64, lx_override=48.0, lz_override=32.0,
                 courant_override=0.1, periodic=1, species=2, emhd_init="spectrum", spectrum_kmax=8,
                 energy_stride=10),
}


def run(keys, graphs, chunks, tmp_path, tag):
    sim = khs.Simulation(khs.Config(**keys, step_graphs=graphs))
    sim.init()
    csv = str(tmp_path / f"{tag}_{graphs}.csv")
    sim.energy_csv_open(csv)
    halted = False
    for n in chunks:
        halted = sim.step(n) or halted
    sim.energy_csv_close()
    out = {n: sim.get(n) for n in sim.state_names()}, sim.step_count, halted, open(csv).read()
    sim.close()
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_step_graphs_match_eager(case, tmp_path):
    keys = CASES[case]
    print("graph case", case, flush=True)
    chunks = [5, 30, 1, 37]                     # unaligned start, several replays, parity changes
    g, e = run(keys, 1, chunks, tmp_path, case), run(keys, 0, chunks, tmp_path, case)
    assert g[1] == e[1] == sum(chunks) and g[2] == e[2] == False
    assert g[3] == e[3]
    for n in g[0]:
        np.testing.assert_array_equal(g[0][n], e[0][n], err_msg=f"{case} {n}")


def test_step_graphs_halt_like_eager():
    keys = dict(BASE, kz_keep_max=4)
    res = []
    for graphs in (1, 0):
        sim = khs.Simulation(khs.Config(**keys, step_graphs=graphs))
        sim.init()
        by = sim.get("By2")
        by[10, 20] = np.nan
        sim.set("By2", by)
        halted = sim.step(60)
        res.append((halted, sim.step_count))
        sim.close()
    stride = khs.Config(**keys).energy_stride
    assert res[0] == res[1] == (True, stride)
