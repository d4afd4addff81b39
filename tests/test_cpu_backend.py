"""The FP64 CPU oracle behind the library's C ABI (oracle/libkhs_cpu.so):
it exports every entry point of include/khs.h, the ctypes binding drives it
like the GPU library, and its steps are the oracle's (bit for bit against the
oracle driven operator by operator through its own interface), with the
reference's energy cadence and halting rule.  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

import khs
import oracle_ref as orc

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPU_LIB = os.path.join(REPO, "oracle", "libkhs_cpu.so")


@pytest.fixture(scope="module")
def cpu():
    if not os.path.exists(CPU_LIB):
        orc.build()
    lib = khs.load_library(CPU_LIB)
    saved = khs._lib
    khs._lib = lib
    yield lib
    khs._lib = saved


def test_cpu_backend_exports_the_whole_abi(cpu):
    hdr = open(os.path.join(REPO, "include", "khs.h")).read()
    declared = set(re.findall(r"\b(khs_[a-z0-9_]+)\s*\(", hdr))
    raw = ctypes.CDLL(CPU_LIB)
    for name in sorted(declared):
        assert hasattr(raw, name), name
    assert b"CPU" in cpu.khs_version()


def test_cpu_backend_steps_are_the_oracle_steps(cpu):
    keys = dict(nx_override=96, nz_override=64, lx_override=9.6, lz_override=6.4, alpha_YM=1, V0=0.05,
                eps_override=1.0, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1, perturb_amp=0.02,
                noise_amp=1e-6, noise_seed=3, kz_keep_max=4, suppress_kz0=1, xi_cut=3.0)
    cfg = khs.Config(**keys)
    sim = khs.Simulation(cfg)
    sim.init()
    st0 = sim.state()
    o = orc.Oracle(orc.params_from_config(cfg))
    o.set_state(st0)
    o.sources()
    assert sim.step(7) is False
    o.step(7)
    for n in orc.STATE:
        np.testing.assert_array_equal(sim.get(n), o.get(n), err_msg=n)
    assert sim.energy() == o.energy() and sim.step_count == 7
    g = sim.gauss()
    np.testing.assert_array_equal(np.array(g["rms_ref"] + g["rms_cont"] + [g["Etot_scale"]]), o.gauss())
    rms, rel = sim.gauss_residual()
    np.testing.assert_array_equal(rms, o.gauss()[0:3])
    np.testing.assert_array_equal(rel, o.gauss()[0:3] / o.gauss()[6])
    # set_array invalidates the sources: a state written from outside steps
    # like the oracle from that state
    arr = sim.get("By2") * 1.5
    sim.set("By2", arr)
    o.set("By2", arr)
    o.sources()
    sim.step(3)
    o.step(3)
    np.testing.assert_array_equal(sim.get("Q2A"), o.get("Q2A"))
    sim.close()


def test_cpu_backend_halts_and_refuses_what_it_lacks(cpu, tmp_path):
    keys = dict(nx_override=64, nz_override=32, lx_override=6.4, lz_override=3.2, run_mode=0, courant_override=0.4,
                energy_stride=4)
    sim = khs.Simulation(khs.Config(**keys))
    sim.init()
    csv = str(tmp_path / "e.csv")
    sim.energy_csv_open(csv)
    n = sim.get("nA")
    n[2, 2] = np.nan
    sim.set("nA", n)
    assert sim.step(10) is True and sim.step_count == 4
    sim.energy_csv_close()
    rows = open(csv).read().splitlines()
    assert rows[0] == "step,time,E_total,E_ratio" and rows[1].startswith("0,0,") and rows[2].startswith("4,")
    with pytest.raises(khs.KhsError) as e:
        sim.checkpoint(str(tmp_path / "c.khs"))
    assert e.value.code == khs.KHS_ERR_STATE and "CPU backend" in str(e.value)
    sim.close()
    with pytest.raises(khs.KhsError):
        khs.Simulation(khs.Config(model="emhd"))


def test_cpu_backend_mode_amplitude_matches_numpy(cpu):
    import sys
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import growth_fit as gf
    keys = dict(nx_override=40, nz_override=32, lx_override=4.0, lz_override=3.2, run_mode=6)
    sim = khs.Simulation(khs.Config(**keys))
    sim.init()
    rng = np.random.default_rng(4)
    a2, a3 = rng.standard_normal((32, 40)), rng.standard_normal((32, 40))
    sim.set("Az2", a2)
    sim.set("Az3", a3)
    for k in (0, 1, 5):
        np.testing.assert_allclose(sim.mode_amplitude(k), gf.circular_amplitude(a2, a3, k), rtol=1e-12)
    sim.close()
