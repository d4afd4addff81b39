"""Binary checkpoint / restart (khs_checkpoint_write / khs_checkpoint_read):
a run restarted from a checkpoint continues with the bits of the
uninterrupted run (YM FP64 with filters, FP32 form, EMHD two fluids), and
foreign, truncated, corrupted or mismatched files are refused with an error
code, leaving the context's state as it was."""
import os

import numpy as np
import pytest

import khs

pytestmark = pytest.mark.gpu


def ym_keys(**kw):
    keys = dict(nx_override=160, nz_override=64, lx_override=16.0, lz_override=6.4, alpha_YM=1.0, V0=0.05,
                eps_override=1.0, courant_override=0.4, run_mode=0, freeze_override=1, perturb_amp=0.01,
                noise_amp=1e-9, noise_seed=3, n_steps=100, xi_sponge=4.0, kz_keep_max=4)
    keys.update(kw)
    return keys


EM2 = dict(model="emhd", nx_override=96, nz_override=64, lx_override=48.0, lz_override=32.0, courant_override=0.1,
           periodic=1, species=2, emhd_init="spectrum", spectrum_kmax=8)

CASES = {
    "ym_fp64_keep_sponge": ym_keys(),
    "ym_fp64_periodic_bands": ym_keys(run_mode=6, kz_keep_max=0, suppress_kz0=1, hyp_diff=1e-3, xi_sponge=0.0),
    "ym_fp32": ym_keys(precision="fp32"),
    "emhd_two_fluids": EM2,
}


def snapshot(sim):
    return {n: sim.get(n) for n in sim.state_names()}


@pytest.mark.parametrize("case", list(CASES))
def test_restart_is_bit_identical(tmp_path, case):
    keys = CASES[case]
    print(f"checkpoint case {case}", flush=True)
    a = khs.Simulation(khs.Config(**keys))
    a.init()
    assert not a.step(6)
    path = str(tmp_path / "run.khs")
    a.checkpoint(path)
    assert not os.path.exists(path + ".tmp")
    assert not a.step(8)
    ref, e0 = snapshot(a), a.energy0
    assert a.step_count == 14
    a.close()

    b = khs.Simulation(khs.Config(**keys))
    b.init()
    assert not b.step(3)                     # a different state that the restore replaces
    b.restore(path)
    assert b.step_count == 6 and b.energy0 == e0
    assert not b.step(8)
    got = snapshot(b)
    b.close()
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=f"{case} {n}")


def test_bad_files_are_refused(tmp_path):
    keys = ym_keys()
    sim = khs.Simulation(khs.Config(**keys))
    sim.init()
    sim.step(2)
    good = str(tmp_path / "good.khs")
    sim.checkpoint(good)
    before = snapshot(sim)
    blob = open(good, "rb").read()

    def refused(data, code):
        p = str(tmp_path / "bad.khs")
        open(p, "wb").write(data)
        with pytest.raises(khs.KhsError) as e:
            sim.restore(p)
        assert e.value.code == code
        assert sim.step_count == 2
        after = snapshot(sim)
        for n in before:
            np.testing.assert_array_equal(after[n], before[n])

    refused(blob[: len(blob) // 2], khs.KHS_ERR_IO)                         # truncated
    refused(blob[:200] + bytes([blob[200] ^ 1]) + blob[201:], khs.KHS_ERR_IO)  # corrupted array byte
    refused(b"NOTACKPT" + blob[8:], khs.KHS_ERR_IO)                          # foreign file
    with pytest.raises(khs.KhsError) as e:
        sim.restore(str(tmp_path / "missing.khs"))
    assert e.value.code == khs.KHS_ERR_IO
    # another grid / precision
    for other in (ym_keys(nx_override=192, lx_override=19.2), ym_keys(precision="fp32")):
        o = khs.Simulation(khs.Config(**other))
        o.init()
        with pytest.raises(khs.KhsError) as e:
            o.restore(good)
        assert e.value.code == khs.KHS_ERR_CONFIG
        o.close()
    # not between steps of a staged step
    sim.fct_x_sweep(0)
    with pytest.raises(khs.KhsError) as e:
        sim.checkpoint(str(tmp_path / "x.khs"))
    assert e.value.code == khs.KHS_ERR_STATE
    sim.close()
