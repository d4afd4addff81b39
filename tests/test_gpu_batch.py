"""Batched runs through captured HIP graphs (khs_batch_*): every graph replay
gives the same bits as stepping each run on its own (eager launches), with
eager work and synchronisations between the replays; a run that halts leaves
the graph and the others continue bit-identically."""
import numpy as np
import pytest

import khs
import oracle_ref as orc
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def small_configs():
    base = dict(alpha_YM=1.0, courant_override=0.4, run_mode=0, freeze_override=1, perturb_amp=1e-3,
                noise_amp=1e-9, xi_cut=4.0, n_steps=1000)
    grids = [(128, 64, 0.05, 1.0, 0, "fp32"), (256, 128, 0.1, 0.7, 1, "fp32"), (96, 32, 0.02, 1.5, 0, "fp64"),
             (64, 64, 0.15, 0.5, 1, "fp32"), (200, 48, 0.08, 1.2, 0, "fp64")]
    out = []
    for i, (nx, nz, v0, a, kz0, prec) in enumerate(grids):
        out.append(dict(base, nx_override=nx, nz_override=nz, lx_override=nx * 0.1, lz_override=nz * 0.1,
                        V0=v0, eps_override=a, k_mode=1 + i % 3, suppress_kz0=kz0, noise_seed=i, precision=prec))
    return out


def test_batch_graph_replays_match_eager():
    cfgs = small_configs()
    batch = khs.Batch(cfgs, lanes=2, segment=10)
    batch.init()
    sims = []
    for c in cfgs:
        s = khs.Simulation(khs.Config(**c))
        s.init()
        sims.append(s)
    other = khs.Simulation(khs.Config(**cfgs[1]))     # unrelated eager work between replays
    other.init()
    for rep in range(3):
        done, halted = batch.run(10)
        assert done == [10 * (rep + 1)] * len(cfgs) and not any(halted)
        other.step(3)
        other.sync()
        for i, s in enumerate(sims):
            s.step(10)
            for n in ("PzA", "Q2B", "By2", "Ez3", "Az2", "nB"):
                assert_bitwise(batch.get(i, n), s.get(n), f"replay {rep} run {i} {n}")
            E, E0 = batch.energy(i)
            assert E == s.energy()
    # odd tail (one eager step per run) and a longer call spanning segments,
    # then replays from the other ping-pong buffer
    done, halted = batch.run(21)
    for i, s in enumerate(sims):
        s.step(21)
        assert done[i] == 51
        assert_bitwise(batch.get(i, "PxA"), s.get("PxA"), f"tail run {i}")
    done, halted = batch.run(20)
    for i, s in enumerate(sims):
        s.step(20)
        assert done[i] == 71
        assert_bitwise(batch.get(i, "Q1B"), s.get("Q1B"), f"after tail run {i}")


def test_batch_halted_run_leaves_graph():
    cfgs = small_configs()[:3]
    batch = khs.Batch(cfgs, lanes=3, segment=4)
    batch.init()
    ref = khs.Simulation(khs.Config(**cfgs[0]))
    ref.init()
    done, halted = batch.run(4)
    ref.step(4)
    arr = batch.get(1, "nA")
    arr[5, 5] = np.nan                               # run 1 goes NaN -> halts at the next check
    batch.set(1, "nA", arr)
    done, halted = batch.run(4)
    ref.step(4)
    assert halted == [False, True, False] and done == [8, 8, 8]
    done, halted = batch.run(8)                      # graph captured again without run 1
    ref.step(8)
    assert halted == [False, True, False] and done == [16, 8, 16]
    for n in ("PzA", "By2", "Az3"):
        assert_bitwise(batch.get(0, n), ref.get(n), f"run 0 {n}")
    with pytest.raises(khs.KhsError):
        batch.get(7, "nA")


def test_batch_concurrent_matches_single_runs():
    """Runs stepped by several host threads on their own streams end in the
    bits of each run stepped alone; per-run step limits stop runs exactly."""
    cfgs = small_configs()
    batch = khs.Batch(cfgs, lanes=2, segment=10)
    batch.init()
    limits = [40, 25, 40, 13, 40]
    for i, n in enumerate(limits):
        batch.set_step_limit(i, n)
    done, halted = batch.run_concurrent(30, threads=3)
    assert done == [min(30, n) for n in limits] and not any(halted)
    done, halted = batch.run_concurrent(30, threads=4)
    assert done == limits and not any(halted)
    for i, c in enumerate(cfgs):
        s = khs.Simulation(khs.Config(**c))
        s.init()
        s.step(limits[i])
        for n in ("PzA", "Q2B", "By2", "Ez3", "Az2", "nB"):
            assert_bitwise(batch.get(i, n), s.get(n), f"concurrent run {i} {n}")
        assert batch.energy(i)[0] == s.energy()
        s.close()
    batch.close()


def test_batch_concurrent_odd_stride_matches_eager():
    """An odd energy stride (the Campaign-25 geometry has 203): the concurrent
    runner's pre-captured graphs from both ping-pong buffers give the bits of
    eager stepping."""
    cfgs = [dict(c, energy_stride=7) for c in small_configs()[:3]]
    batch = khs.Batch(cfgs, lanes=2, segment=10)
    batch.init()
    done, halted = batch.run_concurrent(40, threads=3)
    assert done == [40] * 3 and not any(halted)
    for i, c in enumerate(cfgs):
        s = khs.Simulation(khs.Config(**c, step_graphs=0))
        s.init()
        s.step(40)
        for n in ("PzA", "Q2B", "By2", "Ez3", "Az2", "nB"):
            assert_bitwise(batch.get(i, n), s.get(n), f"odd-stride run {i} {n}")
        s.close()
    batch.close()


def packed_configs():
    """Uneven grids, both precisions excluded (one per batch), walls and
    periodic x, the charge cut, the kz = 0 removal with and without the cut,
    a sponge, warm beams."""
    base = dict(alpha_YM=1.0, courant_override=0.4, freeze_override=1, perturb_amp=1e-3, noise_amp=1e-9,
                n_steps=1000, precision="fp32")
    grids = [(128, 64, 0, dict(xi_cut=4.0)), (256, 128, 0, dict(xi_cut=4.0, suppress_kz0=1)),
             (100, 64, 6, dict(suppress_kz0=1)), (64, 64, 0, dict()),
             (200, 96, 6, dict(xi_sponge=3.0, sigma_sponge=2.0)), (70, 40, 0, dict(xi_cut=2.0, warm_T=0.01))]
    out = []
    for i, (nx, nz, mode, extra) in enumerate(grids):
        out.append(dict(base, nx_override=nx, nz_override=nz, lx_override=nx * 0.1, lz_override=nz * 0.1,
                        V0=0.03 + 0.02 * i, eps_override=0.6 + 0.2 * i, k_mode=1 + i % 3, run_mode=mode,
                        noise_seed=i, **extra))
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_packed_launches_match_single_runs(precision):
    """Runs packed into shared launches give the bits of each run stepped
    alone, over several graph replays with eager work between them, an odd
    tail and a run that halts and leaves the packed graph."""
    cfgs = [dict(c, precision=precision) for c in packed_configs()]
    batch = khs.Batch(cfgs, lanes=2, segment=6)
    batch.init()
    sims = []
    for c in cfgs:
        s = khs.Simulation(khs.Config(**c))
        s.init()
        sims.append(s)
    for rep in range(2):
        done, halted = batch.run_packed(12)
        assert done == [12 * (rep + 1)] * len(cfgs) and not any(halted)
        for i, s in enumerate(sims):
            s.step(12)
            s.sync()
            for n in ("PzA", "Q2B", "By2", "Ez3", "Az2", "nB", "PxA", "Ex1"):
                assert_bitwise(batch.get(i, n), s.get(n), f"packed replay {rep} run {i} {n}")
            assert batch.energy(i)[0] == s.energy()
    done, halted = batch.run_packed(7)                  # a segment of 6 and one eager step
    for i, s in enumerate(sims):
        s.step(7)
        assert done[i] == 31
        assert_bitwise(batch.get(i, "Q3A"), s.get("Q3A"), f"packed tail run {i}")
    arr = batch.get(2, "nA")
    arr[3, 3] = np.nan
    batch.set(2, "nA", arr)
    done, halted = batch.run_packed(12)
    assert halted[2] and not any(h for j, h in enumerate(halted) if j != 2)
    for i, s in enumerate(sims):
        if i == 2:
            continue
        s.step(12)
        assert_bitwise(batch.get(i, "By3"), s.get("By3"), f"packed after halt run {i}")
    for s in sims:
        s.close()
    batch.close()


@pytest.mark.parametrize("shape", list(range(8)))
def test_batch_packed_every_launch_shape(shape):
    """Every packed launch shape (sweep chunks of 8 to 32 rows, one or two
    sweep tiles per workgroup, Maxwell chunks of 2 to 8 rows) gives the bits
    of each run stepped alone."""
    cfgs = [dict(c, precision="fp32" if shape % 2 else "fp64", packed_shape=shape) for c in packed_configs()[:3]]
    batch = khs.Batch(cfgs, lanes=2, segment=4)
    batch.init()
    done, halted = batch.run_packed(8)
    assert done == [8] * len(cfgs) and not any(halted)
    for i, c in enumerate(cfgs):
        s = khs.Simulation(khs.Config(**c))
        s.init()
        s.step(8)
        for n in ("PzA", "Q2B", "By2", "Ez3", "Az2", "nB", "PxA", "Ex1"):
            assert_bitwise(batch.get(i, n), s.get(n), f"shape {shape} run {i} {n}")
        s.close()
    batch.close()


def test_batch_packed_rejects_unsupported_filters():
    cfgs = [dict(c, precision="fp32") for c in packed_configs()[:2]]
    cfgs[1]["kz_keep_max"] = 4
    batch = khs.Batch(cfgs, lanes=1, segment=4)
    batch.init()
    with pytest.raises(khs.KhsError) as e:
        batch.run_packed(4)
    assert e.value.code == khs.KHS_ERR_CONFIG
    batch.close()
