"""Growth-rate parity on the GPU.

* khs_mode_amplitude against the numpy restatement of remote_timeseries.py
  (relative 1e-12).
* Physics: the reference's Campaign-25 kz = 1 run (run_mode 6, alpha = 1,
  V0 = 0.05, EPS = 0.15, NX = 768, NZ = 64, xi_sponge = 20, band filter
  kz = 2..14, z-mean removal, hyperdiffusion 5e-5, 6-field eigenfunction
  seed) grows at the linear eigensolver's rate gamma = 0.0897
  (eigenmode_kz1_a1.00_V0.05_EPS0.15.npz); the reference's own simulation
  measured 0.0889.  Bar: max-R^2 fit within [0.95, 1.02] of the eigenvalue.
  Like the reference's runs of this family, the run ends on the energy
  threshold (E/E0 > 100, here near t = 72, from a colour-2/3 disturbance at
  kz ~ 20 outside the band filter); the fit uses samples up to 5 TU before
  the halt.
* Precision form: the FP32 run's growth rate, fitted over the FP64 run's
  window, within relative 1e-3 of the FP64 rate.
"""
import os
import sys

import numpy as np
import pytest

import khs
from helpers import make_sim, random_state

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import growth_fit as gf  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = os.path.join(GOLDEN, "eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin")

C25_KZ1 = dict(run_mode=6, k_mode=1, alpha_YM=1.0, V0=0.05, eps_override=0.15, xi_sponge=20.0,
               sigma_sponge=5.0, nz_override=64, suppress_kz0=1, kz_suppress_max=0, kz_suppress_hi=14,
               hyp_diff=5e-5, perturb_amp=0.001, target_tu=100)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mode_amplitude_matches_numpy(precision):
    nx, nz = 300, 64
    keys = dict(nx_override=nx, nz_override=nz, lx_override=30.0, lz_override=6.4, run_mode=6,
                precision=precision)
    cfg, sim = make_sim(**keys)
    st = random_state(nx, nz, seed=21)
    sim.set("Az2", st["Az2"])
    sim.set("Az3", st["Az3"])
    a2, a3 = sim.get("Az2"), sim.get("Az3")
    for k in (0, 1, 5, 31):
        got = sim.mode_amplitude(k)
        want = gf.circular_amplitude(a2, a3, k)
        np.testing.assert_allclose(got, want, rtol=1e-12, err_msg=f"k={k}")


def _run(precision):
    cfg = khs.Config(**C25_KZ1, precision=precision)
    sim = khs.Simulation(cfg)
    sim.init(SEED)
    stride = int(round(1.0 / cfg.dt))                  # one sample per time unit
    ts = sim.timeseries(1, cfg.total_steps, stride)
    sim.close()
    return cfg, ts


def test_c25_kz1_growth_rate_fp64_and_fp32():
    g = np.load(os.path.join(GOLDEN, "eigenmode_kz1_a1.00_V0.05_EPS0.15_gamma.npz"), allow_pickle=False)
    gamma_exact = float(g["gamma"][0])
    cfg, r64 = _run("fp64")
    assert cfg.nx == 768 and cfg.total_steps == int(100 / cfg.dt)
    t_end = r64["t"][-1] - 5.0
    assert t_end > 50.0
    g64, t0, t1, r2 = khs.fit_growth_max_r2(r64["t"], r64["amp"], t_max=t_end)
    print(f"FP64 gamma {g64:.6f} over [{t0:.1f}, {t1:.1f}] R2 {r2:.6f}; eigensolver {gamma_exact:.6f}",
          flush=True)
    assert 0.95 <= g64 / gamma_exact <= 1.02
    _, r32 = _run("fp32")
    assert r32["t"][-1] >= t1
    g32 = khs.fit_growth_max_r2(r32["t"], r32["amp"], t_min=t0, t_max=t1)[0]
    g64w = khs.fit_growth_max_r2(r64["t"], r64["amp"], t_min=t0, t_max=t1)[0]
    print(f"FP32 gamma {g32:.6f} vs FP64 {g64w:.6f}: rel err {abs(g32 - g64w) / g64w:.2e}", flush=True)
    assert abs(g32 - g64w) <= 1e-3 * g64w


# C1 as configured: 256 x 128 FP64, walls, keep |kz| <= 4, the production
# window (xi_cut = 5, suppress_kz0), a kz = 1 seed of 2e-5 and 1e-9 noise
C1_WINDOW = dict(nx_override=256, nz_override=128, lx_override=25.6, lz_override=12.8, alpha_YM=1, V0=0.05,
                 eps_override=1.0, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1,
                 perturb_amp=2e-5, noise_amp=1e-9, noise_seed=0, kz_keep_max=4, xi_cut=5.0, suppress_kz0=1)


def test_c1_growth_rate_plateau_matches_eigensolver():
    """The kz = 1 amplitude of C1 grows at the eigensolver's top rate once the
    seed's share of the slower eigenmodes (0.064, 0.054) has decayed: the
    plateau of the local rate (batch_analyze.py:90-152) over a 4000-step run
    (t = 160) is within 3% of gamma = 0.0891.  The first 2000 steps alone
    (t = 80) still beat between the modes (their max-R^2 fit reads 0.080);
    the rate does not move under refinement (0.0800 at dx = 0.1, 0.05 and
    0.025), so neither figure is limited by resolution."""
    from khs import eigenmode

    sim = khs.Simulation(khs.Config(**C1_WINDOW, n_steps=4000))
    sim.init()
    ts = sim.timeseries(1, 4000, 10)
    sim.close()
    assert not ts["halted"]
    pl = khs.fit_growth_plateau(ts["t"][1:], ts["amp"][1:])
    vals, _ = eigenmode.solve(2 * np.pi / 12.8, 1.0, 0.05, 1.0, 256, Lx=25.6, n_eigs=12, sigma=0.4, xi_cut=5.0)
    g_eig = float(vals[0].real)
    print("C1 plateau", pl, "eigensolver", g_eig, flush=True)
    assert pl["has_plateau"] and pl["t0"] > 80.0
    assert abs(pl["gamma"] / g_eig - 1.0) <= 0.03
