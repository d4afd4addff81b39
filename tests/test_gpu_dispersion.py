"""2-D mode spectrum (khs_mode_spectrum; MFMA in the FP32 form) and the
space-time dispersion of EMHD waves.

* spectrum of the random-phase EMHD initial state: every planted mode
  (kx, kz <= K) has coefficient amp nx nz / 2 e^{i phi} (closed form); FP64
  relative 1e-10, FP32 (MFMA, FP32 accumulation) 2e-5 of the mode scale.
* spectrum of an arbitrary field against numpy's FFT (FP64 1e-12, FP32 2e-5
  of sum|f|).
* EMHD two-fluid periodic wave run: the measured frequency of every
  (kx, kz) mode matches omega^2 = 2 sum q^2 n + k_Yee^2 (khs.dispersion)
  within 2e-3 relative.
"""
import numpy as np
import pytest

import khs
import oracle_ref as orc
from khs.dispersion import emhd_wave_omega, peak_frequency, record_spectrum

pytestmark = pytest.mark.gpu


def spectrum_keys(nx, nz, K, amp, seed, precision):
    return dict(model="emhd", nx_override=nx, nz_override=nz, periodic=1, species=2, emhd_init="spectrum",
                spectrum_kmax=K, spectrum_amp=amp, spectrum_seed=seed, smooth_init=0, precision=precision)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
def test_spectrum_of_planted_modes(precision, tol):
    nx, nz, K, amp, seed = 256, 128, 8, 1e-5, 2
    sim = khs.Simulation(khs.Config(**spectrum_keys(nx, nz, K, amp, seed, precision)))
    sim.init()
    F = sim.mode_spectrum("By", 12)
    want = np.zeros((12, 12), complex)
    for kz in range(1, K + 1):
        for kx in range(1, K + 1):
            ph = 2 * np.pi * (orc.noise(seed, 7, kx * 4096 + kz, 0.5) + 0.5)
            want[kz - 1, kx - 1] = amp * nx * nz / 2 * np.exp(1j * ph)
    scale = amp * nx * nz / 2
    assert np.max(np.abs(F - want)) <= tol * scale


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_spectrum_matches_fft(precision, tol):
    nx, nz, K = 192, 64, 20
    sim = khs.Simulation(khs.Config(nx_override=nx, nz_override=nz, lx_override=19.2, lz_override=6.4,
                                    run_mode=6, precision=precision))
    sim.init()
    r = np.random.default_rng(3)
    sim.set("By2", r.standard_normal((nz, nx)))
    f = sim.get("By2")
    F = sim.mode_spectrum("By2", K)
    ref = np.fft.fft2(f)[1:K + 1, 1:K + 1]
    assert np.max(np.abs(F - ref)) <= tol * np.abs(f).sum()


def test_emhd_wave_dispersion():
    nx = nz = 256
    K = 8
    sim = khs.Simulation(khs.Config(**spectrum_keys(nx, nz, K, 1e-5, 2, "fp32"), n_steps=20000))
    sim.init()
    t, F = record_spectrum(sim, "By", K, 20000, 20)
    kx, kz = np.meshgrid(np.arange(1, K + 1), np.arange(1, K + 1))
    want = emhd_wave_omega(kx, kz, nx, nz, 1.0, 1.0, charges=(-1.0, 1.0), densities=(1.0, 1.0))
    got = np.array([[peak_frequency(t, F[:, i, j]) for j in range(K)] for i in range(K)])
    err = np.abs(got / want - 1.0)
    assert err.max() <= 2e-3, f"max rel err {err.max():.2e}"
