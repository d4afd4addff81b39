"""Space-time dispersion of wave runs.

record_spectrum() samples the 2-D Fourier coefficients of one field
(khs_mode_spectrum: z then x contraction on the device, on MFMA in the FP32
form) every `stride` steps; peak_frequency() takes the temporal DFT of each
mode's history (Hann window, zero padding, parabolic peak interpolation) and
returns its angular frequency.

emhd_wave_omega() is the linear dispersion of transverse electromagnetic
waves in the cold multi-fluid plasma of the EMHD model as the emhdgpu2d step
discretises it: omega^2 = 2 sum_s q_s^2 n_s + k_Yee^2 (c = eps0 = m = 1,
fluids at rest).  The factor 2 on the plasma term is the reference's: both
FCT sweeps add the full dt * Lorentz force to the momenta
(FCT_Sweeps.cu:184-232), so each step applies the impulse twice.  k_Yee is
the Yee lattice wavenumber, (2/dx) sin(kx dx / 2) per axis.
"""
import numpy as np


def record_spectrum(sim, name, kmax, n_steps, stride):
    """Returns (t, F) with F[sample, kz-1, kx-1] complex; sample 0 is t = 0."""
    dt = sim.config.dt
    ts = [sim.step_count * dt]
    fs = [sim.mode_spectrum(name, kmax)]
    done = 0
    while done < n_steps:
        k = min(stride, n_steps - done)
        if sim.step(k):
            raise RuntimeError("run halted while recording the spectrum")
        done += k
        ts.append(sim.step_count * dt)
        fs.append(sim.mode_spectrum(name, kmax))
    return np.array(ts), np.array(fs)


def peak_frequency(t, series, pad=64):
    """Angular frequency |omega| of the strongest line in a uniformly sampled
    (real or complex) series; for a complex series the powers at +omega and
    -omega are added, so a standing wave F0 cos(omega t) is found whatever
    the phase of F0."""
    x = np.asarray(series, dtype=complex)
    x = x - x.mean()
    n = len(x)
    w = np.hanning(n)
    full = np.abs(np.fft.fft(x * w, pad * n))
    m = pad * n
    spec = full[: m // 2 + 1].copy()
    spec[1: m // 2] += full[m - 1: m // 2: -1]
    spec[0] = 0.0
    i = int(np.argmax(spec))
    i = min(max(i, 1), len(spec) - 2)
    y0, y1, y2 = np.log(spec[i - 1:i + 2] + 1e-300)
    den = y0 - 2 * y1 + y2
    d = 0.5 * (y0 - y2) / den if den != 0 else 0.0
    df = 2 * np.pi / (pad * n * (t[1] - t[0]))
    return (i + d) * df


def emhd_wave_omega(kx, kz, nx, nz, dx, dz, charges=(-1.0,), densities=(1.0,)):
    kxe = 2.0 / dx * np.sin(np.pi * np.asarray(kx) / nx)
    kze = 2.0 / dz * np.sin(np.pi * np.asarray(kz) / nz)
    wp2 = sum(q * q * n for q, n in zip(charges, densities))
    return np.sqrt(2.0 * wp2 + kxe ** 2 + kze ** 2)
