"""Pins the oracle's column filters (YM_Physics.cu:161-327, 349-584; the
keep-|kz|<=K extension) with planted Fourier modes and closed forms."""
import numpy as np
import pytest

import oracle_ref as orc


def modes(nz, amps):
    z = np.arange(nz)
    v = np.zeros(nz)
    for m, (a, b) in amps.items():
        v += a * np.cos(2 * np.pi * m * z / nz) + b * np.sin(2 * np.pi * m * z / nz)
    return v


def test_band_removes_exactly_its_modes():
    nz = 128
    amps = {0: (0.3, 0), 1: (1.0, -0.5), 2: (0.2, 0.7), 5: (-0.4, 0.1), 9: (0.05, 0.3), 20: (0.6, 0.2)}
    v = modes(nz, amps)
    out = orc.kz_band_column(v, 2, 9)
    keep = {m: ab for m, ab in amps.items() if not 2 <= m <= 9}
    # float twiddles (as the reference): orthogonality holds to ~1e-7
    np.testing.assert_allclose(out, modes(nz, keep), atol=2e-6)


def test_band_nyquist_is_over_subtracted():
    # The reference normalises every mode by 2/NZ, also m = NZ/2, where the
    # projection needs 1/NZ: a band reaching Nyquist subtracts that mode twice.
    nz = 32
    v = np.cos(np.pi * np.arange(nz))
    out = orc.kz_band_column(v, nz // 2, nz // 2)
    np.testing.assert_allclose(out, -v, atol=1e-6)


def _filter_oracle(nz, **kw):
    nx = 3
    o = orc.Oracle(orc.params(nx, nz, 0.1, 0.1, 0.04, **kw))
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    o.set_state(st)
    return o


def test_reference_band_keys_skip_target_mode():
    # kz_suppress_max = 2, kz_suppress_hi = 6: bands [1,2] and [4,6]; k = 3 survives
    # (main_ym.cu:395-401).  Applies to By2 and to the fluid pz alike.
    nz = 64
    o = _filter_oracle(nz, kz_suppress_max=2, kz_suppress_hi=6)
    amps = {m: (1.0 / (m + 1), 0.5) for m in range(0, 9)}
    col = modes(nz, amps)
    o.set("By2", np.tile(col[:, None], (1, 3)))
    o.set("PzA", np.tile(col[:, None], (1, 3)))
    o.filters()
    want = modes(nz, {m: ab for m, ab in amps.items() if m in (0, 3, 7, 8)})
    np.testing.assert_allclose(o.get("By2")[:, 1], want, atol=5e-6)
    np.testing.assert_allclose(o.get("PzA")[:, 2], want, atol=5e-6)


def test_keep_modes_projection_exact_and_idempotent():
    nz, K = 128, 4
    o = _filter_oracle(nz, kz_keep_max=K)
    amps = {m: (np.cos(m), np.sin(2 * m)) for m in range(0, 30)}
    col = modes(nz, amps)
    o.set("Ez3", np.tile(col[:, None], (1, 3)))
    o.set("nB", np.tile(col[:, None], (1, 3)))
    o.filters()
    want = modes(nz, {m: ab for m, ab in amps.items() if m <= K})
    np.testing.assert_allclose(o.get("Ez3")[:, 0], want, atol=1e-13)
    np.testing.assert_allclose(o.get("nB")[:, 0], want, atol=1e-13)
    once = o.get("Ez3")
    o.filters()
    np.testing.assert_allclose(o.get("Ez3"), once, atol=1e-14)
    # colour-1 Az is never filtered (frozen background)
    o.set("Az1", np.tile(col[:, None], (1, 3)))
    o.filters()
    assert np.array_equal(o.get("Az1")[:, 0], col)


def test_zmean_removal_and_colour1_zeroing():
    nz = 64
    o = _filter_oracle(nz, suppress_kz0=1)
    r = np.random.default_rng(0)
    for n in ("Ex2", "Q3B", "By1"):
        o.set(n, 1.0 + r.standard_normal((nz, 3)))
    o.filters()
    assert np.max(np.abs(o.get("Ex2").sum(axis=0))) < 1e-13
    assert np.max(np.abs(o.get("Q3B").sum(axis=0))) < 1e-13
    assert np.all(o.get("By1") == 0.0)


@pytest.mark.parametrize("k", [1, 7, 31])
def test_hyperdiffusion_eigenvalue(k):
    # F -= mu (F[z-2] - 4F[z-1] + 6F - 4F[z+1] + F[z+2]) has the eigenvalue
    # 1 - mu 16 sin^4(pi k / nz) on the mode k (Jacobi evaluation).
    nz, mu = 64, 3e-3
    o = _filter_oracle(nz, hyp_diff=mu)
    z = np.arange(nz)
    col = np.cos(2 * np.pi * k * z / nz + 0.3)
    o.set("Az3", np.tile(col[:, None], (1, 3)))
    o.filters()
    lam = 1 - mu * 16 * np.sin(np.pi * k / nz) ** 4
    np.testing.assert_allclose(o.get("Az3")[:, 1], lam * col, atol=1e-14)


def test_sponge_and_xicut_closed_forms():
    nx, nz, dx, dt, eps = 40, 4, 0.5, 0.05, 1.0
    o = orc.Oracle(orc.params(nx, nz, dx, 0.1, dt, eps=eps, xi_sponge=6.0, sigma_sponge=5.0, xi_cut=8.5))
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    st["By2"] = np.ones((nz, nx))
    st["Q3A"] = np.ones((nz, nx))
    st["By1"] = np.ones((nz, nx))
    o.set_state(st)
    o.filters()
    xi = np.abs((np.arange(nx) * dx - nx * dx / 2) / eps)
    want = np.where(xi > 6.0, np.exp(-5.0 * (xi - 6.0) * dt), 1.0)
    want = np.where(xi > 8.5, 0.0, want)
    np.testing.assert_allclose(o.get("By2")[0], want, rtol=1e-15)
    np.testing.assert_allclose(o.get("Q3A")[2], want, rtol=1e-15)
    assert np.all(o.get("By1") == 1.0)      # colour 1 untouched
