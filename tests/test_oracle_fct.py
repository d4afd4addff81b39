"""Pins the oracle's FCT sweep (FCT_Sweeps.cu:39-177, FCT_DeviceMath.cuh:38-125)
with properties the algorithm must have, each chosen so that a plausible
mistake (wrong flux sign, dropped antidiffusion, swapped limiter argument,
wrong dt/dx scale, broken clamping) fails it."""
import numpy as np
import pytest

import oracle_ref as orc


def test_linear_profile_translates_exactly():
    # Constant velocity, linear profile: diffusion and antidiffusion cancel and
    # the low-order flux moves the ramp by exactly eps cells in the interior.
    n, dx, dt = 64, 0.1, 0.04
    u0 = 0.75                      # eps = u dt/dx = 0.3
    eps = u0 * dt / dx
    x = np.arange(n, dtype=np.float64)
    v = 2.0 + 0.5 * x
    out = orc.fct_row(np.full(n, u0), v, None, dx, dt)
    np.testing.assert_allclose(out[4:-4], v[4:-4] - eps * 0.5, rtol=0, atol=1e-13)


def test_source_enters_as_dt_times_source():
    n, dx, dt = 32, 0.1, 0.04
    v = np.full(n, 1.25)
    s = np.full(n, 3.0)
    out = orc.fct_row(np.zeros(n), v, s, dx, dt)
    np.testing.assert_allclose(out, 1.25 + dt * 3.0, rtol=0, atol=1e-15)


def _ring_sweep(v, u, dx, dt, steps):
    """Periodic ring through the oracle's periodic x-sweep (one row)."""
    n = len(v)
    p = orc.params(n, 4, dx, 0.1, dt, periodic_x=1)
    o = orc.Oracle(p)
    for name in orc.STATE:
        o.set(name, np.zeros((4, n)))
    o.set("nA", np.tile(v, (4, 1)))
    o.set("nB", np.ones((4, n)))
    o.set("PxA", np.tile(u * v, (4, 1)))        # momentum so that ux = px/n = u
    for _ in range(steps):
        # velocities from the current state: ux = px / max(n, 0.05)
        o.set("PxA", np.tile(u * o.get("nA")[0], (4, 1)))
        o.update_velocities()
        o.sources()
        o.x_sweep(0)
    return o.get("nA")[0]


def test_periodic_sweep_conserves_mass():
    r = np.random.default_rng(1)
    n = 96
    v = r.uniform(0.1, 2.0, n)
    u = np.full(n, 0.6)
    out = _ring_sweep(v, u, 0.1, 0.04, 1)
    assert abs(out.sum() - v.sum()) <= 1e-13 * v.sum()


@pytest.mark.parametrize("sign", [+1, -1])
def test_gaussian_moves_eps_cells_per_step(sign):
    n, dx, dt = 200, 0.1, 0.04
    u0 = sign * 0.625                 # eps = 0.25 cells per step
    x = np.arange(n)
    v = 0.2 + np.exp(-0.5 * ((x - 100) / 8.0) ** 2)
    steps = 80
    out = _ring_sweep(v, np.full(n, u0), dx, dt, steps)

    def centre(a):
        w = a - a.min()
        ang = np.angle(np.sum(w * np.exp(2j * np.pi * x / n)))
        return ang / (2 * np.pi) * n

    shift = (centre(out) - centre(v) + n / 2) % n - n / 2
    assert abs(shift - sign * 0.25 * steps) < 0.3


def test_mirror_symmetry_exact():
    # Reversing the row and the velocity must reverse the result bit for bit
    # (the limiter uses the signed LNRHOT differences of both neighbours).
    r = np.random.default_rng(2)
    n, dx, dt = 77, 0.1, 0.04
    u = r.uniform(-1.0, 1.0, n)
    v = r.uniform(0.0, 2.0, n)
    s = r.standard_normal(n)
    a = orc.fct_row(u, v, s, dx, dt)
    b = orc.fct_row(-u[::-1].copy(), v[::-1].copy(), s[::-1].copy(), dx, dt)[::-1]
    assert np.array_equal(a, b)


def test_top_hat_stays_monotone():
    # No new extrema: the limiter keeps the transported top hat within its bounds.
    n, dx, dt = 120, 0.1, 0.04
    x = np.arange(n)
    v = np.where((x > 40) & (x < 70), 1.0, 0.1)
    out = _ring_sweep(v, np.full(n, 0.9), dx, dt, 30)
    assert out.min() >= 0.1 - 1e-12
    assert out.max() <= 1.0 + 1e-12
    # and the unlimited scheme would have overshot: the limiter is active
    assert out.max() > 0.9


def test_zero_velocity_smooth_profile_unchanged():
    # eps = 0: nu = mu = 1/6; unclipped antidiffusion cancels the diffusion.
    n, dx, dt = 50, 0.1, 0.04
    x = np.arange(n, dtype=np.float64)
    v = 1.0 + 0.01 * x + 0.001 * x ** 2          # convex, monotone: no clipping
    out = orc.fct_row(np.zeros(n), v, None, dx, dt)
    np.testing.assert_allclose(out[3:-3], v[3:-3], rtol=0, atol=1e-14)


def test_clamped_ends_replicate_boundary_value():
    # A uniform row with a uniform velocity is a fixed point (clamped ends
    # replicate the end value, so no flux enters or leaves at the row ends).
    n = 40
    out = orc.fct_row(np.full(n, 0.8), np.full(n, 1.5), None, 0.1, 0.04)
    np.testing.assert_allclose(out, 1.5, rtol=0, atol=1e-15)


def test_z_sweep_is_column_fct():
    # The z-sweep (transpose semantics, YM_FCT.cu:93-135) equals the 1-D
    # clamped sweep applied to each column with the stale z velocity.
    nx, nz, dz, dt = 6, 40, 0.1, 0.04
    r = np.random.default_rng(3)
    p = orc.params(nx, nz, 0.1, dz, dt, periodic_x=0)
    o = orc.Oracle(p)
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    st["nA"] = r.uniform(0.5, 1.5, (nz, nx))
    st["PzA"] = r.uniform(-0.5, 0.5, (nz, nx))
    st["Q2A"] = r.standard_normal((nz, nx))
    st["nB"] = np.ones((nz, nx))
    o.set_state(st)
    o.update_velocities()
    o.sources()                    # all sources vanish for this state except sQ (Az = 0 -> 0)
    uz = o.get("uzA")
    o.z_sweep(0)
    got = o.get("Q2A")
    for x in range(nx):
        want = orc.fct_row(uz[:, x], st["Q2A"][:, x], None, dz, dt)
        assert np.array_equal(got[:, x], want)
