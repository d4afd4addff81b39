"""Growth-rate fitters of the library (C++, host) against closed forms and the
numpy restatement of batch_analyze.py in oracle/growth_fit.py."""
import math
import os
import sys

import numpy as np
import pytest

import khs

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import growth_fit as gf  # noqa: E402


def test_oracle_pure_exponential_and_planted_mode():
    t = np.arange(0.0, 40.0, 0.5)
    a = 3e-6 * np.exp(0.0897 * t)
    g, t0, t1, r2 = gf.fit_max_r2(t, a)
    assert abs(g - 0.0897) < 1e-12 and r2 > 1 - 1e-12
    p = gf.fit_plateau(t, a)
    assert p["has_plateau"] and abs(p["gamma"] - 0.0897) < 1e-12
    # a planted mode A cos(2 pi k z / nz + phi(x)) in Az2, A sin(...) in Az3
    # has |F| = nz A at every x; its conjugate vanishes
    nz, nx, k, A = 32, 5, 3, 0.7
    z = np.arange(nz)[:, None]
    ph = np.linspace(0, 2, nx)[None, :]
    amp, conj = gf.circular_amplitude(A * np.cos(2 * np.pi * k * z / nz + ph), A * np.sin(2 * np.pi * k * z / nz + ph), k)
    assert abs(amp - A) < 1e-14 and conj < 1e-14


def test_max_r2_closed_form_and_edges():
    t = np.arange(0.0, 30.0, 1.0)
    g, t0, t1, r2 = khs.fit_growth_max_r2(t, 1e-4 * np.exp(0.123 * t))
    assert abs(g - 0.123) < 1e-12 and abs(r2 - 1.0) < 1e-12
    assert 0.0 <= t0 < t1 <= 29.0            # every window fits exactly: any may win
    g, t0, t1, r2 = khs.fit_growth_max_r2(t[:3], np.ones(3))
    assert g == 0.0 and math.isnan(t0) and math.isnan(t1) and r2 == 0.0
    g, t0, t1, r2 = khs.fit_growth_max_r2(t, np.exp(-0.1 * t))       # nothing grows
    assert g == 0.0 and math.isnan(t0) and r2 == -1e9
    g, t0, t1, r2 = khs.fit_growth_max_r2(t, 1e-4 * np.exp(0.05 * t), t_min=5.0, t_max=20.0)
    assert abs(g - 0.05) < 1e-12 and 5.0 <= t0 < t1 <= 20.0


def _multi_regime(seed):
    # transient decay, exponential growth, saturation, with log-normal noise
    r = np.random.default_rng(seed)
    t = np.arange(0.0, 100.0, 1.0)
    la = np.where(t < 15, -0.05 * t, -0.75 + 0.09 * (t - 15))
    la = np.where(t > 70, -0.75 + 0.09 * 55 + 0.01 * (t - 70), la)
    return t, 1e-6 * np.exp(la + 0.02 * r.standard_normal(t.size))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_max_r2_matches_reference_algorithm(seed):
    t, a = _multi_regime(seed)
    got = khs.fit_growth_max_r2(t, a)
    want = gf.fit_max_r2(t, a)
    assert abs(got[0] - want[0]) <= 1e-9 * abs(want[0])
    assert got[1:3] == tuple(want[1:3])
    assert abs(got[3] - want[3]) < 1e-9
    assert abs(got[0] - 0.09) < 0.01


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_plateau_matches_reference_algorithm(seed):
    t, a = _multi_regime(seed)
    got = khs.fit_growth_plateau(t, a)
    want = gf.fit_plateau(t, a)
    assert got["has_plateau"] == want["has_plateau"]
    if want["has_plateau"]:
        assert abs(got["gamma"] - want["gamma"]) <= 1e-9 * abs(want["gamma"])
        assert (got["t0"], got["t1"]) == (want["t0"], want["t1"])
    # no plateau in pure noise around a constant
    r = np.random.default_rng(seed)
    flat = np.exp(0.3 * r.standard_normal(60))
    assert not khs.fit_growth_plateau(np.arange(60.0), flat)["has_plateau"]
