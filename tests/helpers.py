"""Shared helpers for the parity tests."""
import numpy as np

import khs
import oracle_ref as orc


def random_state(nx, nz, seed, smooth=False):
    """A state that exercises every branch: densities straddling the 0.05
    floor, both signs of every momentum, colour charges off the +-1 axis."""
    r = np.random.default_rng(seed)
    sh = (nz, nx)

    def f(scale):
        a = r.standard_normal(sh) * scale
        if smooth:   # smooth in x and z: low-pass by averaging neighbours
            for _ in range(4):
                a = 0.25 * (np.roll(a, 1, 0) + np.roll(a, -1, 0) + np.roll(a, 1, 1) + np.roll(a, -1, 1))
        return a

    st = {}
    for c in "123":
        st["Ex" + c] = f(1e-2)
        st["Ez" + c] = f(1e-2)
        st["By" + c] = f(1e-2)
        st["Az" + c] = f(1e-1)
    for s, q in (("A", 1.0), ("B", -1.0)):
        st["n" + s] = 1.0 + f(0.3) if smooth else r.uniform(0.02, 1.6, sh)
        st["Px" + s] = f(5e-2)
        st["Pz" + s] = f(5e-2) + 0.05 * q
        st["Q1" + s] = q + f(0.1)
        st["Q2" + s] = f(0.1)
        st["Q3" + s] = f(0.1)
    return st


def make_sim(**keys):
    cfg = khs.Config(**keys)
    sim = khs.Simulation(cfg)
    sim.init()
    return cfg, sim


def make_pair(state=None, **keys):
    cfg, sim = make_sim(**keys)
    o = orc.Oracle(orc.params_from_config(cfg))
    if state is None:
        state = sim.state()
    else:
        for n, a in state.items():
            sim.set(n, a)
        state = sim.state()      # what the device holds (FP32 rounding in that form)
    o.set_state(state)
    return cfg, sim, o


def assert_bitwise(a, b, name=""):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    bad = a.view(np.uint64) != b.view(np.uint64)
    if bad.any():
        d = np.abs(a - b)
        idx = np.unravel_index(np.argmax(np.where(bad, d, -1)), a.shape)
        raise AssertionError(
            f"{name}: {bad.sum()} of {a.size} values differ; max |diff| {np.nanmax(d):.3e} "
            f"at (z,x)={idx}: got {a[idx]!r} want {b[idx]!r}")


def assert_close(a, b, rtol, name="", atol_scale=None):
    """max|a-b| <= rtol * max(|b|) (+ atol_scale)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b))
    tol = rtol * scale + (atol_scale or 0.0)
    assert err <= tol, f"{name}: max|diff| {err:.3e} > {tol:.3e} (scale {scale:.3e})"
