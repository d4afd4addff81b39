"""The keep-|kz|<=K column filter on MFMA (khs_filter_mfma.hip; FP32 form
on f32 MFMA, FP64 form on f64 MFMA) and the per-column kernels it replaces,
against the FP64 oracle's projection (oracle/khs_oracle.cpp keep_modes) on
the state the GPU held just before the filter.

Bar: FP32 form (FP32 accumulation over nz terms) max|diff| <= 1e-5 *
max|value| per array; FP64 form (same sums in another order) 1e-12.  After
the filter no mode above K is left and modes 0..K are those of the input
(FP32 bar).
"""
import numpy as np
import pytest

import khs
import oracle_ref as orc
from helpers import assert_close, make_pair, random_state

pytestmark = pytest.mark.gpu


def keys(nx, nz, K, mfma=1, precision="fp32", **extra):
    k = dict(nx_override=nx, nz_override=nz, lx_override=nx * 0.1, lz_override=nz * 0.1, alpha_YM=1.0, V0=0.05,
             eps_override=1.0, courant_override=0.4, run_mode=6, freeze_override=1, perturb_amp=0.01,
             precision=precision, kz_keep_max=K, mfma_filter=mfma)
    k.update(extra)
    return k


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("mfma", [1, 0])
@pytest.mark.parametrize("shape,K", [((128, 64), 4), ((200, 48), 8), ((256, 256), 8), ((96, 128), 1), ((200, 512), 8)])
def test_keep_filter_vs_oracle(shape, K, mfma, precision):
    nx, nz = shape
    tol = 1e-5 if precision == "fp32" else 1e-12
    cfg, sim, o = make_pair(random_state(nx, nz, seed=nx + K, smooth=True), **keys(nx, nz, K, mfma, precision))
    for s in (0, 1):
        sim.fct_x_sweep(s)
    for s in (0, 1):
        sim.fct_z_sweep(s)
    sim.maxwell()
    pre = {n: sim.get_buf(khs.BUF_NEXT, n) for n in orc.STATE}
    sim.filters()
    o.set_state(pre)
    o.filters()
    for n in orc.STATE:
        assert_close(sim.get_buf(khs.BUF_NEXT, n), o.get(n), tol, f"{precision} K={K} mfma={mfma} {n}")


def test_keep_projection_properties():
    """Independent of the oracle: after the filter every array holds no z
    mode above K, and its modes 0..K equal those of the pre-filter state."""
    nx, nz, K = 128, 128, 8
    cfg, sim, o = make_pair(random_state(nx, nz, seed=5, smooth=True), **keys(nx, nz, K))
    for s in (0, 1):
        sim.fct_x_sweep(s)
    for s in (0, 1):
        sim.fct_z_sweep(s)
    sim.maxwell()
    names = ("Ex2", "Ez3", "By2", "Q2A", "PzB", "nA")
    pre = {n: sim.get_buf(khs.BUF_NEXT, n) for n in names}
    sim.filters()
    for n in names:
        post = sim.get_buf(khs.BUF_NEXT, n)
        sp, s0 = np.fft.rfft(post, axis=0) / nz, np.fft.rfft(pre[n], axis=0) / nz
        scale = np.abs(pre[n]).max()
        assert np.abs(sp[K + 1:]).max() <= 1e-5 * scale, n
        assert np.abs(sp[:K + 1] - s0[:K + 1]).max() <= 1e-5 * scale, n
