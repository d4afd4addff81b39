"""Linear eigenmode solver and eigen-seed exporter (khs.eigenmode) against the
reference's own outputs: the eigensolver run and seed file of Campaign 25
(ymgpu2d/eigenmode_kz1_a1.00_V0.05_EPS0.15.npz, eigenmode_seed_kz1_a1.00_
V0.050_sp20.0.bin) and the 8 x 80 growth-rate grid of
presentation/eigensolver_grid_cache.npz, copied into tests/golden by
make_golden.py."""
import os

import numpy as np
import pytest

from khs import eigenmode as E

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ("b", "ex", "ez", "a", "qA", "qB")


def test_wkb_quartic():
    # the reference run's stored gamma_WKB
    g = np.load(os.path.join(GOLD, "eigenmode_kz1_a1.00_V0.05_EPS0.15_gamma.npz"))
    assert E.wkb_growth_rate(1.0, 1.0, 0.05) == pytest.approx(float(g["gamma_wkb"]), rel=1e-12)
    # kz = 0: w^3 = C, the fastest root grows at C^(1/3) sin(2 pi / 3)
    for alpha, V0, n in [(1.0, 0.05, 0), (3.0, 0.03, 1), (2.0, 0.1, 2)]:
        C = (2 * n + 1) * np.sqrt(alpha ** 3 / 2.0) * V0
        assert E.wkb_growth_rate(0.0, alpha, V0, n) == pytest.approx(np.cbrt(C) * np.sqrt(3.0) / 2.0, rel=1e-12)
    assert np.isnan(E.wkb_growth_rate(1.0, 0.0, 0.05))


def test_growth_rate_and_eigenvector_match_reference_run():
    """Campaign 25 point (alpha 1, V0 0.05, EPS 0.15, kz 1, NX 768, sponge at
    xi 20): gamma and the whole 6 x 768 eigenvector."""
    g = np.load(os.path.join(GOLD, "eigenmode_kz1_a1.00_V0.05_EPS0.15_gamma.npz"))
    gamma, vec = E.growth_rate(1.0, 1.0, 0.05, 0.15, 768, xi_sponge=20.0)
    assert gamma.real == pytest.approx(float(g["gamma"][0]), abs=1e-10)
    assert abs(gamma.imag) < 1e-10
    f = np.load(os.path.join(GOLD, "eigenmode_kz1_a1.00_V0.05_EPS0.15_vec.npz"))
    ref = np.concatenate([f[k] for k in FIELDS])
    overlap = abs(np.vdot(vec, ref)) / (np.linalg.norm(vec) * np.linalg.norm(ref))
    assert overlap > 1.0 - 1e-10
    # the sponge changes the selected mode (a plausible wrong default fails)
    g10, _ = E.growth_rate(1.0, 1.0, 0.05, 0.15, 768, xi_sponge=10.0)
    assert abs(g10.real - float(g["gamma"][0])) > 1e-3


def test_seed_export_reproduces_reference_file(tmp_path):
    path, gamma = E.export_seed(str(tmp_path), 1, 1.0, 0.05, 0.15, 768, 20.0)
    name = "eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin"
    assert os.path.basename(path) == name
    ref = E.read_seed(os.path.join(GOLD, name))
    mine = E.read_seed(path)
    assert mine.shape == ref.shape == (6, 768)
    assert np.max(np.abs(mine - ref)) < 1e-6
    raw_ref = open(os.path.join(GOLD, name), "rb").read()
    assert open(path, "rb").read()[:8] == raw_ref[:8]
    # phase alignment: Az peaks at +1
    assert mine[3].max() == pytest.approx(1.0, abs=1e-7)


def test_seed_filename_tags():
    assert E.seed_filename(2, 3.0, 0.03, 0.15, 5.0) == "eigenmode_seed_kz2_a3.00_V0.030_sp5.0.bin"
    assert E.seed_filename(1, 1.0, 0.05, 0.3, 20.0, Lz=4 * np.pi) == \
        "eigenmode_seed_kz1_a1.00_V0.050_EPS0.300_sp20.0_lz12.5664.bin"
    assert E.seed_filename(1, 1.0, 0.05, 0.15, 0.0, Lx=8 * np.pi).endswith("_sp0.0_lx25.1327.bin")


def test_operator_rows():
    """Entries of L at one interior cell, from the linearised equations."""
    kz, alpha, V0, EPS, N = 1.5, 2.0, 0.05, 0.15, 64
    L = E.build_operator(kz, alpha, V0, EPS, N, xi_sponge=3.0, sigma_sponge=5.0).toarray()
    dx = E.LX_DEFAULT / N
    _, xi = E.grid(N, EPS)
    i = 40
    az1, vz = -V0 * np.log(np.cosh(xi[i])), V0 * np.tanh(xi[i])
    om_a, om_f = kz + alpha * az1, kz - alpha * az1
    damp = max(0.0, abs(xi[i]) - 3.0) * 5.0
    b, ex, ez, a, qa, qb = (k * N + i for k in range(6))
    assert L[b, ex] == -1j * om_f and L[b, ez] == -1.0 / dx and L[b, ez + 1] == 1.0 / dx
    assert L[ex, b] == -1j * om_a
    assert L[ez, b] == 1.0 / dx and L[ez, b - 1] == -1.0 / dx and L[ez, qa] == vz and L[ez, qb] == -vz
    assert L[a, ez] == -1.0
    assert L[qa, a] == 1j * alpha * vz and L[qa, qa] == -1j * vz * om_a - damp
    assert L[qb, qb] == 1j * vz * om_a - damp
    assert L[b, b] == -damp and L[ex, ex] == -damp
    # periodic wrap of the forward / backward differences
    assert L[N - 1, 2 * N] == 1.0 / dx and L[2 * N, N - 1] == -1.0 / dx
    # cut rows are empty and their neighbours lose the ghost coupling
    Lc = E.build_operator(kz, alpha, V0, EPS, N, xi_cut=2.0).toarray()
    cut = np.abs(xi) > 2.0
    j = int(np.argmax(~cut)) - 1             # last cut cell before the live region
    assert not np.any(Lc[[k * N + j for k in range(6)], :])
    assert Lc[2 * N + j + 1, j] == 0.0       # ez row of the first live cell: no b[j]


def test_growth_grid_against_reference_cache():
    """Every 8th kz of the 8 campaign series (NX 384): the cached growth rate
    is the real part of one of the 10 eigenvalues nearest the shift that this
    solver returns, and the selected localised mode is the cached one except
    where this solver's nearest-10 set holds a larger localised eigenvalue
    that the reference's Arnoldi run did not return (ARPACK's start vector
    depends on the call history of the process; at kz = 0.125 the missed
    eigenvalue is the one nearest the shift)."""
    gc = np.load(os.path.join(GOLD, "eigensolver_grid_cache.npz"))
    n, same, dropped = 0, 0, []
    for si, (alpha, V0, sp) in enumerate(gc["series"]):
        for j in range(0, 80, 8):
            kz = float(gc["kz_vals"][j])
            ref = gc["gamma"][si, j]
            vals, vecs = E.solve(kz, alpha, V0, 0.15, 384, xi_sponge=sp)
            k = E.growing_mode(vals, vecs, 384, 0.15, xi_sponge=sp, require_localised=True)
            got = np.nan if k is None else vals[k].real
            n += 1
            if np.isnan(ref):
                assert np.isnan(got), (si, kz)
                same += 1
                continue
            assert np.min(np.abs(vals.real - ref)) < 1e-8, (si, kz, ref)
            if abs(got - ref) < 1e-8:
                same += 1
            else:
                # the cached mode is in the same nearest-10 set, below a
                # localised growing mode the reference's run did not return
                assert got > ref and E.is_localised(vecs[:, k], 384, 0.15, xi_inner=sp)
                dropped.append((si, kz))
    assert n == 80 and same >= 72, (same, dropped)
