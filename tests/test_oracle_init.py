"""Pins the oracle's initial state (kernel_ym_init, ymgpu2d/YM_Init.cu:32-257)
for every run mode with closed forms and invariants that follow from what
each mode is meant to set up, evaluated in float64 numpy (the kernel's float
math agrees to ~1e-6 relative):

  * the seed's polarisation (linear sin, or circular cos + i sin in colours
    2/3), its x envelope (sech, the |xi| < 3 window, the Gaussian of width
    xi_char = 1/sqrt(alpha kz V0)) and its z wavenumber 2 pi k_mode / Lz;
  * the gauge relations the kernel states for its colour-1 background,
    By1 = -dAz1/dx (modes 3 and 5) and dBy1/dx = Jz1 = -2 V0 tanh(xi) w(xi)
    (mode 6 with init_by1_eq, with and without the edge taper), checked with
    finite differences of the oracle's arrays;
  * the counter-streaming beams: n = 1, Q1 = +-1, pz_B = -pz_A.
A swapped sin/cos, a dropped window, a wrong branch or a sign error in a
background fails one of these.
"""
import numpy as np
import pytest

import oracle_ref as orc

NX, NZ, DX, DZ = 128, 32, 0.1, 0.2
V0, EPS, AMP, ALPHA = 0.05, 1.0, 0.02, 1.5


def _init(run_mode, k_mode=2, eps=EPS, init_by1_eq=0, taper=0.0, seeds=None, nx=NX, dx=DX):
    o = orc.Oracle(orc.params(nx, NZ, dx, DZ, 0.04, V0=V0, eps=eps, alpha=ALPHA))
    o.init(k_mode, AMP, V0, eps, run_mode, ALPHA, init_by1_eq, taper, **(seeds or {}))
    st = o.state()
    o.close()
    x = np.arange(nx) * dx
    z = np.arange(NZ) * DZ
    kz = 2 * np.pi * k_mode / (NZ * DZ)
    return st, x, z, kz


def _beams(st, vz):
    one = np.ones((NZ, st["nA"].shape[1]))
    assert np.array_equal(st["nA"], one) and np.array_equal(st["nB"], one)
    assert np.array_equal(st["Q1A"], one) and np.array_equal(st["Q1B"], -one)
    assert np.array_equal(st["PzB"], -st["PzA"])
    assert np.all(st["PxA"] == 0) and np.all(st["PxB"] == 0)
    # float math: tanhf(xi1) - tanhf(xi2) - 1 cancels to ~1 float ulp of 1
    np.testing.assert_allclose(st["PzA"], np.tile(vz, (NZ, 1)), rtol=2e-6, atol=3e-7 * V0)


def _close(a, b, scale, rtol=3e-6):
    np.testing.assert_allclose(a, b, rtol=0, atol=rtol * scale)


def _circular(st, env, z, kz, names=("By2", "By3")):
    """names[0] + i names[1] == env(x) e^{i kz z}"""
    want2 = np.outer(np.cos(kz * z), env)
    want3 = np.outer(np.sin(kz * z), env)
    s = max(np.max(np.abs(env)), 1e-300)
    _close(st[names[0]], want2, s)
    _close(st[names[1]], want3, s)


def test_mode1_circular_seed_and_log_cosh_background():
    st, x, z, kz = _init(1)
    xi = (x - NX * DX / 2) / EPS
    env = AMP * V0 / np.cosh(xi) * (np.abs(xi) < 3)
    _circular(st, env, z, kz)
    _close(st["Az1"], np.tile(-V0 * np.log(np.cosh(xi)), (NZ, 1)), V0 * np.max(np.abs(xi)))
    _beams(st, V0 * np.tanh(xi))
    for n in ("By1", "Az2", "Az3", "Ex1", "Ez2", "Q2A", "Q3B"):
        assert np.all(st[n] == 0), n


def test_mode2_colour1_seed_without_background():
    st, x, z, kz = _init(2)
    xi = (x - NX * DX / 2) / EPS
    env = AMP * V0 / np.cosh(xi) * (np.abs(xi) < 3)
    _close(st["By1"], np.outer(np.sin(kz * z), env), AMP * V0)
    for n in ("Az1", "By2", "By3", "Az2"):
        assert np.all(st[n] == 0), n
    _beams(st, V0 * np.tanh(xi))


def test_mode3_double_tanh_gauge_relation():
    eps = 0.8
    st, x, z, kz = _init(3, eps=eps)
    L = NX * DX
    xi1, xi2 = (x - L / 4) / eps, (x - 3 * L / 4) / eps
    _beams(st, V0 * (np.tanh(xi1) - np.tanh(xi2) - 1))
    # By1 = -dAz1/dx: central difference of the oracle's own Az1 (O(dx^2))
    az1 = st["Az1"][0]
    d = -(az1[2:] - az1[:-2]) / (2 * DX)
    np.testing.assert_allclose(st["By1"][0, 1:-1], d, rtol=0, atol=2e-3 * V0 / eps)
    _circular(st, AMP * V0 * (1 / np.cosh(xi1) + 1 / np.cosh(xi2)), z, kz)


def test_mode4_step_shear_cosine_background():
    st, x, z, kz = _init(4)
    L = NX * DX
    s = np.where((x < L / 3) | (x >= 2 * L / 3), 1.0, -1.0)
    _beams(st, V0 * s)
    _close(st["Az1"], np.tile(-V0 * np.cos(6 * np.pi * x / L), (NZ, 1)), V0)
    _close(st["By2"], np.tile((AMP * V0 * np.sin(kz * z))[:, None], (1, NX)), AMP * V0)
    assert np.all(st["By3"] == 0) and np.all(st["By1"] == 0)


def test_mode5_thin_tanh_bounded_cosine_gauge_relation():
    eps = 0.5
    st, x, z, kz = _init(5, eps=eps)
    L = NX * DX
    xi5 = (x - L / 2) / eps
    _beams(st, V0 * np.tanh(xi5))
    _close(st["Az1"], np.tile(-V0 * np.cos(2 * np.pi * x / L), (NZ, 1)), V0)
    az1 = st["Az1"][0]
    d = -(az1[2:] - az1[:-2]) / (2 * DX)              # By1 = -dAz1/dx
    np.testing.assert_allclose(st["By1"][0, 1:-1], d, rtol=0, atol=1e-3 * V0 * 2 * np.pi / L)
    _circular(st, AMP * V0 / np.cosh(xi5), z, kz)     # no |xi| < 3 window in this mode


def test_mode6_gaussian_seed_width_from_alpha_kz_v0():
    k_mode = 3
    st, x, z, kz = _init(6, k_mode=k_mode)
    xi = (x - NX * DX / 2) / EPS
    amp = np.hypot(st["Az2"], st["Az3"])[0]
    # Az2 + i Az3 = amp(x) e^{i kz z}
    _circular(st, amp, z, kz, names=("Az2", "Az3"))
    # log amp is a parabola in xi with curvature -alpha kz V0 (xi_char^-2)
    sel = np.abs(xi) < 4
    c = np.polyfit(xi[sel], np.log(amp[sel]), 2)
    np.testing.assert_allclose(c[0], -0.5 * ALPHA * kz * V0, rtol=1e-5)
    np.testing.assert_allclose(np.exp(c[2]), AMP * V0, rtol=1e-5)
    for n in ("By1", "By2", "By3", "Q2A", "Q3B"):
        assert np.all(st[n] == 0), n
    _beams(st, V0 * np.tanh(xi))


def test_mode6_eigen_seed_profiles_scale_each_field():
    rng = np.random.default_rng(5)
    prof = {k: rng.uniform(-1, 1, NX) for k in ("seed_by", "seed_az", "seed_qA", "seed_qB")}
    st, x, z, kz = _init(6, seeds=prof)
    a = AMP * V0
    _circular(st, a * prof["seed_az"], z, kz, names=("Az2", "Az3"))
    _circular(st, a * prof["seed_by"], z, kz, names=("By2", "By3"))
    _circular(st, a * prof["seed_qA"], z, kz, names=("Q2A", "Q3A"))
    _circular(st, a * prof["seed_qB"], z, kz, names=("Q2B", "Q3B"))


@pytest.mark.parametrize("taper", [0.0, 12.0])
def test_mode6_by1_equilibrium_carries_the_beam_current(taper):
    # dBy1/dx = Jz1 = -(nA vzA - nB vzB) Q1... = -2 V0 tanh(xi) w(xi): the
    # colour-1 equilibrium of YM_Init.cu:113-120 and :204-227.  The tapered
    # form is evaluated at xi + dx/2, so the backward difference lands on
    # Jz1(xi_i) to O(dx^2); the untapered one at the cell centre (O(dx)).
    nx, dx, eps = 512, 0.1, 1.0
    st, x, z, kz = _init(6, init_by1_eq=1, taper=taper, nx=nx, dx=dx, eps=eps)
    xi = (x - nx * dx / 2) / eps
    by1 = st["By1"][0]
    dby = (by1[1:] - by1[:-1]) / dx
    vz = st["PzA"][0]
    if taper > 0:
        w = 0.5 * (1 - np.tanh((np.abs(xi) - taper) / 3.0))
        np.testing.assert_allclose(vz, V0 * np.tanh(xi) * w, rtol=0, atol=2e-6 * V0)
        jz = -2 * V0 * np.tanh(xi) * w
        inner = np.abs(xi[1:]) > 1.0                  # the taper form is exact where tanh ~ 1
        np.testing.assert_allclose(dby[inner], jz[1:][inner], rtol=0, atol=2e-3 * V0)
    else:
        jz_half = -2 * V0 * np.tanh(0.5 * (xi[1:] + xi[:-1]))
        np.testing.assert_allclose(dby, jz_half, rtol=0, atol=2e-3 * V0)
        # zero mean up to the box-end correction of <log cosh> ~ Lx/(4 eps) - ln 2
        assert abs(np.mean(by1)) < 5e-3 * V0 * nx * dx


def _write_seed(path, profiles, n_fields=None):
    prof = np.asarray(profiles, dtype=np.float32)
    with open(path, "wb") as f:
        np.array([prof.shape[0] if n_fields is None else n_fields, prof.shape[1]], dtype=np.int32).tofile(f)
        prof.tofile(f)


@pytest.mark.parametrize("nx", [50, 97, 256, 1000])
def test_seed_interpolation_reproduces_linear_profiles(tmp_path, nx):
    # main_ym.cu:199-206 resamples each nx_file-point profile onto NX points
    # by linear interpolation: a linear profile a + b j must come back as
    # a + b t, t = i (nx_file - 1) / (NX - 1), at any NX (float rounding);
    # the end points are exact and NX = nx_file is the identity.
    nfile = 50
    j = np.arange(nfile)
    coef = [(0.5, -0.01), (-1.0, 0.03), (0.0, 0.02), (1.0, -0.02), (0.25, 0.005), (-0.3, 0.011)]
    path = tmp_path / "lin.bin"
    _write_seed(path, [a + b * j for a, b in coef])
    prof = orc.seed_profiles(str(path), nx)
    assert list(prof) == orc.SEED_FIELDS
    t = np.arange(nx) * (nfile - 1) / (nx - 1)
    for (a, b), f in zip(coef, orc.SEED_FIELDS):
        v = prof[f]
        np.testing.assert_allclose(v, a + b * t, rtol=0, atol=4e-7 * (abs(a) + abs(b) * nfile))
        assert v[0] == np.float32(a) and v[-1] == np.float32(a + b * (nfile - 1))
        if nx == nfile:
            assert np.array_equal(v, (a + b * j).astype(np.float32).astype(np.float64))


def test_seed_interpolation_of_a_kink_and_partial_files(tmp_path):
    # a profile that is 0 except f[k] = 1 interpolates to the hat function
    # max(0, 1 - |t - k|); a file with 4 fields sets only by, ex, ez, az;
    # a file that claims more fields than it holds, or a bad header, fails.
    nfile, nx, k = 11, 41, 4
    f = np.zeros(nfile)
    f[k] = 1.0
    path = tmp_path / "kink.bin"
    _write_seed(path, [f, f, f, f])
    prof = orc.seed_profiles(str(path), nx)
    assert list(prof) == ["by", "ex", "ez", "az"]
    t = np.arange(nx) * (nfile - 1) / (nx - 1)
    np.testing.assert_allclose(prof["az"], np.maximum(0, 1 - np.abs(t - k)), rtol=0, atol=1e-6)
    _write_seed(tmp_path / "short.bin", [f, f], n_fields=3)
    assert orc.seed_profiles(str(tmp_path / "short.bin"), nx) is None
    _write_seed(tmp_path / "one.bin", np.zeros((1, 1)))
    assert orc.seed_profiles(str(tmp_path / "one.bin"), nx) is None
    assert orc.seed_profiles(str(tmp_path / "missing.bin"), nx) is None


def test_reference_seed_file_at_its_own_resolution():
    # tests/golden holds a seed file of the reference (ym_eigenmode.py
    # --export-seed, normalised to max|az| = 1): read at its own nx_file the
    # profiles are the stored floats; resampled, max|az| stays 1 to 1e-3
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin")
    hdr = np.fromfile(path, dtype=np.int32, count=2)
    raw = np.fromfile(path, dtype=np.float32, offset=8).reshape(hdr[0], hdr[1])
    prof = orc.seed_profiles(path, int(hdr[1]))
    for k, f in enumerate(orc.SEED_FIELDS[:hdr[0]]):
        assert np.array_equal(prof[f], raw[k].astype(np.float64))
    assert abs(np.max(np.abs(prof["az"])) - 1.0) < 1e-6
    p2 = orc.seed_profiles(path, 500)
    assert abs(np.max(np.abs(p2["az"])) - 1.0) < 1e-3
