"""Pins the oracle's Maxwell, current and source operators (YM_Physics.cu,
YM_Fluid.cu) with closed forms and symmetries."""
import numpy as np
import pytest

import oracle_ref as orc

FIELD_COLOUR = ["Ex", "Ez", "By", "Az"]


def _vacuum(nx, nz, dx, dz, dt):
    """Charge-free fluid (Q = 0, p = 0, n = 1): J = 0 and the fluid is static,
    so a full step is the bare Yee update."""
    p = orc.params(nx, nz, dx, dz, dt, alpha=0.0, periodic_x=1, freeze_az1=0)
    o = orc.Oracle(p)
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    st["nA"] = np.ones((nz, nx))
    st["nB"] = np.ones((nz, nx))
    return o, st


@pytest.mark.parametrize("axis", ["z", "x"])
def test_yee_plane_wave_matches_amplification_matrix(axis):
    # One Fourier mode of the leapfrog Yee scheme evolves by the 2x2 matrix
    # M = [[1, -g1], [-g2, 1 + g1 g2]] per step (E updated first, then B with
    # the new E).  The oracle must reproduce M^N exactly up to rounding.
    n, d, dt, m, steps = 64, 0.1, 0.04, 3, 50
    th = 2 * np.pi * m / n
    j = np.arange(n)
    if axis == "z":
        o, st = _vacuum(4, n, 0.1, d, dt)
        # Ampere: Ex[j] += -dt (By[j] - By[j-1]) / dz ; Faraday: By[j] -= dt (Ex[j+1] - Ex[j]) / dz
        g1 = dt / d * (1 - np.exp(-1j * th))
        g2 = dt / d * (np.exp(1j * th) - 1)
        e_name = "Ex2"
        shape = lambda prof: np.tile(prof[:, None], (1, 4))
        take = lambda a: a[:, 0]
    else:
        o, st = _vacuum(n, 4, d, 0.1, dt)
        # Ampere: Ez[j] += dt (By[j] - By[j-1]) / dx ; Faraday: By[j] -= -dt (Ez[j+1] - Ez[j]) / dx
        g1 = -dt / d * (1 - np.exp(-1j * th))
        g2 = -dt / d * (np.exp(1j * th) - 1)
        e_name = "Ez2"
        shape = lambda prof: np.tile(prof[None, :], (4, 1))
        take = lambda a: a[0, :]
    a0, b0 = 0.0 + 0j, 1.0 + 0j
    st["By2"] = shape(np.real(b0 * np.exp(1j * th * j)))
    o.set_state(st)
    o.sources()
    o.step(steps)
    M = np.array([[1, -g1], [-g2, 1 + g1 * g2]])
    aN, bN = np.linalg.matrix_power(M, steps) @ np.array([a0, b0])
    np.testing.assert_allclose(take(o.get("By2")), np.real(bN * np.exp(1j * th * j)), atol=1e-12)
    np.testing.assert_allclose(take(o.get(e_name)), np.real(aN * np.exp(1j * th * j)), atol=1e-12)
    # the scheme is neutrally stable: |eigenvalues| = 1 at this Courant number
    assert np.allclose(np.abs(np.linalg.eigvals(M)), 1.0, atol=1e-12)


def _rotation(seed):
    r = np.random.default_rng(seed)
    q, _ = np.linalg.qr(r.standard_normal((3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _rotate(st, R):
    out = dict(st)
    for f in FIELD_COLOUR:
        v = np.stack([st[f + c] for c in "123"])
        w = np.einsum("ab,bzx->azx", R, v)
        for k, c in enumerate("123"):
            out[f + c] = w[k]
    for s in "AB":
        v = np.stack([st[f"Q{c}{s}"] for c in "123"])
        w = np.einsum("ab,bzx->azx", R, v)
        for k, c in enumerate("123"):
            out[f"Q{c}{s}"] = w[k]
    return out


def test_colour_rotation_covariance_of_maxwell_and_sources():
    # currents, Ampere, potential, Faraday and the precession source are
    # built from epsilon^abc and dot products: they commute with any SO(3)
    # colour rotation.  A wrong sign in one cross-product component breaks it.
    from helpers import random_state
    nx, nz = 24, 20
    st = random_state(nx, nz, seed=4, smooth=True)
    R = _rotation(5)
    res = []
    for s in (st, _rotate(st, R)):
        o = orc.Oracle(orc.params(nx, nz, 0.1, 0.1, 0.04, alpha=1.7, periodic_x=1, freeze_az1=0))
        o.set_state(s)
        o.update_velocities()
        o.sources()
        o.currents()
        o.ampere()
        o.potential()
        o.faraday()
        out = o.state()
        for c in "123":
            out["J" + c] = o.get("Jx" + c)
            out["sQ" + c] = o.get(f"sQ{c}A")
        out["srcpz"] = o.get("srcpzA")
        out["srcpx"] = o.get("srcpxA")
        res.append(out)
    rot = _rotate(res[0], R)
    for f in FIELD_COLOUR:
        for c in "123":
            np.testing.assert_allclose(res[1][f + c], rot[f + c], atol=1e-13, err_msg=f + c)
    for name in ("J", "sQ"):
        v = np.stack([res[0][name + c] for c in "123"])
        w = np.einsum("ab,bzx->azx", R, v)
        for k, c in enumerate("123"):
            np.testing.assert_allclose(res[1][name + c], w[k], atol=1e-13, err_msg=name + c)
    # the Lorentz force on the fluid is a colour scalar
    np.testing.assert_allclose(res[1]["srcpz"], res[0]["srcpz"], atol=1e-13)
    np.testing.assert_allclose(res[1]["srcpx"], res[0]["srcpx"], atol=1e-13)


def test_uniform_drift_current_and_sources():
    nx, nz = 8, 8
    o = orc.Oracle(orc.params(nx, nz, 0.1, 0.1, 0.04, alpha=2.0, periodic_x=1, freeze_az1=0))
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    one = np.ones((nz, nx))
    st.update(nA=2.0 * one, PxA=0.5 * one, PzA=-1.0 * one, Q1A=1.0 * one, Q2A=0.25 * one,
              nB=1.0 * one, PxB=0.125 * one, PzB=0.5 * one, Q1B=-1.0 * one, Q3B=0.5 * one,
              Ex1=0.5 * one, Ez2=0.25 * one, By3=2.0 * one, Az1=1.0 * one, Az2=-0.5 * one)
    o.set_state(st)
    o.update_velocities()
    o.sources()
    o.currents()
    vA = np.array([0.25, -0.5])      # p / n
    vB = np.array([0.125, 0.5])
    QA, QB = np.array([1.0, 0.25, 0.0]), np.array([-1.0, 0.0, 0.5])
    for c in range(3):
        Jx = -(2.0 * QA[c] * vA[0] + 1.0 * QB[c] * vB[0])
        Jz = -(2.0 * QA[c] * vA[1] + 1.0 * QB[c] * vB[1])
        assert np.all(o.get(f"Jx{c+1}") == Jx) and np.all(o.get(f"Jz{c+1}") == Jz)
    # Lorentz: F_x = -sum Q^a (E_x^a + v_z B^a), F_z = -sum Q^a (E_z^a - v_x B^a)
    E_x, E_z, B = np.array([0.5, 0, 0]), np.array([0, 0.25, 0]), np.array([0, 0, 2.0])
    assert np.allclose(o.get("srcpxA"), -(QA @ (E_x + vA[1] * B)))
    assert np.allclose(o.get("srcpzA"), -(QA @ (E_z - vA[0] * B)))
    # precession: sQ = alpha v_z (Q x A)
    A = np.array([1.0, -0.5, 0.0])
    sq = 2.0 * vA[1] * np.cross(QA, A)
    for c in range(3):
        assert np.allclose(o.get(f"sQ{c+1}A"), sq[c])
    # |Q| is conserved to first order by precession: Q . sQ = 0
    assert abs(QA @ sq) < 1e-15


def test_energy_closed_form():
    nx, nz = 10, 6
    o = orc.Oracle(orc.params(nx, nz, 0.1, 0.1, 0.04))
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    one = np.ones((nz, nx))
    st.update(nA=2.0 * one, PxA=1.0 * one, PzA=2.0 * one, nB=0.5 * one, PzB=1.0 * one,
              Ex2=3.0 * one, By1=1.0 * one, Az2=7.0 * one, Q1A=one)
    o.set_state(st)
    want = nx * nz * (0.5 * 5.0 / 2.0 + 0.5 * 1.0 / 0.5 + 0.5 * (9.0 + 1.0))
    assert o.energy() == want


def test_gauss_closed_form():
    # Ex^1 = sin(2 pi x / Lx) on a periodic grid, rho = 0: G^1 = backward
    # difference of Ex^1 exactly, G^2 = G^3 = 0.
    nx, nz, dx = 32, 4, 0.25
    o = orc.Oracle(orc.params(nx, nz, dx, 0.1, 0.04, alpha=1.0, periodic_x=1))
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    x = np.arange(nx)
    ex = np.sin(2 * np.pi * x / nx)
    st["Ex1"] = np.tile(ex, (nz, 1))
    o.set_state(st)
    g = o.gauss()
    div = (ex - np.roll(ex, 1)) / dx
    assert abs(g[0] - np.sqrt(np.mean(div ** 2))) < 1e-14
    assert g[1] == 0.0 and g[2] == 0.0
    assert abs(g[6] - np.sqrt(np.mean(np.abs(ex) ** 2)) / dx) < 1e-14


def _gauss_state(nx, nz):
    st = {n: np.zeros((nz, nx)) for n in orc.STATE}
    st["nA"] = np.ones((nz, nx))
    st["nB"] = np.ones((nz, nx))
    return st


def test_gauss_non_abelian_term_of_uniform_fields():
    # gauss_check.py:74-84: with uniform Az^b and Ez^c (no divergence) the
    # non-Abelian term is the constant alpha eps^{abc} Az^b Ez^c.  Az = (b1,
    # a2, 0) and Ez = (e1, 0, e3) give nab = alpha (a2 e3, -b1 e3, -a2 e1).
    # Without charge G^a = nab^a (rms = |nab^a|); with rho^a = nab^a planted
    # through Q^a_A (nA = 1) G^a is exactly 0 with the reference's sign and
    # 2 nab^a with rho = -nQ.  A transposed index or the sign of any one term
    # fails one of these.
    nx, nz, al = 8, 6, 1.5
    a2, b1, e1, e3 = 0.7, -0.3, 0.4, 2.0
    nab = np.array([al * (a2 * e3), al * (-(b1 * e3)), al * (-(a2 * e1))])
    o = orc.Oracle(orc.params(nx, nz, 0.1, 0.2, 0.04, alpha=al, periodic_x=1))
    one = np.ones((nz, nx))
    st = _gauss_state(nx, nz)
    st.update(Az1=b1 * one, Az2=a2 * one, Ez1=e1 * one, Ez3=e3 * one)
    o.set_state(st)
    g = o.gauss()
    np.testing.assert_allclose(g[:3], np.abs(nab), rtol=1e-15, atol=0)
    np.testing.assert_allclose(g[3:6], g[:3], rtol=0, atol=0)       # rho = 0: both sign conventions agree
    assert abs(g[6] - (e1 + e3) / 0.1) < 1e-12                      # Etot_scale = rms(sum |E|) / dx
    st.update(Q1A=nab[0] * one, Q2A=nab[1] * one, Q3A=nab[2] * one, Q1B=0 * one)
    o.set_state(st)
    g = o.gauss()
    assert np.all(g[:3] == 0.0)
    np.testing.assert_allclose(g[3:6], 2 * np.abs(nab), rtol=1e-15, atol=0)


def test_gauss_centres_ez_from_the_face_below():
    # Ez^c lives at z + 1/2; gauss_check.py:71-72 averages it onto the cell
    # centre from the faces z - 1/2 and z + 1/2: 0.5 (Ez[z] + Ez[z-1]).  With
    # Az^2 = a at one cell only, G^1 is nonzero in that cell only, so its rms
    # fixes alpha a 0.5 (Ez3[z0] + Ez3[z0-1]) and tells z-1 from z+1.
    nx, nz, al, a = 6, 8, 1.0, 2.0
    rng = np.random.default_rng(3)
    ez3_col = rng.uniform(-1.0, 1.0, nz)
    o = orc.Oracle(orc.params(nx, nz, 0.1, 0.1, 0.04, alpha=al, periodic_x=1))
    st = _gauss_state(nx, nz)
    st["Ez3"] = np.tile(ez3_col[:, None], (1, nx))
    x0, z0 = 2, 5
    st["Az2"][z0, x0] = a
    o.set_state(st)
    g = o.gauss()
    want = abs(al * a * 0.5 * (ez3_col[z0] + ez3_col[z0 - 1])) / np.sqrt(nx * nz)
    wrong = abs(al * a * 0.5 * (ez3_col[z0] + ez3_col[(z0 + 1) % nz])) / np.sqrt(nx * nz)
    assert abs(want - wrong) > 1e-3
    assert abs(g[0] - want) < 1e-15
    # colour 3 sees only the z divergence of Ez3 (a column profile): its
    # backward difference with the periodic wrap of np.roll (gauss_check.py:65-67)
    div = (ez3_col - np.roll(ez3_col, 1)) / 0.1
    assert abs(g[2] - np.sqrt(np.mean(div ** 2))) < 1e-13


def test_gauss_charge_sign_and_both_conventions():
    # rho^a = Q^a_A nA + Q^a_B nB (gauss_check.py:87-89).  nA = 2, nB = 3,
    # Q^1 = (+1, -1) gives rho^1 = -1; Q^2 = (0.5, 0.25) gives rho^2 = +1.75.
    # A uniform non-Abelian term alpha Az^2 Ez^3 = +1 on colour 1 makes the
    # sign visible: G^1 = 1 - rho^1 = 2 with the reference's sign and
    # 1 + rho^1 = 0 with the continuity-consistent one (rho = -nQ).
    nx, nz = 5, 4
    o = orc.Oracle(orc.params(nx, nz, 0.1, 0.1, 0.04, alpha=1.0, periodic_x=1))
    one = np.ones((nz, nx))
    st = _gauss_state(nx, nz)
    st.update(nA=2.0 * one, nB=3.0 * one, Q1A=one, Q1B=-one, Q2A=0.5 * one, Q2B=0.25 * one,
              Az2=0.5 * one, Ez3=2.0 * one)
    o.set_state(st)
    g = o.gauss()
    # colour 1: nab = +1, rho = -1;  colour 2: nab = -Az1 Ez3 = 0, rho = +1.75
    np.testing.assert_allclose(g[:3], [2.0, 1.75, 0.0], rtol=1e-15, atol=0)
    np.testing.assert_allclose(g[3:6], [0.0, 1.75, 0.0], rtol=1e-15, atol=0)


def test_gauss_divergence_is_backward_in_x_with_periodic_wrap():
    # Ex^a at x + 1/2: dEx/dx at cell x is (Ex[x] - Ex[x-1]) / DX, with x - 1
    # wrapped (np.roll, gauss_check.py:65-66).  A single nonzero column Ex^2
    # = v at x0 = nx - 1 gives +v/DX at x0 and, through the wrap, -v/DX at
    # x = 0.  A charge rho^2 = -v/DX at x = 0 cancels the second exactly, so
    # the rms is (v/DX)/sqrt(nx); a forward difference or a missing wrap
    # leaves two or three nonzero columns instead.
    nx, nz, dx, v = 7, 3, 0.25, 0.8
    o = orc.Oracle(orc.params(nx, nz, dx, 0.1, 0.04, alpha=1.0, periodic_x=1))
    st = _gauss_state(nx, nz)
    x0 = nx - 1
    st["Ex2"][:, x0] = v
    st["Q2A"][:, 0] = -v / dx
    o.set_state(st)
    g = o.gauss()
    assert abs(g[1] - (v / dx) / np.sqrt(nx)) < 1e-14


@pytest.mark.parametrize("periodic", [0, 1])
def test_enforce_boundaries_and_floored_velocities(periodic):
    """kernel_enforce_boundaries (YM_Fluid.cu:13-30): densities below 0.05 are
    raised to 0.05 (others untouched); with walls, px of both species and
    Ez^1..3 are zero in the first and last columns, nothing else changes;
    periodic x leaves the walls alone.  kernel_update_velocities (:3-11):
    u = p / max(n, 0.05), the floor applying to the velocity only."""
    from helpers import random_state
    nx, nz = 12, 6
    st = random_state(nx, nz, seed=8)
    st["nA"][2, 3] = 0.01
    st["nA"][4, 0] = -0.3
    st["nB"][1, 11] = 0.0499
    st["nB"][3, 5] = 0.05
    o = orc.Oracle(orc.params(nx, nz, 0.1, 0.1, 0.04, periodic_x=periodic))
    o.set_state(st)
    o.update_velocities()
    for sp in "AB":
        n_safe = np.maximum(st["n" + sp], 0.05)
        np.testing.assert_array_equal(o.get("ux" + sp), st["Px" + sp] / n_safe)
        np.testing.assert_array_equal(o.get("uz" + sp), st["Pz" + sp] / n_safe)
    o.enforce_boundaries()
    walls = [] if periodic else [0, nx - 1]
    for name in orc.STATE:
        want = np.array(st[name], dtype=np.float64)
        if name in ("nA", "nB"):
            want = np.where(want < 0.05, 0.05, want)
        if name in ("PxA", "PxB", "Ez1", "Ez2", "Ez3"):
            want[:, walls] = 0.0
        np.testing.assert_array_equal(o.get(name), want, err_msg=name)
    assert o.get("nA")[2, 3] == 0.05 and o.get("nB")[3, 5] == 0.05 and o.get("nB")[1, 11] == 0.05
