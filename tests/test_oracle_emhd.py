"""Pins the EMHD oracle (oracle/khs_oracle_emhd.cpp, the emhdgpu2d step) with
closed forms: Yee plane waves of both polarisations, the one-sided wall
derivatives, the Lorentz force and current of uniform states, the energy sum,
the FCT sweeps' wiring onto the already-pinned fct_row, and the das_kaw
initial state."""
import numpy as np
import pytest

import oracle_ref as orc

DT = 0.04


def _em(nx, nz, dx=0.1, dz=0.1, dt=DT, periodic=1, nsp=1, q0=-1.0, q1=1.0):
    return orc.EmOracle(orc.EmParams(nx, nz, periodic, nsp, dx, dz, dt, 1.0, 1.0, q0, q1))


def _zeros(o):
    st = {n: np.zeros((o.nz, o.nx)) for n in orc.em_state_names(o.nsp)}
    st["n"] = np.ones((o.nz, o.nx))
    if o.nsp > 1:
        st["n2"] = np.ones((o.nz, o.nx))
    return st


def _random(o, seed):
    r = np.random.default_rng(seed)
    sh = (o.nz, o.nx)
    st = {n: r.standard_normal(sh) * 1e-2 for n in orc.EM_FIELDS}
    for suf in [""] + (["2"] if o.nsp > 1 else []):
        st["n" + suf] = r.uniform(0.02, 1.6, sh)
        for c in ("px", "py", "pz"):
            st[c + suf] = r.standard_normal(sh) * 5e-2
    return st


# ---------------------------------------------------------------------------
# Yee update (EMHD_Physics.cu:21-98)
# ---------------------------------------------------------------------------
# (axis, E, B, sign): one Fourier mode of the leapfrog evolves per step by
# M = [[1, -g1], [-g2, 1 + g1 g2]] with g1 = s dt/d (1 - e^{-i th}) and
# g2 = s dt/d (e^{i th} - 1); s = +1 when E += -dt dB (backward difference)
# and B -= +dt dE (forward difference), s = -1 for the opposite signs.
MODES = [("z", "Ex", "By", +1), ("z", "Ey", "Bx", -1), ("x", "Ez", "By", -1), ("x", "Ey", "Bz", +1)]


@pytest.mark.parametrize("axis,e,b,sgn", MODES)
def test_yee_plane_wave_and_potential(axis, e, b, sgn):
    n, d, m, steps = 64, 0.1, 3, 40
    th = 2 * np.pi * m / n
    j = np.arange(n)
    if axis == "z":
        o = _em(4, n, 0.1, d, q0=0.0)
        shape = lambda prof: np.tile(prof[:, None], (1, 4))
        take = lambda a: a[:, 0]
    else:
        o = _em(n, 4, d, 0.1, q0=0.0)
        shape = lambda prof: np.tile(prof[None, :], (4, 1))
        take = lambda a: a[0, :]
    st = _zeros(o)
    st[b] = shape(np.cos(th * j))
    o.set_state(st)
    o.lorentz()
    g1 = sgn * DT / d * (1 - np.exp(-1j * th))
    g2 = sgn * DT / d * (np.exp(1j * th) - 1)
    M = np.array([[1, -g1], [-g2, 1 + g1 * g2]])
    v = np.array([0j, 1.0 + 0j])
    asum = 0j
    for _ in range(steps):
        v = M @ v
        asum += v[0]
    o.step(steps)
    np.testing.assert_allclose(take(o.get(b)), np.real(v[1] * np.exp(1j * th * j)), atol=1e-12)
    np.testing.assert_allclose(take(o.get(e)), np.real(v[0] * np.exp(1j * th * j)), atol=1e-12)
    # vector potential: A -= dt E after every Ampere update
    np.testing.assert_allclose(take(o.get("A" + e[1])), np.real(-DT * asum * np.exp(1j * th * j)), atol=1e-12)
    # the fluid is charge-free and static: untouched
    assert np.all(o.get("n") == 1.0) and not np.any(o.get("pz"))


def test_walls_one_sided_derivatives():
    # walled x (the reference): By linear in x -> Ez = dt c^2 slope at every
    # x including both walls (forward difference at x = 0); By stays put.
    nx, nz, dx = 12, 4, 0.1
    o = _em(nx, nz, dx=dx, periodic=0, q0=0.0)
    st = _zeros(o)
    slope = 0.3
    st["By"] = np.tile(slope * dx * np.arange(nx), (nz, 1))
    o.set_state(st)
    o.lorentz()
    o.step(1)
    np.testing.assert_allclose(o.get("Ez"), DT * slope, rtol=1e-12)
    np.testing.assert_allclose(o.get("By"), st["By"], atol=1e-15)
    # Ey = 1 inside: the walls are zeroed first (main_coupled.cu:48-63), then
    # Faraday's forward difference (backward at x = nx-1) gives Bz at the edges.
    o = _em(nx, nz, dx=dx, periodic=0, q0=0.0)
    st = _zeros(o)
    st["Ey"] = np.ones((nz, nx))
    o.set_state(st)
    o.lorentz()
    o.step(1)
    ey = np.ones(nx)
    ey[[0, -1]] = 0.0
    np.testing.assert_array_equal(o.get("Ey"), np.tile(ey, (nz, 1)))
    bz = np.zeros(nx)
    bz[0], bz[-2], bz[-1] = -DT / dx, DT / dx, DT / dx
    np.testing.assert_allclose(o.get("Bz"), np.tile(bz, (nz, 1)), atol=1e-15)
    assert not np.any(o.get("Bx"))


def test_enforce_boundaries():
    o = _em(8, 5, periodic=0, nsp=2)
    st = _random(o, 1)
    o.set_state(st)
    o.enforce_boundaries()
    for suf in ("", "2"):
        np.testing.assert_array_equal(o.get("n" + suf), np.maximum(st["n" + suf], 0.05))
        want = st["px" + suf].copy()
        want[:, [0, -1]] = 0.0
        np.testing.assert_array_equal(o.get("px" + suf), want)
        np.testing.assert_array_equal(o.get("pz" + suf), st["pz" + suf])
    for f in ("Ey", "Ez", "Bx"):
        want = st[f].copy()
        want[:, [0, -1]] = 0.0
        np.testing.assert_array_equal(o.get(f), want)
    for f in ("Ex", "By", "Bz"):
        np.testing.assert_array_equal(o.get(f), st[f])


# ---------------------------------------------------------------------------
# coupling (main_coupled.cu:99-136) and energy (:197-222)
# ---------------------------------------------------------------------------
def test_uniform_lorentz_current_energy():
    nx, nz = 6, 5
    o = _em(nx, nz, nsp=2, q0=-1.0, q1=1.0)
    one = np.ones((nz, nx))
    E, B = np.array([0.3, -0.2, 0.5]), np.array([0.7, 1.1, -0.4])
    p1, n1 = np.array([0.2, -0.1, 0.4]), 2.0
    p2, n2 = np.array([-0.3, 0.25, 0.05]), 0.5
    st = {}
    for k, c in enumerate("xyz"):
        st["E" + c] = E[k] * one
        st["B" + c] = B[k] * one
        st["A" + c] = 0 * one
        st["p" + c] = p1[k] * one
        st["p" + c + "2"] = p2[k] * one
    st["n"], st["n2"] = n1 * one, n2 * one
    o.set_state(st)
    o.lorentz()
    o.currents()
    for s, q, p, n in (("", -1.0, p1, n1), ("2", 1.0, p2, n2)):
        f = q * (E + np.cross(p / n, B))
        for k, c in enumerate("xyz"):
            np.testing.assert_allclose(o.get("s" + c + s), f[k], rtol=1e-14)
    J = -p1 + p2
    for k, c in enumerate("xyz"):
        np.testing.assert_allclose(o.get("J" + c), J[k], rtol=1e-14, atol=1e-16)
    want = nx * nz * (0.5 * p1 @ p1 / n1 + 0.5 * p2 @ p2 / n2 + 0.5 * (E @ E + B @ B))
    assert abs(o.energy() - want) <= 1e-13 * want


def test_energy_density_floor():
    o = _em(4, 4)
    st = _zeros(o)
    st["n"] = np.zeros((4, 4))
    st["px"] = np.full((4, 4), 1e-6)
    o.set_state(st)
    # n_safe = max(n, 1e-10) in the energy (not the 0.05 velocity floor)
    assert abs(o.energy() - 16 * 0.5 * 1e-12 / 1e-10) < 1e-12


# ---------------------------------------------------------------------------
# FCT sweeps (FCT_Sweeps.cu:184-232) in terms of the pinned fct_row
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("periodic", [0, 1])
def test_sweeps_are_fct_rows_with_lorentz_sources(periodic):
    nx, nz = 20, 14
    o = _em(nx, nz, periodic=periodic, nsp=2)
    st = _random(o, 7)
    o.set_state(st)
    o.lorentz()
    o.update_velocities()
    ux, uz = o.get("ux"), o.get("uz")
    src = {"n": None, "px": o.get("sx"), "py": o.get("sy"), "pz": o.get("sz")}
    np.testing.assert_array_equal(ux, st["px"] / np.maximum(st["n"], 0.05))
    o.x_sweep(0)
    xs = {}
    for v, s in src.items():
        want = np.empty((nz, nx))
        for z in range(nz):
            if periodic:
                pad = lambda a: np.concatenate([a[-3:], a, a[:3]])
                r = orc.fct_row(pad(ux[z]), pad(st[v][z]), None if s is None else pad(s[z]), o.p.dx, o.p.dt)
                want[z] = r[3:-3]
            else:
                want[z] = orc.fct_row(ux[z], st[v][z], None if s is None else s[z], o.p.dx, o.p.dt)
        np.testing.assert_array_equal(o.get(v), want, err_msg=v)
        xs[v] = want
    # species 2 untouched by the species-1 sweep
    np.testing.assert_array_equal(o.get("n2"), st["n2"])
    o.z_sweep(0)
    for v, s in src.items():
        want = np.empty((nz, nx))
        for x in range(nx):
            # the z sweep uses the velocity set before the x sweep (stale uz)
            want[:, x] = orc.fct_row(uz[:, x], xs[v][:, x], None if s is None else s[:, x], o.p.dz, o.p.dt)
        np.testing.assert_array_equal(o.get(v), want, err_msg=v)


# ---------------------------------------------------------------------------
# init_das_kaw_equilibrium (main_coupled.cu:65-94) and smoothing (:22-35)
# ---------------------------------------------------------------------------
def test_daskaw_init_closed_form():
    nx, nz, k = 64, 48, 2
    o = _em(nx, nz, dx=1.0, dz=1.0, periodic=0, nsp=2)
    o.init_daskaw(k, 1e-3, 0)
    eps, V0, xc = 10.0, 0.1, nx / 2.0
    x = np.arange(nx)[None, :]
    z = np.arange(nz)[:, None]
    xe = (x + 0.5 - xc) / eps
    b0 = -V0 * eps * np.log(np.cosh(xe))
    np.testing.assert_allclose(o.get("pz"), np.broadcast_to(V0 * np.tanh((x - xc) / eps), (nz, nx)), rtol=1e-15)
    np.testing.assert_allclose(o.get("Ex"), np.broadcast_to(V0 * np.tanh(xe) * b0, (nz, nx)), rtol=1e-14)
    by = b0 + 1e-3 * V0 / np.cosh(xe) * np.cos(2 * np.pi * k / nz * (z + 0.5))
    np.testing.assert_allclose(o.get("By"), by, rtol=1e-14, atol=1e-17)
    assert np.all(o.get("n") == 1.0) and np.all(o.get("n2") == 1.0)
    for f in ("Ey", "Ez", "Bx", "Bz", "Ax", "Ay", "Az", "px", "py"):
        assert not np.any(o.get(f)), f


def test_smoothing_eigenvalues():
    # 5-point smoothing (1/2 centre, 1/8 per neighbour; z periodic, x clamped)
    # of f = g(x) cos(th (z + 1/2)): the z neighbours give (cos th / 4) f, so
    # the result is ((1/2 + cos(th)/4) g + X g) cos(...), X g = (g[x+1] + g[x-1]) / 8
    # with clamped ends; the z mean (th = 0) likewise with 3/4.
    nx, nz, k = 64, 48, 3
    o0 = _em(nx, nz, dx=1.0, dz=1.0, periodic=0)
    o0.init_daskaw(k, 1e-3, 0)
    o1 = _em(nx, nz, dx=1.0, dz=1.0, periodic=0)
    o1.init_daskaw(k, 1e-3, 1)
    th = 2 * np.pi * k / nz
    cz = np.cos(th * (np.arange(nz)[:, None] + 0.5))

    def X(g):
        gp = np.concatenate([g[1:], g[-1:]])
        gm = np.concatenate([g[:1], g[:-1]])
        return 0.125 * (gp + gm)

    def split(a):
        mean = a.mean(axis=0)
        return mean, 2.0 / nz * ((a - mean) * cz).sum(axis=0)

    m0, g0 = split(o0.get("By"))
    m1, g1 = split(o1.get("By"))
    np.testing.assert_allclose(m1, 0.75 * m0 + X(m0), rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(g1, (0.5 + 0.25 * np.cos(th)) * g0 + X(g0), rtol=1e-9, atol=1e-18)
    ex0, ex1 = o0.get("Ex")[0], o1.get("Ex")
    np.testing.assert_allclose(ex1, np.tile(0.75 * ex0 + X(ex0), (nz, 1)), rtol=1e-13, atol=1e-18)
    # fields without smoothing untouched
    np.testing.assert_array_equal(o1.get("pz"), o0.get("pz"))
