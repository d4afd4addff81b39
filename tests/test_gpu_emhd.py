"""GPU parity of the EMHD model (emhdgpu2d time step) against the FP64 EMHD
oracle (run on an MI355X).

Bar: the FP64 stencil operators (both FCT sweeps with their Lorentz sources,
boundaries, currents, Ampere, potential, Faraday) and whole steps are
bit-identical to the oracle; the initial state (host double math) is
bit-identical; the energy (compensated GPU sum vs serial) agrees to relative
1e-12; the FP32 form agrees to relative 2e-5 of each array's scale.
"""
import numpy as np
import pytest

import khs
import oracle_ref as orc
from helpers import assert_bitwise, assert_close

pytestmark = pytest.mark.gpu


def em_keys(nx, nz, periodic=0, species=1, **extra):
    keys = dict(model="emhd", nx_override=nx, nz_override=nz, lx_override=nx * 0.5, lz_override=nz * 0.5,
                courant_override=0.1, periodic=periodic, species=species, k_mode=2, perturb_amp=0.01)
    keys.update(extra)
    return keys


def em_random(nx, nz, nsp, seed, smooth=False):
    r = np.random.default_rng(seed)
    sh = (nz, nx)

    def f(scale):
        a = r.standard_normal(sh) * scale
        if smooth:
            for _ in range(4):
                a = 0.25 * (np.roll(a, 1, 0) + np.roll(a, -1, 0) + np.roll(a, 1, 1) + np.roll(a, -1, 1))
        return a

    st = {n: f(1e-2) for n in orc.EM_FIELDS}
    for suf in [""] + (["2"] if nsp > 1 else []):
        st["n" + suf] = 1.0 + f(0.3) if smooth else r.uniform(0.02, 1.6, sh)
        for c in ("px", "py", "pz"):
            st[c + suf] = f(5e-2)
    return st


def em_pair(state, **keys):
    cfg = khs.Config(**keys)
    sim = khs.Simulation(cfg)
    sim.init()
    o = orc.EmOracle(orc.em_params_from_config(cfg))
    if state is not None:
        for n, a in state.items():
            sim.set(n, a)
    o.set_state(sim.state())
    return cfg, sim, o


def fluid_names(nsp):
    f = ["n", "px", "py", "pz"]
    return f + ([n + "2" for n in f] if nsp > 1 else [])


@pytest.mark.parametrize("smooth", [1, 0])
def test_daskaw_init_bitwise(smooth):
    keys = dict(model="emhd", k_mode=3, perturb_amp=0.05, smooth_init=smooth, species=2)
    cfg = khs.Config(**keys)
    assert (cfg.nx, cfg.nz, cfg.dx, cfg.dt) == (256, 256, 1.0, 0.002)
    assert cfg.output_dir() == "coupled_k3"
    sim = khs.Simulation(cfg)
    sim.init()
    o = orc.EmOracle(orc.em_params_from_config(cfg))
    o.init_daskaw(3, 0.05, smooth)
    for n in orc.em_state_names(2):
        assert_bitwise(sim.get(n), o.get(n), f"init {n}")


CASES = {
    "walled_1sp": (96, 64, 0, 1),
    "walled_1sp_odd": (130, 70, 0, 1),
    "periodic_2sp": (128, 48, 1, 2),
    "walled_2sp_odd": (70, 36, 0, 2),
}


@pytest.mark.parametrize("case", list(CASES))
def test_staged_operators_bitwise(case):
    nx, nz, periodic, nsp = CASES[case]
    cfg, sim, o = em_pair(em_random(nx, nz, nsp, seed=nx + nz), **em_keys(nx, nz, periodic, nsp))
    o.lorentz()
    o.update_velocities()
    for s in range(nsp):
        sim.fct_x_sweep(s)
        o.x_sweep(s)
    for n in fluid_names(nsp):
        assert_bitwise(sim.get_buf(khs.BUF_STAGE, n), o.get(n), f"{case} x-sweep {n}")
    for s in range(nsp):
        sim.fct_z_sweep(s)
        o.z_sweep(s)
    o.enforce_boundaries()
    for n in fluid_names(nsp):
        assert_bitwise(sim.get_buf(khs.BUF_NEXT, n), o.get(n), f"{case} z-sweep {n}")
    sim.maxwell()
    o.currents()
    o.ampere()
    o.potential()
    o.faraday()
    for n in orc.EM_FIELDS:
        assert_bitwise(sim.get_buf(khs.BUF_NEXT, n), o.get(n), f"{case} maxwell {n}")
    sim.commit()
    with pytest.raises(khs.KhsError):
        sim.fct_x_sweep(nsp)            # species out of range


@pytest.mark.parametrize("case", list(CASES))
def test_step_bitwise(case):
    nx, nz, periodic, nsp = CASES[case]
    cfg, sim, o = em_pair(em_random(nx, nz, nsp, seed=3, smooth=True), **em_keys(nx, nz, periodic, nsp))
    o.lorentz()
    sim.step(8)
    o.step(8)
    for n in orc.em_state_names(nsp):
        assert_bitwise(sim.get(n), o.get(n), f"{case} {n}")


@pytest.mark.parametrize("zc,pf,zg", [(16, 1, 0), (64, 0, 1), (32, 1, 1), (16, 0, 1), (32, 2, 0)])
def test_tile_variants_bitwise(zc, pf, zg):
    nx, nz, periodic, nsp = 130, 70, 0, 2
    keys = em_keys(nx, nz, periodic, nsp, z_chunk=zc, z_prefetch=pf, z_groups=zg, x_segment=64, x_rows=16,
                   autotune=0)
    cfg, sim, o = em_pair(em_random(nx, nz, nsp, seed=4, smooth=True), **keys)
    t = sim.tuning()
    assert (t["x_segment"], t["x_rows"], t["z_chunk"], t["z_prefetch"]) == (64, 16, zc, pf)
    o.lorentz()
    sim.step(5)
    o.step(5)
    for n in orc.em_state_names(nsp):
        assert_bitwise(sim.get(n), o.get(n), f"z{zc} pf{pf} {n}")


@pytest.mark.parametrize("xzv", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("shape,periodic,nsp", [((130, 70), 0, 2), ((200, 96), 1, 1), ((64, 40), 1, 2),
                                                ((116, 150), 0, 1)])
def test_fused_sweep_variants_bitwise(xzv, shape, periodic, nsp):
    """Fused x+z sweep (register x sweep in the z march, Lorentz force once
    per cell): walls and periodic x, one and two species, one and several
    column tiles and z chunks, give the oracle's bits."""
    nx, nz = shape
    keys = em_keys(nx, nz, periodic, nsp, fused_sweep=1, xz_variant=xzv, autotune=0)
    cfg, sim, o = em_pair(em_random(nx, nz, nsp, seed=nx + xzv, smooth=True), **keys)
    o.lorentz()
    sim.step(5)
    o.step(5)
    t = sim.tuning()
    assert (t["fused_sweep"], t["xz_variant"]) == (1, xzv)
    for n in orc.em_state_names(nsp):
        assert_bitwise(sim.get(n), o.get(n), f"xz{xzv} {shape} {n}")


@pytest.mark.parametrize("mk", [0, 1, 2, 3])
@pytest.mark.parametrize("shape,periodic,nsp", [((130, 70), 0, 2), ((200, 96), 1, 1), ((64, 40), 1, 2),
                                                ((124, 33), 0, 1), ((125, 64), 1, 2)])
def test_maxwell_variants_bitwise(mk, shape, periodic, nsp):
    """EMHD Maxwell as tiles and as a z march (62 columns per wavefront, both
    walls' one-sided differences, periodic x and z wrap, chunk heights that do
    and do not divide nz, one and two fluids) gives the oracle's bits, as an
    operator and in whole steps."""
    nx, nz = shape
    keys = em_keys(nx, nz, periodic, nsp, maxwell_kernel=mk, autotune=0)
    cfg, sim, o = em_pair(em_random(nx, nz, nsp, seed=nx * 3 + mk, smooth=True), **keys)
    o.lorentz()
    o.update_velocities()
    for s in range(nsp):
        sim.fct_x_sweep(s)
        o.x_sweep(s)
    for s in range(nsp):
        sim.fct_z_sweep(s)
        o.z_sweep(s)
    o.enforce_boundaries()
    sim.maxwell()
    o.currents()
    o.ampere()
    o.potential()
    o.faraday()
    for n in orc.EM_FIELDS:
        assert_bitwise(sim.get_buf(khs.BUF_NEXT, n), o.get(n), f"maxwell {mk} {shape} {n}")
    sim.commit()
    o2 = orc.EmOracle(orc.em_params_from_config(cfg))
    o2.set_state(sim.state())
    o2.lorentz()
    sim.step(4)
    o2.step(4)
    for n in orc.em_state_names(nsp):
        assert_bitwise(sim.get(n), o2.get(n), f"maxwell {mk} {shape} steps {n}")


def test_daskaw_steps_bitwise():
    """The reference run itself (256x256, dx = 1, dt = 0.002), 20 steps from init."""
    cfg = khs.Config(model="emhd", k_mode=1, n_steps=20)
    sim = khs.Simulation(cfg)
    sim.init()
    o = orc.EmOracle(orc.em_params_from_config(cfg))
    o.init_daskaw(1, 0.05, 1)
    o.lorentz()
    sim.step(20)
    o.step(20)
    for n in orc.em_state_names(1):
        assert_bitwise(sim.get(n), o.get(n), f"daskaw {n}")
    assert abs(sim.energy() - o.energy()) <= 1e-12 * abs(o.energy())


@pytest.mark.parametrize("periodic,nsp", [(0, 1), (1, 2)])
def test_derived_bitwise(periodic, nsp):
    nx, nz = 150, 40
    cfg, sim, o = em_pair(em_random(nx, nz, nsp, seed=7), **em_keys(nx, nz, periodic, nsp))
    o.lorentz()
    o.update_velocities()
    o.currents()
    names = ["Jx", "Jy", "Jz", "ux", "uz", "sx", "sy", "sz"]
    if nsp > 1:
        names += ["ux2", "uz2", "sx2", "sy2", "sz2"]
    for n in names:
        assert_bitwise(sim.get(n), o.get(n), n)
    Eo = o.energy()
    assert_bitwise(sim.get("energy_density"), o.get("em_energy_density"), "energy density")
    assert abs(sim.energy() - Eo) <= 1e-12 * abs(Eo)


@pytest.mark.parametrize("periodic,nsp", [(0, 2), (1, 1)])
def test_operator_entry_points_bitwise(periodic, nsp):
    """khs_fluid_velocities, khs_sources (sx, sz, sy) and
    khs_enforce_boundaries against the EMHD oracle."""
    nx, nz = 150, 40
    st = em_random(nx, nz, nsp, seed=27 + nsp)
    st["n"][2, :5] = 0.01
    cfg, sim, o = em_pair(st, **em_keys(nx, nz, periodic, nsp))
    o.lorentz()
    o.update_velocities()
    for s in range(nsp):
        tag = "2" if s else ""
        ux, uz = sim.fluid_velocities(s)
        assert_bitwise(ux, o.get("ux" + tag), "ux" + tag)
        assert_bitwise(uz, o.get("uz" + tag), "uz" + tag)
        sx, sz, sy = sim.sources(s)
        assert_bitwise(sx, o.get("sx" + tag), "sx" + tag)
        assert_bitwise(sz, o.get("sz" + tag), "sz" + tag)
        assert_bitwise(sy[0], o.get("sy" + tag), "sy" + tag)
    sim.enforce_boundaries()
    o.enforce_boundaries()
    for n in orc.em_state_names(nsp):
        assert_bitwise(sim.get(n), o.get(n), f"enforce_boundaries {n}")


def test_fp32_form_close():
    nx, nz = 128, 64
    cfg, sim, o = em_pair(em_random(nx, nz, 2, seed=9, smooth=True),
                          **em_keys(nx, nz, 1, 2, precision="fp32"))
    o.lorentz()
    sim.step(4)
    o.step(4)
    for n in orc.em_state_names(2):
        assert_close(sim.get(n), o.get(n), 2e-5, f"fp32 {n}")


def test_spectrum_init():
    """Random-phase spectrum init (extension): By = amp sum_{kx,kz<=K}
    cos(2 pi (kx x/nx + kz z/nz) + phi), phases from the counter generator."""
    nx, nz, K, amp, seed = 64, 32, 3, 1e-4, 5
    cfg = khs.Config(**em_keys(nx, nz, 1, 2, emhd_init="spectrum", spectrum_kmax=K, spectrum_amp=amp,
                               spectrum_seed=seed, smooth_init=0))
    sim = khs.Simulation(cfg)
    sim.init()
    x = np.arange(nx)[None, :]
    z = np.arange(nz)[:, None]
    want = np.zeros((nz, nx))
    for kx in range(1, K + 1):
        for kz in range(1, K + 1):
            ph = 2 * np.pi * (orc.noise(seed, 7, kx * 4096 + kz, 0.5) + 0.5)
            want += np.cos(2 * np.pi * (kx * x / nx + kz * z / nz) + ph)
    assert_close(sim.get("By"), amp * want, 1e-12, "spectrum By")
    assert np.all(sim.get("n") == 1.0) and np.all(sim.get("n2") == 1.0)
    for f in ("Ex", "Ey", "Ez", "Bx", "Bz", "px", "pz2"):
        assert not np.any(sim.get(f)), f


def test_snapshot_and_errors(tmp_path):
    cfg, sim, o = em_pair(None, **em_keys(64, 32, 0, 1))
    sim.write_snapshot(str(tmp_path), 7)
    lines = (tmp_path / "coupled_00007.csv").read_text().splitlines()
    assert lines[0] == "X,Z,By,Pz" and len(lines) == 1 + 64 * 32
    x, z, by, pz = (float(v) for v in lines[1 + 3 + 64].split(","))
    assert (x, z) == (1.5, 0.5)
    assert abs(by - sim.get("By")[1, 3]) <= 1e-5 * abs(by) + 1e-300
    with pytest.raises(khs.KhsError):
        sim.gauss()
    with pytest.raises(KeyError):
        sim.get("Ex1")
    with pytest.raises(khs.KhsError):
        sim.get(85)                    # second fluid absent with species = 1
