"""GPU parity of the HIP library against the FP64 CPU oracle (run on an MI355X).

Bar (stated per test):
  * FP64 stencil operators (FCT sweeps, sources, currents, Maxwell, sponge,
    xi_cut, hyperdiffusion) and whole steps without DFT filters: bit-identical
    to the oracle (both keep the reference's expression order, no FMA
    contraction).
  * DFT band / z-mean filters: the GPU sums each column in a different order
    than the reference's warp tree -> max|diff| <= 1e-12 * max|value|.
  * energy / Gauss reductions (compensated on the GPU, serial in the oracle):
    relative 1e-12.
  * FP32 precision form: relative 2e-5 of each array's scale after 4 steps.
"""
import numpy as np
import pytest

import khs
import oracle_ref as orc
from helpers import assert_bitwise, assert_close, make_pair, make_sim, random_state

pytestmark = pytest.mark.gpu

FLUID = ["nA", "PxA", "PzA", "Q1A", "Q2A", "Q3A", "nB", "PxB", "PzB", "Q1B", "Q2B", "Q3B"]
FIELDS = ["Ex1", "Ex2", "Ex3", "Ez1", "Ez2", "Ez3", "By1", "By2", "By3", "Az1", "Az2", "Az3"]

C1 = dict(nx_override=256, nz_override=128, lx_override=25.6, lz_override=12.8, alpha_YM=1, V0=0.05,
          eps_override=1.0, courant_override=0.4, k_mode=1, run_mode=0, freeze_override=1,
          perturb_amp=2e-5, noise_amp=1e-9, noise_seed=0, n_steps=2000)


def grid(nx, nz, periodic, **extra):
    keys = dict(nx_override=nx, nz_override=nz, lx_override=nx * 0.1, lz_override=nz * 0.1,
                alpha_YM=1.0, V0=0.05, eps_override=1.0, courant_override=0.4, run_mode=6 if periodic else 0,
                freeze_override=0, perturb_amp=0.01)
    keys.update(extra)
    return keys


SHAPES = [(64, 48), (300, 40), (256, 128)]


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6])
def test_init_bitwise(mode):
    keys = dict(C1)
    keys.update(run_mode=mode, nx_override=96, nz_override=64, init_by1_eq=1 if mode == 6 else 0,
                vz_edge_taper=3.0 if mode == 6 else 0.0)
    cfg, sim = make_sim(**keys)
    o = orc.Oracle(orc.params_from_config(cfg))
    o.init_from_config(cfg)
    for n in orc.STATE:
        want = o.get(n)
        if cfg.noise_amp > 0 and n in ("Ex2", "Ex3", "Ez2", "Ez3", "By2", "By3"):
            cells = np.arange(cfg.nx * cfg.nz, dtype=np.uint64)
            want = want + np.array([orc.noise(0, orc.IDS[n], int(c), 1e-9) for c in cells]).reshape(want.shape)
        assert_bitwise(sim.get(n), want, f"mode {mode} {n}")


@pytest.mark.parametrize("nx", [96, 77, 300])
def test_seeded_init_resampled_bitwise(tmp_path, nx):
    # run_mode 6 with an eigenfunction seed file of 77 points resampled onto
    # NX (main_ym.cu:178-217): the library's reader + init against the
    # oracle's reader + init, bit for bit
    rng = np.random.default_rng(nx)
    prof = rng.uniform(-1, 1, (6, 77)).astype(np.float32)
    path = tmp_path / "seed.bin"
    with open(path, "wb") as f:
        np.array([6, 77], dtype=np.int32).tofile(f)
        prof.tofile(f)
    keys = dict(C1, run_mode=6, nx_override=nx, nz_override=64, lx_override=nx * 0.1, noise_amp=0.0,
                perturb_amp=0.01)
    cfg, sim = make_sim(**keys)
    sim.init(str(path))                # again, from the seed file
    o = orc.Oracle(orc.params_from_config(cfg))
    f32 = np.float32
    o.init(cfg.k_mode, float(f32(cfg.perturb_amp)), float(f32(cfg.V0)), cfg.eps, cfg.run_mode,
           float(f32(cfg.alpha)), cfg.init_by1_eq, float(f32(cfg.vz_edge_taper)),
           **orc.init_seed_args(orc.seed_profiles(str(path), cfg.nx)))
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"seeded nx {nx} {n}")
    assert np.any(sim.get("Q2B") != 0)


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("shape", SHAPES)
def test_staged_operators_bitwise(periodic, shape):
    """x-sweep, z-sweep (+ boundaries), Maxwell: each stage against the oracle."""
    nx, nz = shape
    keys = grid(nx, nz, periodic, warm_T=0.01 if periodic else 0.0)
    st = random_state(nx, nz, seed=nx + nz + periodic)
    cfg, sim, o = make_pair(st, **keys)
    o.update_velocities()
    o.sources()
    # x sweeps
    sim.fct_x_sweep(0)
    sim.fct_x_sweep(1)
    o.x_sweep(0)
    o.x_sweep(1)
    for n in FLUID:
        assert_bitwise(sim.get_buf(khs.BUF_STAGE, n), o.get(n), f"x-sweep {n}")
    # z sweeps (+ density floor and x walls of enforce_boundaries)
    sim.fct_z_sweep(0)
    sim.fct_z_sweep(1)
    o.z_sweep(0)
    o.z_sweep(1)
    o.enforce_boundaries()
    for n in FLUID:
        assert_bitwise(sim.get_buf(khs.BUF_NEXT, n), o.get(n), f"z-sweep {n}")
    # Maxwell
    sim.maxwell()
    o.currents()
    o.ampere()
    o.potential()
    o.faraday()
    for n in FIELDS:
        assert_bitwise(sim.get_buf(khs.BUF_NEXT, n), o.get(n), f"maxwell {n}")
    sim.commit()


@pytest.mark.parametrize("periodic", [0, 1])
def test_derived_bitwise(periodic):
    nx, nz = 300, 40
    cfg, sim, o = make_pair(random_state(nx, nz, seed=7), **grid(nx, nz, periodic, warm_T=0.02))
    o.update_velocities()
    o.sources()
    o.currents()
    for n in ["Jx1", "Jx2", "Jx3", "Jz1", "Jz2", "Jz3", "uxA", "uzA", "uxB", "uzB", "srcpxA", "srcpzA",
              "srcpxB", "srcpzB", "sQ1A", "sQ2A", "sQ3A", "sQ1B", "sQ2B", "sQ3B"]:
        assert_bitwise(sim.get(n), o.get(n), n)
    o.energy()
    assert_bitwise(sim.get("energy_density"), o.get("energy_density"), "energy density")


@pytest.mark.parametrize("periodic", [0, 1])
def test_operator_entry_points_bitwise(periodic):
    """khs_fluid_velocities, khs_sources and khs_enforce_boundaries against
    the oracle's update_velocities, sources and enforce_boundaries."""
    nx, nz = 130, 40
    st = random_state(nx, nz, seed=17 + periodic)
    # densities below the floor and wall momenta / fields for the boundary operator
    st["nA"][3, :7] = 0.01
    st["nB"][5, -4:] = -0.2
    cfg, sim, o = make_pair(st, **grid(nx, nz, periodic, warm_T=0.02))
    o.update_velocities()
    o.sources()
    for s, tag in ((0, "A"), (1, "B")):
        ux, uz = sim.fluid_velocities(s)
        assert_bitwise(ux, o.get("ux" + tag), "ux" + tag)
        assert_bitwise(uz, o.get("uz" + tag), "uz" + tag)
        px, pz, sq = sim.sources(s)
        assert_bitwise(px, o.get("srcpx" + tag), "srcpx" + tag)
        assert_bitwise(pz, o.get("srcpz" + tag), "srcpz" + tag)
        for k in range(3):
            assert_bitwise(sq[k], o.get(f"sQ{k + 1}{tag}"), f"sQ{k + 1}{tag}")
    sim.enforce_boundaries()
    o.enforce_boundaries()
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"enforce_boundaries {n}")
    with pytest.raises(khs.KhsError):
        sim.fluid_velocities(2)


STEP_CASES = {
    "walled_c1_shape": dict(C1, n_steps=0),
    "periodic_mode6_sponge_cut_hyp_warm": grid(128, 64, 1, xi_sponge=4.0, sigma_sponge=5.0, xi_cut=5.5,
                                               hyp_diff=5e-3, warm_T=0.01, freeze_override=1),
    "periodic_unfrozen": grid(96, 32, 1, freeze_override=0, alpha_YM=2.0),
    "walled_unfrozen_odd": grid(70, 36, 0, freeze_override=0),
}


@pytest.mark.parametrize("fused", [1, 0, -1])
@pytest.mark.parametrize("case", list(STEP_CASES))
def test_step_bitwise(case, fused):
    """Whole steps (fused x+z sweep kernel, the split x / z kernels, and the
    autotuner's choice between them)."""
    keys = dict(STEP_CASES[case], fused_sweep=fused)
    nx, nz = keys["nx_override"], keys["nz_override"]
    st = random_state(nx, nz, seed=3, smooth=True)
    cfg, sim, o = make_pair(st, **keys)
    o.sources()
    sim.step(6)
    o.step(6)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"{case} {n}")


FILTER_CASES = {
    "zmean_bands": grid(64, 64, 1, suppress_kz0=1, kz_suppress_max=2, kz_suppress_hi=9, hyp_diff=1e-3),
    "keep4_c1": dict(C1, kz_keep_max=4, nx_override=64),
    "bands_nyquist": grid(48, 32, 0, kz_suppress_max=3, kz_suppress_hi=16),
    "zmean_bands_tall": grid(70, 512, 0, suppress_kz0=1, kz_suppress_max=2, kz_suppress_hi=20),
    "keep_tall_float_twiddles": grid(64, 200, 1, kz_keep_max=6, mfma_filter=0),
    # z-mean only: the one-launch cut + z-mean + colour-1 zeroing kernel
    "zmean_only_ragged": grid(70, 32, 0, suppress_kz0=1),
    "zmean_xicut": grid(96, 64, 0, suppress_kz0=1, xi_cut=3.0),
    "zmean_sponge_periodic": grid(128, 128, 1, suppress_kz0=1, xi_sponge=4.0, sigma_sponge=1.0),
    # the production window with the MFMA keep projection: the z-mean removal
    # is folded into the projection (kz = 0 dropped from the 12 colour-2/3 arrays)
    "keep8_zmean_xicut_mfma": grid(128, 64, 0, kz_keep_max=8, suppress_kz0=1, xi_cut=5.0),
    "keep4_zmean_bands_mfma": grid(96, 128, 1, kz_keep_max=4, suppress_kz0=1, kz_suppress_max=1,
                                   kz_suppress_hi=6, hyp_diff=1e-3),
}


@pytest.mark.parametrize("case", list(FILTER_CASES))
def test_step_filters_close(case):
    """Column DFT filters: coefficient sums in another order -> 1e-12 relative."""
    keys = FILTER_CASES[case]
    nx, nz = keys["nx_override"], keys["nz_override"]
    cfg, sim, o = make_pair(random_state(nx, nz, seed=5, smooth=True), **keys)
    o.sources()
    sim.step(3)
    o.step(3)
    for n in orc.STATE:
        assert_close(sim.get(n), o.get(n), 1e-12, f"{case} {n}")


def test_energy_gauss_close():
    nx, nz = 300, 40
    cfg, sim, o = make_pair(random_state(nx, nz, seed=11), **grid(nx, nz, 1))
    E, Eo = sim.energy(), o.energy()
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    g, go = sim.gauss(), o.gauss()
    got = np.array(g["rms_ref"] + g["rms_cont"] + [g["Etot_scale"]])
    np.testing.assert_allclose(got, go, rtol=1e-12)
    rms, rel = sim.gauss_residual()             # gauss_check.py's rms and relative residual
    np.testing.assert_allclose(rms, go[0:3], rtol=1e-12)
    np.testing.assert_allclose(rel, np.asarray(go[0:3]) / go[6], rtol=1e-12)


def test_fp32_form_close():
    """FP32 bulk arithmetic vs the FP64 oracle on a smooth state, 4 steps."""
    nx, nz = 128, 64
    keys = grid(nx, nz, 1, precision="fp32", xi_sponge=4.0, hyp_diff=1e-3)
    cfg, sim, o = make_pair(random_state(nx, nz, seed=9, smooth=True), **keys)
    o.sources()
    sim.step(4)
    o.step(4)
    for n in orc.STATE:
        assert_close(sim.get(n), o.get(n), 2e-5, f"fp32 {n}")


def test_precision_form_keep8_step_close():
    """The precision form's whole step (FP32 sweeps, Maxwell and sources, the
    keep |kz| <= 8 projection on FP32 MFMA with the z-mean removal folded in,
    the xi_cut window) against the FP64 oracle over 4 steps: 5e-5 of each
    array's scale (FP32 rounding of the state, 2e-5 without the filter, plus
    the FP32 contraction of 64 rows)."""
    nx, nz = 256, 64
    keys = grid(nx, nz, 0, precision="fp32", kz_keep_max=8, suppress_kz0=1, xi_cut=5.0)
    cfg, sim, o = make_pair(random_state(nx, nz, seed=13, smooth=True), **keys)
    assert sim.tuning()["fused_sweep"] in (0, 1)
    o.sources()
    sim.step(4)
    o.step(4)
    for n in orc.STATE:
        assert_close(sim.get(n), o.get(n), 5e-5, f"fp32 keep8 {n}")


def test_precision_form_constraint_and_energy_track_fp64():
    """The precision form's errors as the FP32 form reports them: from the same
    seeded KH state (V0 = 0.1, a = 0.5, keep |kz| <= 8, the production window)
    FP32 and FP64 runs of 400 steps keep E/E0 within 1e-6 of each other and
    the Gauss-law residual rms of every colour within 1e-3 relative, or
    within FP32's resolution of the divergence (2e-6 of the field scale, the
    level of a residual that is round-off in both runs).  The FP64 run is the
    oracle-exact path."""
    keys = dict(C1, V0=0.1, eps_override=0.5, kz_keep_max=8, suppress_kz0=1, xi_cut=5.0)
    res = {}
    for prec in ("fp64", "fp32"):
        sim = khs.Simulation(khs.Config(**keys, precision=prec))
        sim.init()
        e0 = sim.energy()
        assert not sim.step(400)
        g = sim.gauss()
        res[prec] = (sim.energy() / e0, np.array(g["rms_ref"]), g["Etot_scale"])
        sim.close()
    (r64, g64, scale), (r32, g32, _) = res["fp64"], res["fp32"]
    print("E/E0", r64, r32, "gauss", g64, g32, "scale", scale, flush=True)
    assert abs(r32 - r64) <= 1e-6, (r32, r64)
    assert np.all(g64 > 0)
    np.testing.assert_allclose(g32, g64, rtol=1e-3, atol=2e-6 * scale)


def test_full_size_step_bitwise():
    """One step at the 2048x1024 headline size, bit-identical to the oracle."""
    keys = dict(C1, nx_override=2048, nz_override=1024, lx_override=204.8, lz_override=102.4,
                noise_seed=1, n_steps=5000)
    cfg, sim, o = make_pair(**keys)
    o.sources()
    sim.step(2)
    o.step(2)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"2048x1024 {n}")


@pytest.mark.parametrize("xs,xr,zc,pf,mk,xv,zg", [(64, 1, 16, 0, 0, 1, 0), (128, 4, 64, 1, 1, 0, 1), (256, 16, 32, 0, 2, 1, 1),
                                                  (64, 16, 32, 1, 3, 0, 0), (128, 1, 16, 1, 1, 1, 1), (256, 4, 64, 0, 2, 0, 1)])
@pytest.mark.parametrize("periodic", [1, 0])
def test_tile_variants_bitwise(xs, xr, zc, pf, mk, xv, zg, periodic):
    """Every tile / kernel variant the autotuner may pick gives the same bits."""
    keys = grid(300, 136, periodic, warm_T=0.01, x_segment=xs, x_rows=xr, z_chunk=zc, z_prefetch=pf,
                maxwell_kernel=mk, x_serial=xv, z_groups=zg, autotune=0, xi_sponge=8.0, xi_cut=12.0)
    st = random_state(300, 136, seed=31, smooth=True)
    cfg, sim, o = make_pair(st, **keys)
    t = sim.tuning()
    assert t == dict(x_segment=xs, x_rows=xr, z_chunk=zc, z_prefetch=pf, maxwell=mk, x_serial=xv,
                     xz_variant=t["xz_variant"], z_groups=zg, x_lane_cells=0, zm_fused=0, fused_sweep=0,
                     step_fusion=0, mxz_variant=t["mxz_variant"])
    o.sources()
    sim.step(3)
    o.step(3)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"x{xs}r{xr} z{zc} pf{pf} {n}")


@pytest.mark.parametrize("cells", [4, 8])
@pytest.mark.parametrize("shape,periodic,warm", [((300, 136), 1, 0.01), ((256, 64), 0, 0.0), ((70, 36), 0, 0.01),
                                                 ((1000, 40), 0, 0.01), ((1000, 40), 1, 0.0), ((500, 24), 1, 0.01)])
def test_x_register_form_bitwise(cells, shape, periodic, warm):
    """Register-form x sweep (lane runs of 4 / 8 cells, DPP neighbours; vector
    and per-cell loads, one and several wave segments per row, both ends)
    gives the oracle's bits."""
    nx, nz = shape
    keys = grid(nx, nz, periodic, warm_T=warm, x_lane_cells=cells, autotune=0, xi_sponge=8.0)
    cfg, sim, o = make_pair(random_state(nx, nz, seed=nx + cells, smooth=True), **keys)
    t = sim.tuning()
    assert (t["x_lane_cells"], t["x_segment"]) == (cells, 62 * cells)
    o.sources()
    sim.step(3)
    o.step(3)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"x lane cells {cells} {shape} {n}")


@pytest.mark.parametrize("zc", [16, 32, 64])
@pytest.mark.parametrize("shape,periodic,warm", [((300, 136), 1, 0.01), ((70, 36), 0, 0.0), ((128, 200), 0, 0.01)])
def test_z_two_row_prefetch_bitwise(zc, shape, periodic, warm):
    """z march with loads two rows ahead (chunk ends, clamped z ends) gives
    the oracle's bits."""
    nx, nz = shape
    keys = grid(nx, nz, periodic, z_chunk=zc, z_prefetch=2, autotune=0, warm_T=warm, xi_sponge=8.0)
    cfg, sim, o = make_pair(random_state(nx, nz, seed=nz + zc, smooth=True), **keys)
    t = sim.tuning()
    assert (t["z_chunk"], t["z_prefetch"], t["z_groups"]) == (zc, 2, 0)
    o.sources()
    sim.step(3)
    o.step(3)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"z{zc} pf2 {shape} {n}")


@pytest.mark.parametrize("zm", [1, 2, 3])
@pytest.mark.parametrize("shape,periodic,freeze,warm", [((300, 136), 1, 1, 0.01), ((70, 36), 0, 1, 0.0),
                                                        ((256, 64), 0, 0, 0.01), ((130, 100), 1, 0, 0.0),
                                                        ((64, 16), 0, 1, 0.0)])
def test_fused_z_maxwell_bitwise(zm, shape, periodic, freeze, warm):
    """Fused z sweep + Maxwell kernel (three wavefronts per workgroup, LDS
    hand-over, row 0 recomputed for the last chunk) gives the oracle's bits:
    chunk heights that do and do not divide nz, a single chunk, walls and
    periodic x, warm sources, sponge and cut, Az1 frozen or not."""
    nx, nz = shape
    keys = grid(nx, nz, periodic, zm_fused=zm, autotune=0, xi_sponge=8.0, xi_cut=12.0, freeze_override=freeze,
                warm_T=warm)
    cfg, sim, o = make_pair(random_state(nx, nz, seed=nx * 3 + zm, smooth=True), **keys)
    assert sim.tuning()["zm_fused"] == zm
    o.sources()
    sim.step(3)
    o.step(3)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"zm {zm} {shape} {n}")


@pytest.mark.parametrize("mk", [4, 5, 6, 7, 8, 9, 10])
@pytest.mark.parametrize("shape,periodic,freeze", [((300, 136), 1, 1), ((70, 36), 0, 1), ((256, 64), 0, 0),
                                                   ((130, 100), 1, 0)])
def test_maxwell_march_bitwise(mk, shape, periodic, freeze):
    """Maxwell as a z march (63 output columns per wavefront, chunk ends that
    wrap in z, walls and periodic x, sponge and xi cut, Az1 frozen or not)
    gives the oracle's bits."""
    nx, nz = shape
    keys = grid(nx, nz, periodic, maxwell_kernel=mk, autotune=0, xi_sponge=8.0, xi_cut=12.0,
                freeze_override=freeze)
    cfg, sim, o = make_pair(random_state(nx, nz, seed=nx * 7 + mk, smooth=True), **keys)
    assert sim.tuning()["maxwell"] == mk
    o.sources()
    sim.step(3)
    o.step(3)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"maxwell {mk} {shape} {n}")


@pytest.mark.parametrize("xzv", list(range(15)))
@pytest.mark.parametrize("shape,periodic,warm", [((300, 136), 1, 0.01), ((70, 36), 0, 0.0), ((256, 64), 0, 0.0),
                                                 ((1000, 40), 0, 0.01), ((116, 300), 1, 0.0)])
def test_fused_sweep_variants_bitwise(xzv, shape, periodic, warm):
    """Every fused x+z sweep kernel variant (LDS-staged x sweep; register x
    sweep inside the z march: several wave columns per row, both x ends,
    columns a multiple of the 58 a wave stores, several z chunks) gives the
    oracle's bits."""
    nx, nz = shape
    if xzv >= 13 and warm > 0:
        pytest.skip("the ring sweep (13, 14) takes no warm terms")
    keys = grid(nx, nz, periodic, warm_T=warm, fused_sweep=1, xz_variant=xzv, autotune=0)
    cfg, sim, o = make_pair(random_state(nx, nz, seed=nx + xzv, smooth=True), **keys)
    assert sim.tuning()["xz_variant"] == xzv
    o.sources()
    sim.step(4)
    o.step(4)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"xz{xzv} {nx}x{nz} {n}")


MXZ_CASES = {
    # frozen Az1, xi_cut: the headline's step on the C1 shape
    "walled_c1_xicut": dict(C1, n_steps=0, xi_cut=5.0),
    "periodic_sponge_unfrozen": grid(300, 136, 1, xi_sponge=4.0, sigma_sponge=5.0, freeze_override=0, alpha_YM=2.0),
    "walled_odd_cut_unfrozen": grid(70, 36, 0, xi_cut=2.0, freeze_override=0),
    # many wave columns, both walls, chunks taller than the grid
    "walled_wide_short": grid(1000, 20, 0, xi_cut=30.0),
    # several chunks, the last one ragged; columns not a multiple of 58
    "periodic_tall_nofilter": grid(116, 300, 1),
    "walled_58_columns": grid(116, 64, 0, xi_sponge=3.0, sigma_sponge=2.0),
    # one wave column holds both walls / wraps around the periodic grid
    "walled_narrow": grid(40, 24, 0, xi_cut=1.0),
    "periodic_narrow_short": grid(50, 12, 1, xi_sponge=1.0, sigma_sponge=1.0, freeze_override=0),
}


@pytest.mark.parametrize("v", list(range(9)))
@pytest.mark.parametrize("case", list(MXZ_CASES))
def test_maxwell_sweep_launch_bitwise(case, v):
    """Maxwell of one step and the sweeps of the next in one launch (every
    variant, forced): 7 steps with an energy check every 5 -- a 5-step
    segment replayed as a captured graph, then 2 eager steps -- give the
    oracle's bits, the sponge / xi_cut of the charges applied inside the
    launch included."""
    keys = dict(MXZ_CASES[case], fused_sweep=1, step_fusion=1, mxz_variant=v, autotune=0,
                energy_stride=5)
    nx, nz = keys["nx_override"], keys["nz_override"]
    cfg, sim, o = make_pair(random_state(nx, nz, seed=nx + 7 * v, smooth=True), **keys)
    o.sources()
    assert not sim.step(7), "halted on the energy check"
    o.step(7)
    t = sim.tuning()
    assert t["step_fusion"] == 1 and t["mxz_variant"] == v
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"mxz{v} {case} {n}")


def test_maxwell_sweep_launch_eager_segments():
    """The same launch without graphs and with segments that start in the
    middle of an energy stride (khs_step(3) twice, then 1, then 4), FP64,
    against the oracle; and the tuner's choice (step_fusion = -1) as well."""
    for fusion in (1, -1):
        keys = grid(300, 136, 0, xi_cut=8.0, fused_sweep=1, step_fusion=fusion, step_graphs=0,
                    energy_stride=4)
        cfg, sim, o = make_pair(random_state(300, 136, seed=41, smooth=True), **keys)
        o.sources()
        for k in (3, 3, 1, 4):
            assert not sim.step(k), "halted on the energy check"
            o.step(k)
        assert sim.tuning()["step_fusion"] in ((1,) if fusion == 1 else (0, 1))
        for n in orc.STATE:
            assert_bitwise(sim.get(n), o.get(n), f"fusion {fusion} {n}")


@pytest.mark.parametrize("extra,want", [(dict(fused_sweep=1), 1), (dict(fused_sweep=0), 0),
                                        (dict(fused_sweep=-1, z_chunk=32), 0), (dict(fused_sweep=-1), None),
                                        (dict(fused_sweep=-1, autotune=0), 1)])
def test_sweep_form_choice(extra, want):
    """The sweep form in use (khs_get_tuning): forced by fused_sweep = 0 / 1,
    split when a split-sweep tile shape is forced, fused without tuning, the
    timed choice otherwise; every form gives the oracle's bits."""
    keys = grid(300, 136, 1, **extra)
    cfg, sim, o = make_pair(random_state(300, 136, seed=17, smooth=True), **keys)
    o.sources()
    sim.step(2)
    o.step(2)
    got = sim.tuning()["fused_sweep"]
    assert got in (0, 1) and (want is None or got == want)
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"{extra} {n}")


def test_determinism_and_staged_equals_fused():
    keys = grid(256, 128, 1, hyp_diff=1e-3, xi_sponge=6.0)
    st = random_state(256, 128, seed=21, smooth=True)
    _, a, _ = make_pair(st, **keys)
    _, b, _ = make_pair(st, **keys)
    a.step(5)
    for _ in range(5):
        b.fct_x_sweep(0); b.fct_x_sweep(1); b.fct_z_sweep(0); b.fct_z_sweep(1)
        b.maxwell(); b.filters(); b.commit()
    for n in orc.STATE:
        assert_bitwise(a.get(n), b.get(n), n)


def test_nan_halts_and_errors():
    cfg, sim = make_sim(**grid(64, 32, 1, energy_stride=2))
    v = sim.get("By2")
    v[3, 5] = np.nan
    sim.set("By2", v)
    assert sim.step(10) is True
    assert sim.step_count == 2
    with pytest.raises(khs.KhsError) as e:
        sim.fct_z_sweep(0)
    assert e.value.code == khs.KHS_ERR_STATE
    with pytest.raises(khs.KhsError):
        sim.fct_x_sweep(2)
    with pytest.raises(khs.KhsError):
        sim.get(99)


@pytest.mark.parametrize("graphs", [1, 0])
def test_stride_energy_summed_by_last_maxwell(graphs):
    """With Maxwell fused to the next step's sweeps, the energy of each check
    is summed by the stride's last Maxwell launch as it writes the state: the
    value the halting rule saw (khs_last_energy) equals the separate energy
    reduction of the same state and the oracle's energy to 1e-13, and the
    state itself stays the oracle's to the bit."""
    keys = grid(300, 136, 0, xi_cut=8.0, fused_sweep=1, step_fusion=1, maxwell_kernel=10,
                step_graphs=graphs, energy_stride=5)
    cfg, sim, o = make_pair(random_state(300, 136, seed=17, smooth=True), **keys)
    o.sources()
    assert sim.last_energy_check()[1] == -1
    for n_steps in (5, 5):
        assert not sim.step(n_steps), "halted on the energy check"
        o.step(n_steps)
        E, at = sim.last_energy_check()
        assert at == sim.step_count
        E_sep, E_orc = sim.energy(), o.energy()
        assert abs(E - E_sep) <= 1e-13 * abs(E_sep), (E, E_sep)
        assert abs(E - E_orc) <= 1e-13 * abs(E_orc), (E, E_orc)
    assert sim.tuning()["step_fusion"] == 1
    for n in orc.STATE:
        assert_bitwise(sim.get(n), o.get(n), f"energy-fused {n}")
