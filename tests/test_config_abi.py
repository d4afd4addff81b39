"""CPU tests of the host side of libkhs: config parsing and derivation
(YM_Config.cu, main_ym.cu:51-83, 143-176, 244-249), the C ABI surface, and the
oracle's initial state against closed forms.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import khs
import oracle_ref as orc

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")


def test_library_exports_every_declared_symbol():
    lib = khs.load_library()
    hdr = open(os.path.join(REPO, "include", "khs.h")).read()
    declared = set(re.findall(r"\b(khs_[a-z0-9_]+)\s*\(", hdr))
    assert len(declared) >= 30
    raw = ctypes.CDLL(khs.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(raw, name), f"libkhs.so does not export {name}"
    assert set(khs.exported_symbols()) == declared
    assert lib.khs_config_sizeof() == ctypes.sizeof(khs.khs_config)


def test_defaults_match_reference():
    c = khs.Config()
    assert (c.nx, c.nz, c.k_mode, c.run_mode) == (768, 64, 1, 6)
    assert c.lx == 6.0 * np.pi and c.lz == 2.0 * np.pi
    assert c.courant == float(np.float32(0.1))          # 0.1f widened (main_ym.cu:57)
    assert c.alpha == float(np.float32(0.5)) and c.v0 == float(np.float32(0.1))
    assert c.eps == c.lx / 6.0
    assert c.periodic_x == 1 and c.freeze_az1 == 1
    assert c.total_steps == 2000000 and c.energy_stride == 10000 and c.export_stride == 20000
    assert c.energy_thresh == 100.0 or c.energy_thresh == 5.0


def test_float_parse_quirk_gives_1999_steps():
    # courant 0.4f * dx(25.6f/256) = 0.0400000011920929, (int)(80/DT) = 1999
    c = khs.Config(nx_override=256, nz_override=128, lx_override=25.6, lz_override=12.8,
                   courant_override=0.4, target_tu=80, run_mode=0)
    assert c.dt == float(np.float32(0.4)) * (float(np.float32(25.6)) / 256)
    assert abs(c.dt - 0.0400000011920929) < 1e-16
    assert c.total_steps == 1999
    assert c.energy_stride == 12 and c.export_stride == 24
    assert c.periodic_x == 0 and c.freeze_az1 == 0 and c.energy_thresh == 5.0
    c2 = khs.Config(nx_override=256, nz_override=128, lx_override=25.6, lz_override=12.8,
                    courant_override=0.4, target_tu=80, n_steps=2000)
    assert c2.total_steps == 2000
    # n_steps alone also selects the physical cadences
    c3 = khs.Config(nx_override=256, nz_override=128, lx_override=25.6, lz_override=12.8,
                    courant_override=0.4, n_steps=2000)
    assert (c3.total_steps, c3.energy_stride, c3.export_stride) == (2000, 12, 24)


def test_ini_file_and_inline_comments(tmp_path):
    p = tmp_path / "run.ini"
    p.write_text("# comment\n k_mode = 2   # inline\nalpha_YM=1.5\nrun_mode = 1\nV0 = 0.05\n"
                 "suppress_kz0 = 1\nnz_override = 128\nhyp_diff = 5e-5\nxi_sponge = 13.0\n"
                 "kz_suppress_hi = 14\nrun_tag = demo\nbogus_key = 3\nno equals sign here\n")
    c = khs.Config.from_ini(str(p))
    assert (c.k_mode, c.run_mode, c.suppress_kz0, c.nz, c.kz_suppress_hi) == (2, 1, 1, 128, 14)
    assert c.alpha_YM == np.float32(1.5) and c.hyp_diff == np.float32(5e-5)
    assert c.run_tag == "demo"
    assert c.energy_thresh == 100.0       # sponge active
    assert c.output_dir() == "ym_k2_a1.500_circ_v0.0500_sp13.0_nkz0_bp14_hd5e-05_demo"


def test_output_dir_name_c1():
    c = khs.Config(nx_override=256, nz_override=128, lx_override=25.6, lz_override=12.8, alpha_YM=1,
                   V0=0.05, eps_override=1.0, courant_override=0.4, k_mode=1, run_mode=0,
                   freeze_override=1, perturb_amp=2e-5)
    assert c.output_dir() == "ym_k1_a1.000_v0.0500_frz_eps1.00_nx256_lx25.60"
    # V0 = 0.1 is stored as 0.1f != 0.1, so the reference always appends _v0.1000
    d = khs.Config(V0=0.1, perturb_amp=0).output_dir()
    assert d == "ym_k1_a0.500_noperturb_circ_az2seed_v0.1000"


@pytest.mark.parametrize("keys,msg", [
    (dict(suppress_kz0=1, nz_override=96), "power of 2"),
    (dict(kz_suppress_max=2, nz_override=40), "multiple of 32"),
    (dict(kz_keep_max=40, nz_override=64), "kz_keep_max"),
    (dict(run_mode=9), "run_mode"),
    (dict(x_lane_cells=6), "x_lane_cells"),
    (dict(z_groups=2), "z_groups"),
    (dict(fused_sweep=2), "fused_sweep"),
    (dict(xz_variant=15), "xz_variant"),
    (dict(xz_variant=13, warm_T=0.01), "xz_variant"),
    (dict(step_fusion=2), "step_fusion"),
    (dict(mxz_variant=9), "mxz_variant"),
    (dict(packed_shape=8), "packed_shape"),
    (dict(maxwell_kernel=11), "maxwell_kernel"),
    (dict(nx_override=32768, nz_override=16384), "grid too large"),
    (dict(model="emhd", nx_override=32768, nz_override=16384), "grid too large"),
    (dict(nx_override=64, nz_override=1 << 23, run_mode=6), "grid too large"),   # 2^29 cells: 2^32 bytes
    (dict(model="emhd", xz_variant=7), "xz_variant"),
])
def test_constraint_errors_are_codes_not_exits(keys, msg):
    with pytest.raises(khs.KhsError) as e:
        khs.Config(**keys)
    assert e.value.code == khs.KHS_ERR_CONFIG and msg in str(e.value)


def test_largest_grid_keeps_the_store_sentinel_out_of_range():
    # halo lanes store at byte offset 0xFFFFFFF0, which the buffer range
    # check must drop: the largest admitted FP64 array is exactly that size
    c = khs.Config(nx_override=32766, nz_override=16385, run_mode=6)     # 2^29 - 2 cells
    assert c.nx * c.nz * 8 == 0xFFFFFFF0


def test_missing_ini_and_unknown_key():
    with pytest.raises(khs.KhsError) as e:
        khs.Config.from_ini("/nonexistent/x.ini")
    assert e.value.code == khs.KHS_ERR_CONFIG
    with pytest.raises(khs.KhsError):
        khs.Config().set("not_a_key", "1")
    with pytest.raises(khs.KhsError):
        khs.Config().set("k_mode", "abc")


def test_reference_example_ini_parses():
    ref = "/root/reference/ymgpu2d/example.ini"
    if not os.path.exists(ref):
        pytest.skip("reference checkout not present")
    c = khs.Config.from_ini(ref)
    assert (c.k_mode, c.run_mode, c.nz, c.kz_suppress_hi, c.suppress_kz0) == (1, 6, 128, 14, 1)
    assert c.alpha_YM == np.float32(2.0) and c.eps_override == np.float32(0.15)


def test_oracle_init_mode0_closed_forms():
    # kernel_ym_init run_mode 0 (YM_Init.cu:96-99, 271-285): n = 1, Q1 = +-1,
    # pz = +-V0 tanh(xi), Az1 = -V0 log cosh(xi), By2 = amp V0 sech(xi) [|xi|<3] sin(kz z)
    nx, nz, dx, dz, V0, eps, amp = 64, 32, 0.1, 0.1, 0.05, 1.0, 2e-5
    o = orc.Oracle(orc.params(nx, nz, dx, dz, 0.04, V0=V0, eps=eps))
    o.init(1, amp, V0, eps, 0, 1.0)
    x = np.arange(nx) * dx
    xi = (x - nx * dx / 2) / eps
    z = np.arange(nz) * dz
    kz = 2 * np.pi / (nz * dz)
    assert np.all(o.get("nA") == 1.0) and np.all(o.get("nB") == 1.0)
    assert np.all(o.get("Q1A") == 1.0) and np.all(o.get("Q1B") == -1.0)
    np.testing.assert_allclose(o.get("PzA")[0], V0 * np.tanh(xi), rtol=3e-7, atol=1e-12)
    np.testing.assert_allclose(o.get("PzB")[5], -V0 * np.tanh(xi), rtol=3e-7, atol=1e-12)
    np.testing.assert_allclose(o.get("Az1")[3], -V0 * np.log(np.cosh(xi)), rtol=1e-6, atol=1e-9)
    win = (np.abs(xi) < 3).astype(float)
    by2 = amp * V0 * np.outer(np.sin(kz * z), win / np.cosh(xi))
    np.testing.assert_allclose(o.get("By2"), by2, rtol=1e-6, atol=1e-6 * amp * V0)
    for n in ("Ex1", "Ez3", "By1", "By3", "Az2", "PxA", "Q2B"):
        assert np.all(o.get(n) == 0.0)


def test_noise_is_uniform_and_keyed():
    v = np.array([orc.noise(7, 10, c, 1e-9) for c in range(20000)])
    assert np.all(np.abs(v) <= 1e-9)
    assert abs(v.mean()) < 2e-11 and abs(v.std() - 1e-9 / np.sqrt(3)) < 2e-11
    w = np.array([orc.noise(8, 10, c, 1e-9) for c in range(100)])
    assert not np.allclose(w, v[:100], rtol=1e-3, atol=0)


def test_emhd_config_matches_reference_run():
    # emhdgpu2d/main_coupled.cu:228-262: 256x256, dx = dz = 1, dt = 0.002,
    # 500000 steps, energy every 500, export every 1000, halt at E/E0 > 5
    c = khs.Config(model="emhd")
    assert (c.nx, c.nz, c.dx, c.dz, c.dt) == (256, 256, 1.0, 1.0, 0.002)
    assert (c.total_steps, c.energy_stride, c.export_stride, c.energy_thresh) == (500000, 500, 1000, 5.0)
    assert (c.periodic_x, c.species, c.charge1, c.em_perturb_amp) == (0, 1, -1.0, 0.05)
    assert c.eps == 10.0 and c.v0 == 0.1 and c.c == 1.0 and c.eps_0 == 1.0
    assert c.output_dir() == "coupled_k1"
    assert khs.Config(model="emhd", k_mode=4, perturb_amp=0).output_dir() == "coupled_k4_noperturb"
    # perturb_amp is a double for EMHD (atof), a float for YM (stof)
    c = khs.Config(model="emhd", perturb_amp="0.1")
    assert c.em_perturb_amp == 0.1 and c.perturb_amp == float(np.float32(0.1))
    for bad in (dict(species=3), dict(emhd_init="spectrum", spectrum_kmax=200)):
        with pytest.raises(khs.KhsError):
            khs.Config(model="emhd", **bad)


def test_emhd_array_ids():
    lib = khs.load_library()
    want = {"Ex": 64, "Az": 72, "n": 76, "pz": 79, "n2": 82, "pz2": 85, "Jx": 96, "uz2": 102, "sz2": 108,
            "em_energy_density": 109, "Ex1": 0, "energy_density": 52}
    for name, i in want.items():
        assert lib.khs_array_id(name.encode()) == i, name
    for name, i in khs.EM_ARRAY_IDS.items():
        if name != "energy_density":
            assert lib.khs_array_id(name.encode()) == i, name
