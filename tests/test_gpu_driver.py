"""The khs_run driver (the ym_coupled command line, main_ym.cu:18-491): an
.ini run writes energy.csv and ym_%06d.csv snapshots into the reference's
output directory, the snapshot matches the library's own writer on a
library run of the same config, and --checkpoint-every / --restart continue
a run with byte-identical output."""
import filecmp
import os
import shutil
import subprocess

import pytest

import khs

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(khs.LIB_PATH), "khs_run")
INI = """# small walled run
nx_override = 128
nz_override = 64
lx_override = 12.8
lz_override = 6.4
alpha_YM = 1.0
V0 = 0.05
eps_override = 1.0
courant_override = 0.4
run_mode = 0
freeze_override = 1
perturb_amp = 0.01
noise_amp = 1e-9
noise_seed = 2
kz_keep_max = 4
xi_sponge = 4.0
n_steps = {n}
energy_stride = 5
export_stride = 10
"""


def run(cwd, n, *extra):
    ini = os.path.join(cwd, "run.ini")
    with open(ini, "w") as f:
        f.write(INI.format(n=n))
    print("khs_run", cwd, n, extra, flush=True)
    r = subprocess.run([DRIVER, ini, *extra], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    cfg = khs.Config.from_ini(ini)
    return cfg, os.path.join(cwd, cfg.output_dir())


def test_driver_outputs_match_library(tmp_path):
    cfg, out = run(str(tmp_path), 20)
    names = sorted(os.listdir(out))
    assert names == ["energy.csv", "ym_000000.csv", "ym_000010.csv", "ym_000020.csv"]
    rows = open(os.path.join(out, "energy.csv")).read().splitlines()
    assert len(rows) == 1 + 1 + 4                 # header, step 0, steps 5..20
    assert rows[-1].startswith("20,")
    sim = khs.Simulation(cfg)
    sim.init()
    sim.step(20)
    lib_dir = str(tmp_path / "lib")
    os.makedirs(lib_dir)
    sim.write_snapshot(lib_dir, 20)
    sim.close()
    assert filecmp.cmp(os.path.join(out, "ym_000020.csv"), os.path.join(lib_dir, "ym_000020.csv"), shallow=False)


def test_driver_checkpoint_restart(tmp_path):
    full, half, cont = (str(tmp_path / d) for d in ("full", "half", "cont"))
    for d in (full, half, cont):
        os.makedirs(d)
    _, out_full = run(full, 20)
    _, out_half = run(half, 10, "--checkpoint-every", "10")
    ckpt = os.path.join(out_half, "checkpoint.khs")
    assert os.path.exists(ckpt)
    shutil.copy(ckpt, os.path.join(cont, "c.khs"))
    _, out_cont = run(cont, 20, "--restart", os.path.join(cont, "c.khs"))
    assert filecmp.cmp(os.path.join(out_full, "ym_000020.csv"), os.path.join(out_cont, "ym_000020.csv"),
                       shallow=False)
    assert not os.path.exists(os.path.join(out_cont, "ym_000000.csv"))


SNAP_COLUMNS = ["By1", "By2", "By3", "Az2", "Az3", "PzA", "PxA", "Q1A", "PzB", "PxB", "Q1B",
                "Ex1", "Ex2", "Ex3", "Ez1", "Ez2", "Ez3", "nA", "nB", "Q2A", "Q3A", "Q2B", "Q3B", "Az1"]


def test_snapshot_matches_an_independent_6_digit_formatter(tmp_path):
    # main_ym.cu:287-306 streams doubles through std::ostream's defaults:
    # %g with 6 significant digits ("1e-09", "-0", "0.05"), X = x*DX, Z = z*DZ,
    # x fastest, the 26 columns of :291-292.  Python's %-formatting is an
    # independent implementation of the same conversion.
    ini = tmp_path / "run.ini"
    ini.write_text(INI.format(n=20))
    cfg = khs.Config.from_ini(str(ini))
    sim = khs.Simulation(cfg)
    sim.init()
    sim.step(7)
    arr = sim.get("PzA")
    arr[0, 0] = -0.0                                 # the sign of zero is printed
    sim.set("PzA", arr)
    sim.write_snapshot(str(tmp_path), 7)
    cols = [sim.get(n) for n in SNAP_COLUMNS]
    lines = ["X,Z," + ",".join(SNAP_COLUMNS)]
    for z in range(cfg.nz):
        for x in range(cfg.nx):
            vals = [x * cfg.dx, z * cfg.dz] + [c[z, x] for c in cols]
            lines.append(",".join("%g" % v for v in vals))
    sim.close()
    got = (tmp_path / "ym_000007.csv").read_text()
    assert got == "\n".join(lines) + "\n"
    assert got.splitlines()[1].split(",")[7] == "-0"
