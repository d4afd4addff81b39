"""Writes the fixtures of tests/golden/ from data files the reference holds
(run in a checkout that has the reference at REF; never on the GPU host).

  eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin   copied byte for byte from
      ymgpu2d/eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin (int32 header
      {6, 768}, then 6 x 768 float32 profiles By, Ex, Ez, Az, QA, QB)
  eigenmode_kz1_a1.00_V0.05_EPS0.15_gamma.npz  the scalars of
      ymgpu2d/eigenmode_kz1_a1.00_V0.05_EPS0.15.npz (the linear eigensolver's
      growth rate at alpha = 1, V0 = 0.05, EPS = 0.15, kz = 1, NX = 768),
      read with numpy.load(allow_pickle=False).
  eigenmode_kz1_a1.00_V0.05_EPS0.15_vec.npz  the six complex profiles b, ex,
      ez, a, qA, qB of the same file (the eigenvector, 6 x 768).
  eigensolver_grid_cache.npz  kz_vals (80), series (8 x (alpha, V0,
      xi_sponge)) and gamma (8 x 80, NaN where no localised mode) of
      ymgpu2d/presentation/eigensolver_grid_cache.npz (NX = 384, EPS = 0.15,
      Lx = 6 pi; presentation/compute_eigensolver_grid.py).
"""
import os
import shutil
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = os.path.join(REF, "ymgpu2d")
    shutil.copyfile(os.path.join(src, "eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin"),
                    os.path.join(HERE, "eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin"))
    d = np.load(os.path.join(src, "eigenmode_kz1_a1.00_V0.05_EPS0.15.npz"), allow_pickle=False)
    np.savez(os.path.join(HERE, "eigenmode_kz1_a1.00_V0.05_EPS0.15_gamma.npz"),
             gamma=np.array([d["gamma"].real, d["gamma"].imag]), gamma_wkb=d["gamma_wkb"], alpha=d["alpha"],
             V0=d["V0"], EPS=d["EPS"], kz_int=d["kz_int"])
    np.savez_compressed(os.path.join(HERE, "eigenmode_kz1_a1.00_V0.05_EPS0.15_vec.npz"),
                        **{k: d[k] for k in ("b", "ex", "ez", "a", "qA", "qB")})
    g = np.load(os.path.join(src, "presentation", "eigensolver_grid_cache.npz"), allow_pickle=False)
    np.savez_compressed(os.path.join(HERE, "eigensolver_grid_cache.npz"),
                        kz_vals=g["kz_vals"], series=g["series"], gamma=g["gamma"])


if __name__ == "__main__":
    main()
