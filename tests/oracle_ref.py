"""ctypes wrapper of the FP64 CPU oracle (oracle/khs_oracle.cpp) — test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module.  It never stands in for the GPU library.

How each oracle function is pinned (tests that fail on a plausible mistake):

  fct_row (FCT_Sweeps.cu:39-177)       test_oracle_fct.py: linear profiles are
      invariant (dropped diffusion/antidiffusion term), total mass is conserved
      on a periodic ring (index/flux-form error), a Gaussian advected at
      constant eps moves eps*steps cells in the right direction (sign / dt/dx
      scale), mirror symmetry x -> -x with u -> -u (limiter sign S = sign of
      LNRHOT difference), positivity/no new extrema for a top hat (limiter),
      nu/mu closed forms at eps = 0 and |eps| >= 1/6.
  x_sweep / z_sweep (YM_FCT.cu)        periodic x-sweep equals the ring sweep
      applied row by row; z-sweep equals fct_row on each column (transpose
      semantics); clamped-end semantics checked on a ramp.
  currents / ampere / potential /      test_oracle_maxwell.py: Abelian vacuum
  faraday (YM_Physics.cu)              (alpha=0, no fluid current) Yee plane
      wave follows the closed-form Yee dispersion relation; SO(3) colour
      rotation covariance of a full non-Abelian step (epsilon^abc signs);
      current of a uniform drifting fluid equals -n Q v exactly.
  lorentz / precession (YM_Fluid.cu)   closed forms on uniform fields, colour
      covariance, sQ orthogonal to Q (|Q| conserved to first order).
  sponge / xicut / zmean / bands /     test_oracle_filters.py: planted Fourier
  keep_modes / hyperdiffuse            modes removed or kept exactly, band
      edges (the skipped k = kz_suppress_max+1), Nyquist 2x over-subtraction,
      idempotence of the projection, hyperdiffusion eigenvalue
      1 - mu*16 sin^4(pi k/nz), sponge decay closed form.
  energy / gauss                        closed forms on constructed states
      (uniform drift, divergence of an analytic field).
  init (YM_Init.cu)                     closed forms: n = 1, Q = +-1, pz = V0 tanh,
      Az1 = -V0 log cosh (float math), seed window |xi| < 3.
"""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB = os.path.join(ORACLE_DIR, "liborc.so")

IDS = {n: i for i, n in enumerate(
    ["Ex1", "Ex2", "Ex3", "Ez1", "Ez2", "Ez3", "By1", "By2", "By3", "Az1", "Az2", "Az3",
     "nA", "PxA", "PzA", "Q1A", "Q2A", "Q3A", "nB", "PxB", "PzB", "Q1B", "Q2B", "Q3B"])}
IDS.update({n: 32 + i for i, n in enumerate(
    ["Jx1", "Jx2", "Jx3", "Jz1", "Jz2", "Jz3", "uxA", "uzA", "uxB", "uzB",
     "srcpxA", "srcpzA", "srcpxB", "srcpzB", "sQ1A", "sQ2A", "sQ3A", "sQ1B", "sQ2B", "sQ3B",
     "energy_density"])})
STATE = [n for n, i in IDS.items() if i < 24]


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("periodic_x", ctypes.c_int), ("freeze_az1", ctypes.c_int),
        ("suppress_kz0", ctypes.c_int), ("kz_suppress_max", ctypes.c_int), ("kz_suppress_hi", ctypes.c_int),
        ("kz_keep_max", ctypes.c_int),
        ("dx", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double), ("c", ctypes.c_double),
        ("eps0", ctypes.c_double), ("alpha", ctypes.c_double), ("V0", ctypes.c_double), ("eps", ctypes.c_double),
        ("xi_sponge", ctypes.c_double), ("sigma_sponge", ctypes.c_double), ("xi_cut", ctypes.c_double),
        ("warm_T", ctypes.c_double), ("hyp_diff", ctypes.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    v, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    lib.orc_create.restype = v
    lib.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
    lib.orc_destroy.argtypes = [v]
    lib.orc_array.restype = dp
    lib.orc_array.argtypes = [v, i]
    for n in ["orc_update_velocities", "orc_enforce_boundaries", "orc_currents", "orc_ampere", "orc_potential",
              "orc_faraday", "orc_filters", "orc_sources"]:
        getattr(lib, n).argtypes = [v]
    for n in ["orc_x_sweep", "orc_z_sweep", "orc_step"]:
        getattr(lib, n).argtypes = [v, i]
    lib.orc_energy.restype = d
    lib.orc_energy.argtypes = [v]
    lib.orc_gauss.argtypes = [v, dp]
    lib.orc_fct_row.argtypes = [i, dp, dp, dp, dp, d, d]
    lib.orc_kz_band_column.argtypes = [dp, i, i, i]
    lib.orc_stream_triad.argtypes = [ctypes.c_long, i]
    lib.orc_stream_triad.restype = d
    lib.orc_init.argtypes = [v, i, d, d, d, i, d, i, d, dp, dp, dp, dp]
    lib.orc_noise.restype = d
    lib.orc_seed_profiles.restype = i
    lib.orc_seed_profiles.argtypes = [ctypes.c_char_p, i, dp, ctypes.POINTER(ctypes.c_int)]
    lib.orc_noise.argtypes = [ctypes.c_uint64, i, ctypes.c_uint64, d]
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params(nx, nz, dx, dz, dt, alpha=1.0, V0=0.05, eps=1.0, periodic_x=0, freeze_az1=1,
           suppress_kz0=0, kz_suppress_max=0, kz_suppress_hi=0, kz_keep_max=0,
           xi_sponge=0.0, sigma_sponge=5.0, xi_cut=0.0, warm_T=0.0, hyp_diff=0.0):
    return OrcParams(nx, nz, periodic_x, freeze_az1, suppress_kz0, kz_suppress_max, kz_suppress_hi,
                     kz_keep_max, dx, dz, dt, 1.0, 1.0, alpha, V0, eps, xi_sponge, sigma_sponge, xi_cut,
                     warm_T, hyp_diff)


def params_from_config(cfg):
    """OrcParams from a derived khs.Config (same double values the GPU uses)."""
    return OrcParams(cfg.nx, cfg.nz, cfg.periodic_x, cfg.freeze_az1, cfg.suppress_kz0, cfg.kz_suppress_max,
                     cfg.kz_suppress_hi, cfg.kz_keep_max, cfg.dx, cfg.dz, cfg.dt, cfg.c, cfg.eps_0, cfg.alpha,
                     cfg.v0, cfg.eps, float(cfg.xi_sponge), float(cfg.sigma_sponge), float(cfg.xi_cut),
                     float(cfg.warm_T), float(cfg.hyp_diff))


class Oracle:
    def __init__(self, p):
        self.lib = load()
        self.p = p
        self.h = self.lib.orc_create(ctypes.byref(p))
        self.nx, self.nz = p.nx, p.nz

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arr(self, name):
        """numpy view (nz, nx) on the oracle's storage."""
        i = IDS[name] if isinstance(name, str) else name
        ptr = self.lib.orc_array(self.h, i)
        return np.ctypeslib.as_array(ptr, shape=(self.nz, self.nx))

    def get(self, name):
        return self.arr(name).copy()

    def set(self, name, a):
        self.arr(name)[...] = np.asarray(a, dtype=np.float64).reshape(self.nz, self.nx)

    def set_state(self, st):
        for n, a in st.items():
            self.set(n, a)

    def state(self):
        return {n: self.get(n) for n in STATE}

    def init(self, k_mode, perturb_amp, V0, eps, run_mode, alpha, init_by1_eq=0, vz_edge_taper=0.0,
             seed_by=None, seed_az=None, seed_qA=None, seed_qB=None):
        def opt(a):
            return None if a is None else _ptr(np.ascontiguousarray(a, dtype=np.float64))
        keep = [seed_by, seed_az, seed_qA, seed_qB]
        self.lib.orc_init(self.h, k_mode, perturb_amp, V0, eps, run_mode, alpha, init_by1_eq,
                          vz_edge_taper, *[opt(a) for a in keep])

    def init_from_config(self, cfg):
        """kernel_ym_init arguments exactly as main_ym.cu:223-225 passes them."""
        f32 = np.float32
        self.init(cfg.k_mode, float(f32(cfg.perturb_amp)), float(f32(cfg.V0)), cfg.eps, cfg.run_mode,
                  float(f32(cfg.alpha)), cfg.init_by1_eq, float(f32(cfg.vz_edge_taper)))

    # operators
    def update_velocities(self): self.lib.orc_update_velocities(self.h)
    def x_sweep(self, s): self.lib.orc_x_sweep(self.h, s)
    def z_sweep(self, s): self.lib.orc_z_sweep(self.h, s)
    def enforce_boundaries(self): self.lib.orc_enforce_boundaries(self.h)
    def currents(self): self.lib.orc_currents(self.h)
    def ampere(self): self.lib.orc_ampere(self.h)
    def potential(self): self.lib.orc_potential(self.h)
    def faraday(self): self.lib.orc_faraday(self.h)
    def filters(self): self.lib.orc_filters(self.h)
    def sources(self): self.lib.orc_sources(self.h)
    def step(self, n=1): self.lib.orc_step(self.h, n)
    def energy(self): return self.lib.orc_energy(self.h)

    def gauss(self):
        out = np.zeros(7)
        self.lib.orc_gauss(self.h, _ptr(out))
        return out


def stream_triad(n=1 << 25, reps=5):
    """Host STREAM-triad bandwidth (GB/s) with the oracle's OpenMP threads."""
    return load().orc_stream_triad(n, reps)


def fct_row(u, v, s, dx, dt):
    lib = load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    sp = None if s is None else _ptr(np.ascontiguousarray(s, dtype=np.float64))
    lib.orc_fct_row(len(v), _ptr(u), _ptr(v), sp, _ptr(out), dx, dt)
    return out


def kz_band_column(col, lo, hi):
    lib = load()
    c = np.ascontiguousarray(col, dtype=np.float64).copy()
    lib.orc_kz_band_column(_ptr(c), len(c), lo, hi)
    return c


def noise(seed, aid, cell, amp):
    return load().orc_noise(seed, aid, cell, amp)


SEED_FIELDS = ["by", "ex", "ez", "az", "qA", "qB"]


def seed_profiles(path, nx):
    """Seed-file profiles resampled onto nx points (main_ym.cu:178-217), as
    {field: array} for the fields the file holds; None if it cannot be read."""
    out = np.zeros(6 * nx)
    present = (ctypes.c_int * 6)()
    if load().orc_seed_profiles(os.fsencode(path), nx, _ptr(out), present) != 0:
        return None
    return {f: out[k * nx:(k + 1) * nx].copy() for k, f in enumerate(SEED_FIELDS) if present[k]}


def init_seed_args(prof):
    """orc_init keyword arguments from seed_profiles() output."""
    return {"seed_" + f: prof.get(f) for f in ("by", "az", "qA", "qB")}


# --------------------------------------------------------------------------
# EMHD oracle (oracle/khs_oracle_emhd.cpp)
# --------------------------------------------------------------------------
EM_IDS = {n: i for i, n in enumerate(["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Ax", "Ay", "Az", "Jx", "Jy", "Jz"])}
for _s, _suf in ((0, ""), (1, "2")):
    for _k, _n in enumerate(["n", "px", "py", "pz", "ux", "uz", "sx", "sy", "sz"]):
        EM_IDS[_n + _suf] = 16 + 12 * _s + _k
EM_IDS["em_energy_density"] = 40
EM_FIELDS = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Ax", "Ay", "Az"]


def em_state_names(nsp):
    f = ["n", "px", "py", "pz"]
    return EM_FIELDS + f + ([n + "2" for n in f] if nsp > 1 else [])


class EmParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int), ("periodic_x", ctypes.c_int), ("nsp", ctypes.c_int),
                ("dx", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double), ("c", ctypes.c_double),
                ("eps0", ctypes.c_double), ("q0", ctypes.c_double), ("q1", ctypes.c_double)]


def em_params_from_config(cfg):
    return EmParams(cfg.nx, cfg.nz, cfg.periodic_x, cfg.species, cfg.dx, cfg.dz, cfg.dt, cfg.c, cfg.eps_0,
                    cfg.charge1, cfg.charge2)


def _load_em():
    lib = load()
    if getattr(lib, "_em_ready", False):
        return lib
    v, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.emo_create.restype = v
    lib.emo_create.argtypes = [ctypes.POINTER(EmParams)]
    lib.emo_destroy.argtypes = [v]
    lib.emo_array.restype = ctypes.POINTER(ctypes.c_double)
    lib.emo_array.argtypes = [v, i]
    for n in ["emo_update_velocities", "emo_enforce_boundaries", "emo_currents", "emo_ampere",
              "emo_potential", "emo_faraday", "emo_lorentz"]:
        getattr(lib, n).argtypes = [v]
    for n in ["emo_x_sweep", "emo_z_sweep", "emo_step"]:
        getattr(lib, n).argtypes = [v, i]
    lib.emo_energy.restype = d
    lib.emo_energy.argtypes = [v]
    lib.emo_init_daskaw.argtypes = [v, i, d, i]
    lib._em_ready = True
    return lib


class EmOracle:
    """The emhdgpu2d step on the CPU.  After set_state / init call lorentz()
    once (main_coupled.cu:282) before stepping, as the reference does."""

    def __init__(self, p):
        self.lib = _load_em()
        self.p = p
        self.h = self.lib.emo_create(ctypes.byref(p))
        self.nx, self.nz, self.nsp = p.nx, p.nz, p.nsp

    def close(self):
        if self.h:
            self.lib.emo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arr(self, name):
        i = EM_IDS[name] if isinstance(name, str) else name
        return np.ctypeslib.as_array(self.lib.emo_array(self.h, i), shape=(self.nz, self.nx))

    def get(self, name):
        return self.arr(name).copy()

    def set(self, name, a):
        self.arr(name)[...] = np.asarray(a, dtype=np.float64).reshape(self.nz, self.nx)

    def set_state(self, st):
        for n, a in st.items():
            self.set(n, a)

    def state(self):
        return {n: self.get(n) for n in em_state_names(self.nsp)}

    def init_daskaw(self, k_mode, perturb_amp, smooth=1):
        self.lib.emo_init_daskaw(self.h, k_mode, perturb_amp, smooth)

    def update_velocities(self): self.lib.emo_update_velocities(self.h)
    def x_sweep(self, s): self.lib.emo_x_sweep(self.h, s)
    def z_sweep(self, s): self.lib.emo_z_sweep(self.h, s)
    def enforce_boundaries(self): self.lib.emo_enforce_boundaries(self.h)
    def currents(self): self.lib.emo_currents(self.h)
    def ampere(self): self.lib.emo_ampere(self.h)
    def potential(self): self.lib.emo_potential(self.h)
    def faraday(self): self.lib.emo_faraday(self.h)
    def lorentz(self): self.lib.emo_lorentz(self.h)
    def step(self, n=1): self.lib.emo_step(self.h, n)
    def energy(self): return self.lib.emo_energy(self.h)
