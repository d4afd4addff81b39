"""ctypes binding of libkhs, the MI355X (gfx950) Kelvin-Helmholtz time-step library.

This is the Python surface the ymgpu2d analysis scripts use instead of parsing
CSV snapshots: the same config keys (ymgpu2d/example.ini), the same operator
boundaries (YM_FCT / YM_Physics / YM_Fluid / YM_Energy) and the same output
files, driven through the C ABI declared in include/khs.h.

    import khs
    cfg = khs.Config.from_ini("run.ini")       # or khs.Config(nx_override=256, ...)
    sim = khs.Simulation(cfg)
    sim.init()
    halted = sim.step(100)
    By2 = sim.get("By2")                         # numpy (nz, nx), reference layout

The library is loaded from this directory (built by kh-simulations_amd/build.py).
There is no CPU fallback: without libkhs.so or without a GPU every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkhs.so")

KHS_OK, KHS_ERR_ARG, KHS_ERR_CONFIG, KHS_ERR_IO, KHS_ERR_DEVICE, KHS_ERR_STATE = 0, -1, -2, -3, -4, -5
FP64, FP32 = 0, 1
SPECIES_A, SPECIES_B = 0, 1
BUF_CURRENT, BUF_STAGE, BUF_NEXT = 0, 1, 2

STATE_NAMES = ["Ex1", "Ex2", "Ex3", "Ez1", "Ez2", "Ez3", "By1", "By2", "By3",
               "Az1", "Az2", "Az3", "nA", "PxA", "PzA", "Q1A", "Q2A", "Q3A",
               "nB", "PxB", "PzB", "Q1B", "Q2B", "Q3B"]
DERIVED_NAMES = ["Jx1", "Jx2", "Jx3", "Jz1", "Jz2", "Jz3", "uxA", "uzA", "uxB", "uzB",
                 "srcpxA", "srcpzA", "srcpxB", "srcpzB", "sQ1A", "sQ2A", "sQ3A",
                 "sQ1B", "sQ2B", "sQ3B", "energy_density"]
ARRAY_IDS = {n: i for i, n in enumerate(STATE_NAMES)}
ARRAY_IDS.update({n: 32 + i for i, n in enumerate(DERIVED_NAMES)})
ARRAY_IDS.update({"pxA": 13, "pzA": 14, "pxB": 19, "pzB": 20})
# EMHD model (model="emhd")
EM_STATE_NAMES = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Ax", "Ay", "Az", "n", "px", "py", "pz",
                  "n2", "px2", "py2", "pz2"]
EM_ARRAY_IDS = {n: 64 + i for i, n in enumerate(EM_STATE_NAMES[:9])}
EM_ARRAY_IDS.update({n: 76 + i for i, n in enumerate(["n", "px", "py", "pz"])})
EM_ARRAY_IDS.update({n: 82 + i for i, n in enumerate(["n2", "px2", "py2", "pz2"])})
EM_ARRAY_IDS.update({n: 96 + i for i, n in enumerate(
    ["Jx", "Jy", "Jz", "ux", "uz", "ux2", "uz2", "sx", "sy", "sz", "sx2", "sy2", "sz2", "energy_density"])})
MODEL_YM, MODEL_EMHD = 0, 1


class KhsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"khs error {code}: {msg}")
        self.code = code


class khs_config(ctypes.Structure):
    """Mirror of struct khs_config (include/khs.h)."""
    _fields_ = [
        ("k_mode", ctypes.c_int), ("alpha_YM", ctypes.c_float), ("perturb_amp", ctypes.c_float),
        ("run_mode", ctypes.c_int), ("V0", ctypes.c_float), ("xi_sponge", ctypes.c_float),
        ("sigma_sponge", ctypes.c_float), ("xi_cut", ctypes.c_float), ("warm_T", ctypes.c_float),
        ("freeze_override", ctypes.c_int), ("suppress_kz0", ctypes.c_int), ("hyp_diff", ctypes.c_float),
        ("kz_suppress_max", ctypes.c_int), ("eps_override", ctypes.c_float), ("kz_suppress_hi", ctypes.c_int),
        ("vz_edge_taper", ctypes.c_float), ("init_by1_eq", ctypes.c_int), ("nz_override", ctypes.c_int),
        ("nx_override", ctypes.c_int), ("courant_override", ctypes.c_float), ("lz_override", ctypes.c_float),
        ("lx_override", ctypes.c_float), ("target_tu", ctypes.c_float),
        ("run_tag", ctypes.c_char * 128), ("seed_profile_file", ctypes.c_char * 512),
        ("n_steps", ctypes.c_int), ("noise_amp", ctypes.c_double), ("noise_seed", ctypes.c_uint64),
        ("kz_keep_max", ctypes.c_int), ("precision", ctypes.c_int),
        ("energy_stride_override", ctypes.c_int), ("export_stride_override", ctypes.c_int),
        ("fused_sweep", ctypes.c_int), ("autotune", ctypes.c_int),
        ("x_segment", ctypes.c_int), ("z_chunk", ctypes.c_int),
        ("model", ctypes.c_int), ("species", ctypes.c_int), ("periodic", ctypes.c_int),
        ("charge1", ctypes.c_double), ("charge2", ctypes.c_double), ("emhd_init", ctypes.c_int),
        ("smooth_init", ctypes.c_int), ("spectrum_kmax", ctypes.c_int), ("spectrum_amp", ctypes.c_double),
        ("spectrum_seed", ctypes.c_uint64), ("em_perturb_amp", ctypes.c_double),
        ("z_groups", ctypes.c_int), ("xz_variant", ctypes.c_int), ("x_serial", ctypes.c_int), ("maxwell_kernel", ctypes.c_int), ("x_rows", ctypes.c_int), ("z_prefetch", ctypes.c_int), ("mfma_filter", ctypes.c_int), ("x_lane_cells", ctypes.c_int), ("snake_order", ctypes.c_int), ("zm_fused", ctypes.c_int), ("step_graphs", ctypes.c_int),
        ("step_fusion", ctypes.c_int), ("mxz_variant", ctypes.c_int), ("packed_shape", ctypes.c_int),
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("periodic_x", ctypes.c_int), ("freeze_az1", ctypes.c_int),
        ("lx", ctypes.c_double), ("lz", ctypes.c_double), ("dx", ctypes.c_double), ("dz", ctypes.c_double),
        ("courant", ctypes.c_double), ("dt", ctypes.c_double),
        ("alpha", ctypes.c_double), ("v0", ctypes.c_double), ("eps", ctypes.c_double),
        ("c", ctypes.c_double), ("eps_0", ctypes.c_double),
        ("total_steps", ctypes.c_int), ("export_stride", ctypes.c_int), ("energy_stride", ctypes.c_int),
        ("energy_thresh", ctypes.c_double),
    ]


_lib = None


def load_library(path=None):
    """Load libkhs.so (raises ImportError with a build hint if absent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} not found: build it with `python kh-simulations_amd/build.py`")
    lib = ctypes.CDLL(p)
    for name, (res, args) in _signatures().items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.khs_config_sizeof() != ctypes.sizeof(khs_config):
        raise ImportError("khs_config layout mismatch between binding and library")
    if path is None:
        _lib = lib
    return lib


def _signatures():
    c_int, c_double, c_char_p, c_size_t, c_void_p = (ctypes.c_int, ctypes.c_double, ctypes.c_char_p,
                                                      ctypes.c_size_t, ctypes.c_void_p)
    P = ctypes.POINTER
    cfgp = P(khs_config)
    dp = P(c_double)
    return {
        "khs_last_error": (c_char_p, []), "khs_version": (c_char_p, []), "khs_config_sizeof": (c_size_t, []),
        "khs_config_defaults": (c_int, [cfgp]), "khs_config_set": (c_int, [cfgp, c_char_p, c_char_p]),
        "khs_load_ini": (c_int, [c_char_p, cfgp]), "khs_derive": (c_int, [cfgp]),
        "khs_output_dir": (c_int, [cfgp, c_char_p, c_size_t]),
        "khs_create": (c_int, [cfgp, P(c_void_p)]), "khs_destroy": (c_int, [c_void_p]),
        "khs_init": (c_int, [c_void_p, c_char_p]), "khs_array_id": (c_int, [c_char_p]),
        "khs_get_array": (c_int, [c_void_p, c_int, dp, c_size_t]),
        "khs_set_array": (c_int, [c_void_p, c_int, dp, c_size_t]),
        "khs_get_array_buf": (c_int, [c_void_p, c_int, c_int, dp, c_size_t]),
        "khs_fct_x_sweep": (c_int, [c_void_p, c_int]), "khs_fct_z_sweep": (c_int, [c_void_p, c_int]),
        "khs_maxwell": (c_int, [c_void_p]), "khs_filters": (c_int, [c_void_p]), "khs_commit": (c_int, [c_void_p]),
        "khs_step": (c_int, [c_void_p, c_int, P(c_int)]),
        "khs_get_step": (c_int, [c_void_p, P(c_int), dp]),
        "khs_energy": (c_int, [c_void_p, dp]), "khs_energy0": (c_int, [c_void_p, dp]),
        "khs_last_energy": (c_int, [c_void_p, dp, P(c_int)]),
        "khs_gauss": (c_int, [c_void_p, dp]),
        "khs_gauss_residual": (c_int, [c_void_p, dp, dp]),
        "khs_write_snapshot": (c_int, [c_void_p, c_char_p, c_int]),
        "khs_energy_csv_open": (c_int, [c_void_p, c_char_p]),
        "khs_energy_csv_row": (c_int, [c_void_p, c_int, c_double]),
        "khs_energy_csv_close": (c_int, [c_void_p]),
        "khs_sync": (c_int, [c_void_p]), "khs_timing_enable": (c_int, [c_void_p, c_int]),
        "khs_timing_reset": (c_int, [c_void_p]),
        "khs_kernel_times": (c_int, [c_void_p, dp, P(c_int), c_int]),
        "khs_kernel_name": (c_char_p, [c_int]),
        "khs_get_tuning": (c_int, [c_void_p, P(c_int)]),
        "khs_fluid_velocities": (c_int, [c_void_p, c_int, dp, dp, c_size_t]),
        "khs_sources": (c_int, [c_void_p, c_int, dp, dp, dp, c_size_t]),
        "khs_enforce_boundaries": (c_int, [c_void_p]),
        "khs_checkpoint_write": (c_int, [c_void_p, c_char_p]),
        "khs_checkpoint_read": (c_int, [c_void_p, c_char_p]),
        "khs_mode_amplitude": (c_int, [c_void_p, c_int, dp]),
        "khs_mode_spectrum": (c_int, [c_void_p, c_int, c_int, dp]),
        "khs_fit_growth_max_r2": (c_int, [dp, dp, c_int, c_double, c_double, c_double, dp]),
        "khs_fit_growth_plateau": (c_int, [dp, dp, c_int, c_double, c_double, c_double, dp]),
        "khs_batch_create": (c_int, [cfgp, c_int, c_int, c_int, P(c_void_p)]),
        "khs_batch_destroy": (c_int, [c_void_p]), "khs_batch_init": (c_int, [c_void_p]),
        "khs_batch_run": (c_int, [c_void_p, c_int, P(c_int), P(c_int)]),
        "khs_batch_run_eager": (c_int, [c_void_p, c_int, P(c_int), P(c_int)]),
        "khs_batch_run_packed": (c_int, [c_void_p, c_int, P(c_int), P(c_int)]),
        "khs_batch_get_array": (c_int, [c_void_p, c_int, c_int, dp, c_size_t]),
        "khs_batch_set_array": (c_int, [c_void_p, c_int, c_int, dp, c_size_t]),
        "khs_batch_energy": (c_int, [c_void_p, c_int, dp, dp]),
        "khs_batch_mode_amplitude": (c_int, [c_void_p, c_int, c_int, dp]),
        "khs_batch_set_step_limit": (c_int, [c_void_p, c_int, c_int]),
        "khs_batch_run_concurrent": (c_int, [c_void_p, c_int, c_int, P(c_int), P(c_int)]),
        "khs_energy_noalloc": (c_int, [P(c_void_p), c_int, c_int, c_int, c_void_p, c_void_p]),
    }


def exported_symbols():
    """Names this binding binds (all entry points of include/khs.h)."""
    return list(_signatures())


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fit_growth_max_r2(t, amp, t_min=None, t_max=None, min_win_frac=0.15):
    """Max-R^2 growth-rate fit (analysis/batch_analyze.py:44-69) -> (gamma, t0, t1, r2)."""
    load_library()
    t = np.ascontiguousarray(t, dtype=np.float64)
    a = np.ascontiguousarray(amp, dtype=np.float64)
    out = np.zeros(4)
    nan = float("nan")
    _check(_lib.khs_fit_growth_max_r2(_dptr(t), _dptr(a), len(t), nan if t_min is None else t_min,
                                      nan if t_max is None else t_max, min_win_frac, _dptr(out)))
    return tuple(float(v) for v in out)


def fit_growth_plateau(t, amp, win_tu=8.0, min_plateau_tu=10.0, flat_tol=0.06):
    """Plateau growth-rate fit (analysis/batch_analyze.py:90-152)."""
    load_library()
    t = np.ascontiguousarray(t, dtype=np.float64)
    a = np.ascontiguousarray(amp, dtype=np.float64)
    out = np.zeros(5)
    _check(_lib.khs_fit_growth_plateau(_dptr(t), _dptr(a), len(t), win_tu, min_plateau_tu, flat_tol, _dptr(out)))
    return {"gamma": float(out[0]), "t0": float(out[1]), "t1": float(out[2]), "duration": float(out[3]),
            "has_plateau": bool(out[4])}


def _check(rc):
    if rc != KHS_OK:
        raise KhsError(rc, _lib.khs_last_error().decode(errors="replace"))


class Config:
    """A run configuration: reference .ini keys plus extension keys."""

    def __init__(self, **keys):
        load_library()
        self._c = khs_config()
        _check(_lib.khs_config_defaults(ctypes.byref(self._c)))
        for k, v in keys.items():
            self.set(k, v)
        if keys:
            self.derive()

    @classmethod
    def from_ini(cls, path):
        self = cls()
        _check(_lib.khs_load_ini(os.fsencode(path), ctypes.byref(self._c)))
        return self

    def set(self, key, value):
        _check(_lib.khs_config_set(ctypes.byref(self._c), key.encode(), str(value).encode()))
        return self

    def derive(self):
        _check(_lib.khs_derive(ctypes.byref(self._c)))
        return self

    def output_dir(self):
        buf = ctypes.create_string_buffer(1024)
        _check(_lib.khs_output_dir(ctypes.byref(self._c), buf, 1024))
        return buf.value.decode()

    @property
    def struct(self):
        return self._c

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        v = getattr(self._c, name)
        return v.decode() if isinstance(v, bytes) else v


class Simulation:
    """One device context (single GPU, single host thread)."""

    def __init__(self, config):
        load_library()
        self.config = config if isinstance(config, Config) else Config(**config)
        h = ctypes.c_void_p()
        _check(_lib.khs_create(ctypes.byref(self.config.struct), ctypes.byref(h)))
        self._h = h
        self.nx, self.nz = self.config.nx, self.config.nz
        self.model = self.config.model

    def close(self):
        if getattr(self, "_h", None):
            _lib.khs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state ----
    def init(self, seed_path=None):
        _check(_lib.khs_init(self._h, os.fsencode(seed_path) if seed_path else None))

    def array_id(self, name):
        if isinstance(name, int):
            return name
        ids = EM_ARRAY_IDS if self.model == MODEL_EMHD else ARRAY_IDS
        if name not in ids:
            raise KeyError(name)
        return ids[name]

    def state_names(self):
        if self.model == MODEL_EMHD:
            return EM_STATE_NAMES[: 13 if self.config.species == 1 else 17]
        return STATE_NAMES

    def get(self, name):
        out = np.empty((self.nz, self.nx), dtype=np.float64)
        _check(_lib.khs_get_array(self._h, self.array_id(name),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
        return out

    def get_buf(self, buf, name):
        """Read a state array from BUF_CURRENT / BUF_STAGE / BUF_NEXT."""
        out = np.empty((self.nz, self.nx), dtype=np.float64)
        _check(_lib.khs_get_array_buf(self._h, int(buf), self.array_id(name),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.nx)
        _check(_lib.khs_set_array(self._h, self.array_id(name),
                                  a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size))

    def state(self):
        return {n: self.get(n) for n in self.state_names()}

    # ---- operators ----
    def fct_x_sweep(self, species):
        _check(_lib.khs_fct_x_sweep(self._h, species))

    def fct_z_sweep(self, species):
        _check(_lib.khs_fct_z_sweep(self._h, species))

    def fluid_velocities(self, species):
        """(ux, uz) of species s (kernel_update_velocities) -> two (nz, nx) arrays."""
        c = self.config
        ux, uz = np.empty((c.nz, c.nx)), np.empty((c.nz, c.nx))
        _check(_lib.khs_fluid_velocities(self._h, int(species), _dptr(ux), _dptr(uz), ux.size))
        return ux, uz

    def sources(self, species):
        """(src_px, src_pz, sQ) of species s; sQ has 3 (YM) or 1 (EMHD: sy) rows."""
        c = self.config
        px, pz = np.empty((c.nz, c.nx)), np.empty((c.nz, c.nx))
        nq = 1 if c.model == MODEL_EMHD else 3
        sq = np.empty((nq, c.nz, c.nx))
        _check(_lib.khs_sources(self._h, int(species), _dptr(px), _dptr(pz), _dptr(sq), px.size))
        return px, pz, sq

    def enforce_boundaries(self):
        _check(_lib.khs_enforce_boundaries(self._h))

    def maxwell(self):
        _check(_lib.khs_maxwell(self._h))

    def filters(self):
        _check(_lib.khs_filters(self._h))

    def commit(self):
        _check(_lib.khs_commit(self._h))

    def step(self, n=1):
        halted = ctypes.c_int(0)
        _check(_lib.khs_step(self._h, int(n), ctypes.byref(halted)))
        return bool(halted.value)

    @property
    def step_count(self):
        s = ctypes.c_int()
        _check(_lib.khs_get_step(self._h, ctypes.byref(s), None))
        return s.value

    # ---- diagnostics ----
    def energy(self):
        e = ctypes.c_double()
        _check(_lib.khs_energy(self._h, ctypes.byref(e)))
        return e.value

    def last_energy_check(self):
        """(E, step) of the last energy check of khs_step (step -1: none yet)."""
        e, st = ctypes.c_double(), ctypes.c_int()
        _check(_lib.khs_last_energy(self._h, ctypes.byref(e), ctypes.byref(st)))
        return e.value, st.value

    @property
    def energy0(self):
        e = ctypes.c_double()
        _check(_lib.khs_energy0(self._h, ctypes.byref(e)))
        return e.value

    def gauss(self):
        out = (ctypes.c_double * 7)()
        _check(_lib.khs_gauss(self._h, out))
        v = list(out)
        return {"rms_ref": v[0:3], "rms_cont": v[3:6], "Etot_scale": v[6]}

    def gauss_residual(self):
        """(rms[3], rel[3]) as analysis/gauss_check.py reports them (khs_gauss_residual)."""
        rms, rel = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        _check(_lib.khs_gauss_residual(self._h, rms, rel))
        return list(rms), list(rel)

    def mode_amplitude(self, kz):
        """Circular amplitude of Az2 + i Az3 at z mode kz -> (amp, amp_conj)."""
        out = (ctypes.c_double * 2)()
        _check(_lib.khs_mode_amplitude(self._h, int(kz), out))
        return out[0], out[1]

    def mode_spectrum(self, name, kmax):
        """Complex 2-D Fourier coefficients [kz-1, kx-1] (kx, kz = 1..kmax) of a state array."""
        out = np.zeros(2 * kmax * kmax)
        _check(_lib.khs_mode_spectrum(self._h, self.array_id(name), int(kmax),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return (out[0::2] + 1j * out[1::2]).reshape(kmax, kmax)

    def timeseries(self, kz, n_steps, sample_stride):
        """Step n_steps, sampling the mode-kz amplitude at t = 0 and every
        sample_stride steps (stops early if the run halts)."""
        ts, amps, conj = [], [], []
        dt = self.config.dt
        a, c = self.mode_amplitude(kz)
        ts.append(self.step_count * dt); amps.append(a); conj.append(c)
        done, halted = 0, False
        while done < n_steps and not halted:
            k = min(sample_stride, n_steps - done)
            halted = self.step(k)
            done += k
            a, c = self.mode_amplitude(kz)
            ts.append(self.step_count * dt); amps.append(a); conj.append(c)
        return {"t": np.array(ts), "amp": np.array(amps), "amp_conj": np.array(conj), "halted": halted}

    # ---- files ----
    def checkpoint(self, path):
        """Binary FP64 checkpoint of the committed state (khs_checkpoint_write)."""
        _check(_lib.khs_checkpoint_write(self._h, os.fsencode(path)))

    def restore(self, path):
        """Load a checkpoint of the same configuration; step and E0 follow it."""
        _check(_lib.khs_checkpoint_read(self._h, os.fsencode(path)))

    def write_snapshot(self, directory, step):
        _check(_lib.khs_write_snapshot(self._h, os.fsencode(directory), int(step)))

    def energy_csv_open(self, path):
        _check(_lib.khs_energy_csv_open(self._h, os.fsencode(path)))

    def energy_csv_close(self):
        _check(_lib.khs_energy_csv_close(self._h))

    # ---- timing ----
    def sync(self):
        _check(_lib.khs_sync(self._h))

    def timing(self, on=True):
        _check(_lib.khs_timing_enable(self._h, 1 if on else 0))

    def timing_reset(self):
        _check(_lib.khs_timing_reset(self._h))

    def tuning(self):
        """Tile shapes and kernel variants in use (khs_get_tuning)."""
        out = (ctypes.c_int * 13)()
        _check(_lib.khs_get_tuning(self._h, out))
        return dict(x_segment=out[0], x_rows=out[1], z_chunk=out[2], z_prefetch=out[3], maxwell=out[4],
                    x_serial=out[5], xz_variant=out[6], z_groups=out[7], x_lane_cells=out[8], zm_fused=out[9],
                    fused_sweep=out[10], step_fusion=out[11], mxz_variant=out[12])

    def kernel_times(self):
        n = 16
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        _check(_lib.khs_kernel_times(self._h, ms, cnt, n))
        out = {}
        for i in range(n):
            nm = _lib.khs_kernel_name(i)
            if nm is None:
                break
            out[nm.decode()] = (ms[i], cnt[i])
        return out


class Batch:
    """Independent runs stepped together through captured HIP graphs
    (khs_batch_*).  configs: list of Config."""

    def __init__(self, configs, lanes=4, segment=50):
        load_library()
        self.configs = [c if isinstance(c, Config) else Config(**c) for c in configs]
        n = len(self.configs)
        arr = (khs_config * n)(*[c.struct for c in self.configs])
        h = ctypes.c_void_p()
        _check(_lib.khs_batch_create(arr, n, int(lanes), int(segment), ctypes.byref(h)))
        self._h = h
        self.n = n

    def close(self):
        if getattr(self, "_h", None):
            _lib.khs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        _check(_lib.khs_batch_init(self._h))

    def _run(self, fn, nsteps):
        done = (ctypes.c_int * self.n)()
        halted = (ctypes.c_int * self.n)()
        _check(fn(self._h, int(nsteps), done, halted))
        return list(done), [bool(h) for h in halted]

    def run(self, nsteps):
        """Graph-captured stepping -> (steps per run, halted flags)."""
        return self._run(_lib.khs_batch_run, nsteps)

    def run_eager(self, nsteps):
        return self._run(_lib.khs_batch_run_eager, nsteps)

    def run_packed(self, nsteps):
        """Graph-captured segments with every step of all runs in three shared
        launches (khs_batch_run_packed)."""
        return self._run(_lib.khs_batch_run_packed, nsteps)

    def run_concurrent(self, nsteps, threads=16):
        """Runs stepped by `threads` host threads on their own streams (16
        keeps more small runs' kernels in flight than 8 at HIP's default 4
        hardware queues)."""
        return self._run(lambda h, n, d, hl: _lib.khs_batch_run_concurrent(h, n, int(threads), d, hl), nsteps)

    def get(self, run, name):
        if not 0 <= run < self.n:
            raise KhsError(-1, f"run index {run} out of range")
        c = self.configs[run]
        ids = EM_ARRAY_IDS if c.model == MODEL_EMHD else ARRAY_IDS
        out = np.empty((c.nz, c.nx), dtype=np.float64)
        _check(_lib.khs_batch_get_array(self._h, int(run), ids[name] if isinstance(name, str) else name,
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
        return out

    def set(self, run, name, arr):
        c = self.configs[run]
        ids = EM_ARRAY_IDS if c.model == MODEL_EMHD else ARRAY_IDS
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(c.nz, c.nx)
        _check(_lib.khs_batch_set_array(self._h, int(run), ids[name] if isinstance(name, str) else name,
                                        a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size))

    def energy(self, run):
        e, e0 = ctypes.c_double(), ctypes.c_double()
        _check(_lib.khs_batch_energy(self._h, int(run), ctypes.byref(e), ctypes.byref(e0)))
        return e.value, e0.value

    def set_step_limit(self, run, max_step):
        """run stops at max_step (<= 0: no limit) without being marked halted."""
        _check(_lib.khs_batch_set_step_limit(self._h, int(run), int(max_step)))

    def mode_amplitude(self, run, kz):
        """Circular Az2 + i Az3 amplitude of run `run` at z mode kz -> (amp, amp_conj)."""
        out = (ctypes.c_double * 2)()
        _check(_lib.khs_batch_mode_amplitude(self._h, int(run), int(kz), out))
        return out[0], out[1]
