"""khs_energy_noalloc: the energy reduction on caller-owned device arrays with
the library's static scratch (no device allocation) gives the bits of
khs_energy on the same state, on two calls in a row, the second on a smaller
grid; bad arguments are error codes."""
import ctypes

import numpy as np
import pytest

import khs
from helpers import make_sim, random_state

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libkhs itself is linked against (already loaded by
    khs.load_library), for the test's own device buffers."""
    khs.load_library()
    with open("/proc/self/maps") as f:
        path = next(l.split()[-1] for l in f if "libamdhip64.so" in l)
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc_energy(sim, nx, nz, precision):
    """Copy the context's state into device buffers of the test, then call
    khs_energy_noalloc on them (null stream) and read the sum back."""
    hip = _hip()
    dtype = np.float64 if precision == 0 else np.float32
    names = sim.state_names()                  # the order of the khs_array ids
    assert len(names) == 24
    bufs = []
    try:
        for n in names + ["_out"]:
            p = ctypes.c_void_p()
            nbytes = 16 if n == "_out" else nx * nz * np.dtype(dtype).itemsize
            assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
            bufs.append(p)
            if n != "_out":
                host = np.ascontiguousarray(sim.get(n), dtype=dtype)
                assert hip.hipMemcpy(p, host.ctypes.data, nbytes, 1) == 0          # host to device
        lib = khs.load_library()
        ptrs = (ctypes.c_void_p * 24)(*[b.value for b in bufs[:24]])
        rc = lib.khs_energy_noalloc(ptrs, nx, nz, precision, None, bufs[24])
        assert rc == khs.KHS_OK, lib.khs_last_error()
        out = np.full(2, np.nan)
        assert hip.hipMemcpy(out.ctypes.data, bufs[24], 16, 2) == 0                # device to host, waits
        return float(out[0])
    finally:
        for p in bufs:
            hip.hipFree(p)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_energy_matches_khs_energy_twice(precision):
    code = 0 if precision == "fp64" else 1
    for nx, nz, seed in ((300, 160, 1), (40, 24, 2)):            # the second call on a smaller grid
        keys = dict(nx_override=nx, nz_override=nz, lx_override=nx * 0.1, lz_override=nz * 0.1, run_mode=6,
                    precision=precision)
        cfg, sim = make_sim(**keys)
        for n, a in random_state(nx, nz, seed).items():
            sim.set(n, a)
        want = sim.energy()
        got = _noalloc_energy(sim, nx, nz, code)
        print(f"{precision} {nx}x{nz}: khs_energy {want!r} khs_energy_noalloc {got!r}", flush=True)
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64)
        sim.close()


def test_noalloc_energy_rejects_bad_arguments():
    lib = khs.load_library()
    ptrs = (ctypes.c_void_p * 24)()                               # all null
    out = ctypes.c_double()
    assert lib.khs_energy_noalloc(ptrs, 8, 8, 0, None, ctypes.byref(out)) == khs.KHS_ERR_ARG
    assert b"null" in lib.khs_last_error()
    assert lib.khs_energy_noalloc(ptrs, 0, 8, 0, None, ctypes.byref(out)) == khs.KHS_ERR_ARG
    assert lib.khs_energy_noalloc(ptrs, 1 << 20, 1 << 10, 0, None, ctypes.byref(out)) == khs.KHS_ERR_ARG
    assert b"limit" in lib.khs_last_error()
