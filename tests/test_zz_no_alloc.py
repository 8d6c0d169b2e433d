"""vi_rates_noalloc (the right-hand side without any device allocation, workspace in a static
per-GPU device array) against vi_rates through the C ABI: bit-identical on two calls in a row,
the second on a smaller grid, then the first again; no handle memory grows; handles whose
workspace exceeds the static array are refused."""
import ctypes

import numpy as np
import pytest

from helpers import small_cfg

pytestmark = pytest.mark.gpu


def _rates(fn, sim, sv):
    dv = np.full((2, sim.nz, sim.nx), np.nan, dtype=sim.dtype)
    ds = np.full((3, sim.nz, sim.nx), np.nan, dtype=sim.dtype)
    dm = np.full((3, sim.nrelax, sim.nz, sim.nx), np.nan, dtype=sim.dtype) if sim.nrelax else None
    svp = np.ascontiguousarray(sv, dtype=sim.dtype)
    from viscoelasticinversion_amd import _lib
    _lib.check(fn(sim.handle, svp.ctypes.data_as(ctypes.c_void_p),
                  dv.ctypes.data_as(ctypes.c_void_p), ds.ctypes.data_as(ctypes.c_void_p),
                  dm.ctypes.data_as(ctypes.c_void_p) if dm is not None else None))
    return dv, ds, dm


def _sim(cfg, prec, seed):
    from viscoelasticinversion_amd.configs import build_simulation
    sim = build_simulation(cfg, prec)
    rng = np.random.default_rng(seed)
    for k in range(5):
        sim.set(k, rng.standard_normal((cfg["nz"], cfg["nx"])))
    if cfg["nrelax"]:
        for k in (5, 6, 7):
            sim.set(k, rng.standard_normal((cfg["nrelax"], cfg["nz"], cfg["nx"])))
    if cfg["cpml"]:
        for k in range(8, 16):
            sim.set(k, rng.standard_normal(sim.get(k).shape))
    return sim


def test_noalloc_rates_equal_rates():
    from viscoelasticinversion_amd import _lib
    lib = _lib.lib()
    big_cfg = small_cfg(nx=300, nz=210, nrelax=3, cpml=True, fs=True, seed=21, nt=8)
    small = small_cfg(nx=90, nz=61, nrelax=1, cpml=True, fs=True, seed=22, nt=8)
    small["free_surface"] = "afda"
    big = _sim(big_cfg, "float32", 1)
    little = _sim(small, "float64", 2)
    print("[case] vi_rates_noalloc: 300x210 L=3 FP32, then 90x61 L=1 FP64 H-AFDA, then again",
          flush=True)
    for sim, sv in ((big, np.linspace(-1, 1, 5)), (little, np.linspace(2, 3, 5)),
                    (big, np.linspace(-3, 1, 5))):
        before = sim.device_bytes()
        got = _rates(lib.vi_rates_noalloc, sim, sv)
        assert sim.device_bytes() == before  # nothing allocated for the handle
        want = _rates(lib.vi_rates, sim, sv)
        for g, w in zip(got, want):
            if w is not None:
                assert np.array_equal(g, w)
        assert np.abs(got[0]).max() > 0 and np.abs(got[1]).max() > 0
    big.close()
    little.close()


def test_noalloc_rates_refuse_a_handle_beyond_the_workspace():
    from viscoelasticinversion_amd import _lib
    from viscoelasticinversion_amd.stepper import Simulation
    # FP64, L = 6: 23 planes of 1024 x 948 padded cells = 179 MB > 128 MiB
    sim = Simulation(900, 900, 5.0, 5.0, 4e-4, "float64", 6)
    sim.set_model(2000.0, 2e9, 1e9, 0.02)
    sim.set_relaxation(np.geomspace(1e-3, 1e-1, 6))
    dv = np.zeros((2, 900, 900))
    ds = np.zeros((3, 900, 900))
    rc = _lib.lib().vi_rates_noalloc(sim.handle, None, dv.ctypes.data_as(ctypes.c_void_p),
                                     ds.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == _lib.VI_EINVAL
    assert b"static" in _lib.lib().vi_last_error()
    sim.close()
