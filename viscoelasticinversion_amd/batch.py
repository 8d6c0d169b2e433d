"""Device-resident models and batches of shots (the C ABI's vi_model / vi_batch).

A ``Model`` is uploaded once and shared by many shots, each reading a column window of it.
A ``Batch`` advances a set of ``Simulation`` objects with one kernel launch per half step
covering all their tiles; optionally it replays a captured hipGraph of a block of steps.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


class Model:
    """Device-resident model arrays of shape (nz, nx): rho, lam, mu[, tau] (isotropic) or,
    with ``vti`` = (c11, c13, c33, c55), rho and the VTI stiffnesses [, tau]; lam and mu
    are then ignored (pass None)."""

    def __init__(self, rho, lam, mu, tau=None, precision="float32", vti=None):
        self.dtype = np.dtype(np.float32 if precision in ("float32", np.float32, "f32") else np.float64)
        self.vti = vti is not None
        planes = (rho, *vti) if self.vti else (rho, lam, mu)
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=self.dtype)) for a in planes]
        self.nz, self.nx = arrs[0].shape
        for a in arrs[1:]:
            if a.shape != arrs[0].shape:
                raise ValueError("model arrays must share one shape")
        t = None if tau is None else np.ascontiguousarray(np.asarray(tau, dtype=self.dtype))
        if t is not None and t.shape != arrs[0].shape:
            raise ValueError("tau must have the model's shape")
        h = ctypes.c_void_p()
        P = ctypes.c_void_p
        prec = L.VI_F32 if self.dtype == np.float32 else L.VI_F64
        tp = t.ctypes.data_as(P) if t is not None else None
        if self.vti:
            L.check(L.lib().vi_model_create_vti(self.nx, self.nz, prec,
                                                *(a.ctypes.data_as(P) for a in arrs), tp,
                                                ctypes.byref(h)))
        else:
            L.check(L.lib().vi_model_create(self.nx, self.nz, prec,
                                            *(a.ctypes.data_as(P) for a in arrs), tp,
                                            ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_dict(cls, m, precision="float32"):
        if "c11" in m:
            return cls(m["rho"], None, None, m.get("tau"), precision,
                       vti=(m["c11"], m["c13"], m["c33"], m["c55"]))
        return cls(m["rho"], m["lam"], m["mu"], m.get("tau"), precision)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().vi_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Advance several simulations together (same results as stepping each alone)."""

    def __init__(self, sims=()):
        h = ctypes.c_void_p()
        L.check(L.lib().vi_batch_create(ctypes.byref(h)))
        self._h = h
        self.sims = []
        if sims:
            self.set(sims)

    def set(self, sims):
        self.sims = list(sims)
        arr = (ctypes.c_void_p * max(len(self.sims), 1))(*[s.handle for s in self.sims])
        L.check(L.lib().vi_batch_set(self._h, arr, len(self.sims)))

    def set_tiling(self, rows):
        L.check(L.lib().vi_batch_set_tiling(self._h, int(rows)))

    @property
    def tiling(self):
        return int(L.lib().vi_batch_get_tiling(self._h))

    def autotune(self, trial_steps=10):
        chosen = ctypes.c_int()
        us = (ctypes.c_double * 4)()
        L.check(L.lib().vi_batch_autotune(self._h, int(trial_steps), ctypes.byref(chosen), us))
        return chosen.value, dict(zip((1, 2, 4, 8), list(us)))

    def set_graph(self, steps_per_graph):
        L.check(L.lib().vi_batch_set_graph(self._h, int(steps_per_graph)))

    def set_cache_block(self, nbytes):
        """Advance groups of shots with at most ``nbytes`` of working set one after another
        through each block of steps (0: every shot in every launch); results are unchanged."""
        L.check(L.lib().vi_batch_set_cache_block(self._h, int(nbytes)))

    @property
    def cache_block(self):
        return int(L.lib().vi_batch_get_cache_block(self._h))

    def set_streams(self, n):
        """n = 1..4 (default 2): n launch groups at once on n streams, each at most 1/n of
        the cache block; 1: one at a time."""
        L.check(L.lib().vi_batch_set_streams(self._h, int(n)))

    def step(self, nsteps=1):
        L.check(L.lib().vi_batch_step(self._h, int(nsteps)))

    def sync(self):
        L.check(L.lib().vi_batch_sync(self._h))

    def profile(self, every=1):
        """Time the launches of every ``every``-th step with dispatch events (0 = off)."""
        L.check(L.lib().vi_batch_profile_enable(self._h, int(every)))

    def profile_read(self, kernel):
        ms, n = ctypes.c_double(), ctypes.c_long()
        L.check(L.lib().vi_batch_profile_read(self._h, int(kernel), ctypes.byref(ms),
                                              ctypes.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            L.lib().vi_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mem_info():
    """(free, total) device bytes; free includes the large buffers the library keeps for
    reuse (freed when an allocation needs the memory)."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    L.check(L.lib().vi_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def cached_bytes():
    """Device bytes the library keeps for reuse: large buffers released by simulations,
    handed to the next allocation they fit instead of back to the runtime."""
    b = ctypes.c_size_t()
    L.check(L.lib().vi_cached_bytes(ctypes.byref(b)))
    return b.value


def release_cached():
    """Free the buffers the library keeps for reuse (for other users of the device)."""
    L.check(L.lib().vi_release_cached())
