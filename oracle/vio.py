"""ctypes wrapper of the CPU oracle ``libvio.so`` (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import
this module, as the checker of the HIP product path. The product package never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvio.so")
_lib = None

VX, VZ, SXX, SZZ, SXZ, MXX, MZZ, MXZ = range(8)
PSI_VX_X, PSI_VZ_X, PSI_SXX_X, PSI_SXZ_X, PSI_VX_Z, PSI_VZ_Z, PSI_SXZ_Z, PSI_SZZ_Z = range(8, 16)
FUSED, MULTIPASS = 0, 1
FS_NONE, FS_IMAGE, FS_AFDA, FS_HERMITIAN = 0, 1, 2, 3


def fs_mode(free_surface) -> int:
    """False/None/"none" -> 0, True/"image" -> 1 (stress imaging), "afda" -> 2 (H-AFDA),
    "hermitian" -> 3 (H-AFDA with formula #3 for d vz/dz on row 5)."""
    if isinstance(free_surface, str):
        return {"none": FS_NONE, "image": FS_IMAGE, "afda": FS_AFDA,
                "hermitian": FS_HERMITIAN}[free_surface]
    return int(free_surface or 0)


def afda_weights(h, counts=None):
    """The oracle's H-AFDA one-sided weights / h, shape (13, 8): table rows A0..A3, B0..B2,
    C1..C3, D0..D2 over the field rows 4..11 (vio.c). ``counts`` = points per table row
    (None: the default counts)."""
    out = np.zeros((13, 8))
    cp = None
    if counts is not None:
        cp = (ctypes.c_int * 13)(*[int(c) for c in counts])
    if lib().vio_afda_weights(float(h), cp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
        raise ValueError(f"afda point counts out of range: {counts}")
    return out


def afda_hermitian(h, n=5):
    """Formula #3 weights (vio.c): [alpha, beta_0 / h, ..., beta_7 / h] for d vz/dz at z = h
    from d vz/dz(0) and vz at (m+1/2) h, m < n."""
    out = np.zeros(9)
    if lib().vio_afda_hermitian(float(h), int(n), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
        raise ValueError(f"hermitian point count out of range: {n}")
    return out


def afda_default_points():
    out = (ctypes.c_int * 13)()
    lib().vio_afda_default_points(out)
    return list(out)


class Desc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nrelax", ctypes.c_int),
        ("dx", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("free_surface", ctypes.c_int), ("cpml_thickness", ctypes.c_int),
        ("cpml_d0", ctypes.c_double), ("cpml_alpha_max", ctypes.c_double),
        ("cpml_kappa_max", ctypes.c_double), ("cpml_power", ctypes.c_double),
        ("cpml_left", ctypes.c_int), ("cpml_right", ctypes.c_int),
        ("cpml_top", ctypes.c_int), ("cpml_bottom", ctypes.c_int),
        ("afda_points", ctypes.c_int * 13),
        ("afda_hermitian_points", ctypes.c_int),
        ("fd_order_x", ctypes.c_int), ("fd_order_z", ctypes.c_int),
    ]


def build() -> None:
    """Compile libvio.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        vp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
        _lib.vio_afda_weights.restype = ctypes.c_int
        _lib.vio_afda_weights.argtypes = [ctypes.c_double, ip, ctypes.POINTER(ctypes.c_double)]
        _lib.vio_afda_hermitian.restype = ctypes.c_int
        _lib.vio_afda_hermitian.argtypes = [ctypes.c_double, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double)]
        _lib.vio_afda_default_points.restype = None
        _lib.vio_afda_default_points.argtypes = [ip]
        for suf, ct in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
            rp = ctypes.POINTER(ct)
            f = getattr(_lib, "vio_create" + suf)
            f.restype = vp
            f.argtypes = [ctypes.POINTER(Desc), rp, rp, rp, rp, rp]
            f = getattr(_lib, "vio_create_vti" + suf)
            f.restype = vp
            f.argtypes = [ctypes.POINTER(Desc), rp, rp, rp, rp, rp, rp, rp]
            f = getattr(_lib, "vio_set_sources" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [vp, ctypes.c_int, ip, ip, ip, ctypes.c_int, rp]
            f = getattr(_lib, "vio_set_receivers" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [vp, ctypes.c_int, ip, ip, ip]
            f = getattr(_lib, "vio_run" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, rp]
            for name in ("vio_get", "vio_set"):
                f = getattr(_lib, name + suf)
                f.restype = ctypes.c_int
                f.argtypes = [vp, ctypes.c_int, rp]
            f = getattr(_lib, "vio_step_index" + suf)
            f.restype = ctypes.c_long
            f.argtypes = [vp]
            f = getattr(_lib, "vio_rates" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [vp, rp, rp, rp, rp]
            f = getattr(_lib, "vio_destroy" + suf)
            f.restype = None
            f.argtypes = [vp]
        _lib.vio_fd_weights.restype = ctypes.c_int
        _lib.vio_fd_weights.argtypes = [ctypes.c_double, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
        _lib.vio_fd_coeffs.restype = None
        _lib.vio_fd_coeffs.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        _lib.vio_cpml_axis.restype = None
        _lib.vio_cpml_axis.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 5 + [
            ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp]
        _lib.vio_cpml_axis_f32.restype = None
        _lib.vio_cpml_axis_f32.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 5 + [
            ctypes.c_int, ctypes.c_int, fp, fp, fp, fp, fp]
    return _lib


def fd_weights(h: float, order: int) -> np.ndarray:
    """Staggered weights / h of the given order (2, 4, 6, 8), padded to 4 with zeros."""
    out = np.zeros(4)
    if lib().vio_fd_weights(float(h), int(order), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
        raise ValueError(f"unsupported order {order}")
    return out


def fd_coeffs(h: float) -> np.ndarray:
    out = np.zeros(4)
    lib().vio_fd_coeffs(h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def cpml_axis(n, thickness, d0, alpha_max, kappa_max, power, dt, lo, hi, dtype=np.float64):
    """Return dict of a, b, kappa, d, alpha (arrays of length n), computed in ``dtype``
    (float64 or float32) as the reference computes its CPML coefficients."""
    dt_ = np.dtype(dtype)
    arrs = {k: np.zeros(n, dtype=dt_) for k in ("a", "b", "kappa", "d", "alpha")}
    ct = ctypes.c_float if dt_ == np.float32 else ctypes.c_double
    fn = lib().vio_cpml_axis_f32 if dt_ == np.float32 else lib().vio_cpml_axis
    P = ctypes.POINTER(ct)
    fn(n, thickness, d0, alpha_max, kappa_max, power, dt, int(lo), int(hi),
       *(arrs[k].ctypes.data_as(P) for k in ("a", "b", "kappa", "d", "alpha")))
    return arrs


class Oracle:
    """One oracle simulation. ``dtype`` is np.float32 or np.float64.

    ``model`` holds rho, lam, mu (and tau when nrelax > 0) as [nz][nx] arrays, or for a VTI
    medium rho, c11, c13, c33, c55 (and tau); ``taun`` the
    L relaxation times; ``cpml`` a dict with thickness, d0, alpha_max, kappa_max, power,
    left, right, top, bottom (or None).
    """

    def __init__(self, nx, nz, dx, dz, dt, model, dtype=np.float64, nrelax=0, taun=None,
                 free_surface=False, cpml=None, afda_points=None, hermitian_points=None,
                 fd_order=(8, 8)):
        self.dtype = np.dtype(dtype)
        self.suf = "_f32" if self.dtype == np.float32 else "_f64"
        self.ct = ctypes.c_float if self.suf == "_f32" else ctypes.c_double
        self.nx, self.nz, self.L = nx, nz, nrelax
        d = Desc()
        d.nx, d.nz, d.nrelax = nx, nz, nrelax
        d.dx, d.dz, d.dt = dx, dz, dt
        d.free_surface = fs_mode(free_surface)
        if afda_points is not None:
            for k, c in enumerate(afda_points):
                d.afda_points[k] = int(c)
        if hermitian_points is not None:
            d.afda_hermitian_points = int(hermitian_points)
        d.fd_order_x, d.fd_order_z = (int(o) for o in fd_order)
        if cpml:
            d.cpml_thickness = int(cpml["thickness"])
            d.cpml_d0 = float(cpml["d0"])
            d.cpml_alpha_max = float(cpml.get("alpha_max", 0.0))
            d.cpml_kappa_max = float(cpml.get("kappa_max", 1.0))
            d.cpml_power = float(cpml.get("power", 2.0))
            d.cpml_left = int(cpml.get("left", True))
            d.cpml_right = int(cpml.get("right", True))
            d.cpml_top = int(cpml.get("top", False))
            d.cpml_bottom = int(cpml.get("bottom", True))
        self.T = d.cpml_thickness
        self._keep = []

        def arr(a):
            a = np.ascontiguousarray(a, dtype=self.dtype)
            self._keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(self.ct))

        tau = model.get("tau", np.zeros((nz, nx)))
        tn = np.zeros(max(nrelax, 1)) if taun is None else np.asarray(taun)
        if "c11" in model:
            self.h = getattr(lib(), "vio_create_vti" + self.suf)(
                ctypes.byref(d), arr(model["rho"]), arr(model["c11"]), arr(model["c13"]),
                arr(model["c33"]), arr(model["c55"]), arr(tau), arr(tn))
        else:
            self.h = getattr(lib(), "vio_create" + self.suf)(
                ctypes.byref(d), arr(model["rho"]), arr(model["lam"]), arr(model["mu"]),
                arr(tau), arr(tn))
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        self._keep.clear()

    def _p(self, a):
        return a.ctypes.data_as(ctypes.POINTER(self.ct))

    def set_sources(self, ix, iz, kind, wavelets):
        ix = np.ascontiguousarray(ix, dtype=np.int32)
        iz = np.ascontiguousarray(iz, dtype=np.int32)
        kind = np.ascontiguousarray(kind, dtype=np.int32)
        w = np.ascontiguousarray(wavelets, dtype=self.dtype).reshape(len(ix), -1)
        IP = ctypes.POINTER(ctypes.c_int)
        rc = getattr(lib(), "vio_set_sources" + self.suf)(
            self.h, len(ix), ix.ctypes.data_as(IP), iz.ctypes.data_as(IP),
            kind.ctypes.data_as(IP), w.shape[1], self._p(w))
        if rc:
            raise ValueError("bad sources")

    def set_receivers(self, ix, iz, comp):
        ix = np.ascontiguousarray(ix, dtype=np.int32)
        iz = np.ascontiguousarray(iz, dtype=np.int32)
        comp = np.ascontiguousarray(comp, dtype=np.int32)
        IP = ctypes.POINTER(ctypes.c_int)
        rc = getattr(lib(), "vio_set_receivers" + self.suf)(
            self.h, len(ix), ix.ctypes.data_as(IP), iz.ctypes.data_as(IP),
            comp.ctypes.data_as(IP))
        if rc:
            raise ValueError("bad receivers")
        self.nrec = len(ix)

    def run(self, nsteps, mode=FUSED, nthreads=0):
        nrec = getattr(self, "nrec", 0)
        tr = np.zeros((nsteps, max(nrec, 1)), dtype=self.dtype)
        rc = getattr(lib(), "vio_run" + self.suf)(self.h, nsteps, mode, nthreads, self._p(tr))
        if rc:
            raise RuntimeError("oracle run failed")
        return tr[:, :nrec]

    def _shape(self, field):
        if field < MXX:
            return (self.nz, self.nx)
        if field < PSI_VX_X:
            return (self.L, self.nz, self.nx)
        if field < PSI_VX_Z:
            return (self.nz, 2 * self.T)
        return (2 * self.T, self.nx)

    def get(self, field):
        out = np.zeros(self._shape(field), dtype=self.dtype)
        if getattr(lib(), "vio_get" + self.suf)(self.h, field, self._p(out)):
            raise ValueError("field not present")
        return out

    def set(self, field, values):
        a = np.ascontiguousarray(values, dtype=self.dtype).reshape(self._shape(field))
        if getattr(lib(), "vio_set" + self.suf)(self.h, field, self._p(a)):
            raise ValueError("field not present")

    def rates(self, src_values=None):
        """Rates of the next half steps at the current state (the state is not changed):
        dv (2, nz, nx) = (vx, vz), ds (3, nz, nx) = (sxx, szz, sxz), dm (3, L, nz, nx) or
        None; ``src_values`` = one sample per source (None: no source term)."""
        dv = np.zeros((2, self.nz, self.nx), dtype=self.dtype)
        ds = np.zeros((3, self.nz, self.nx), dtype=self.dtype)
        dm = np.zeros((3, self.L, self.nz, self.nx), dtype=self.dtype) if self.L else None
        sv = None
        if src_values is not None:
            sv = np.ascontiguousarray(src_values, dtype=self.dtype).ravel()
        rc = getattr(lib(), "vio_rates" + self.suf)(
            self.h, self._p(sv) if sv is not None else None, self._p(dv), self._p(ds),
            self._p(dm) if dm is not None else None)
        if rc:
            raise RuntimeError("oracle rates failed")
        return dv, ds, dm

    @property
    def step_index(self):
        return getattr(lib(), "vio_step_index" + self.suf)(self.h)

    def close(self):
        if self.h:
            getattr(lib(), "vio_destroy" + self.suf)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
