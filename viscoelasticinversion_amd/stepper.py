"""Plain single-shot stepper: a thin object over one C-ABI handle (vi_create ... vi_destroy).

Arrays cross the boundary as numpy arrays of shape (nz, nx) -- x fastest, which is the
memory order of the reference's column-major (Nx, Nz) Julia arrays -- in the handle's
precision. Every call goes to the HIP library; there is no CPU execution path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

FIELDS = {"vx": L.VI_VX, "vz": L.VI_VZ, "sxx": L.VI_SXX, "szz": L.VI_SZZ, "sxz": L.VI_SXZ,
          "mxx": L.VI_MXX, "mzz": L.VI_MZZ, "mxz": L.VI_MXZ,
          "psi_vx_x": L.VI_PSI_VX_X, "psi_vz_x": L.VI_PSI_VZ_X, "psi_sxx_x": L.VI_PSI_SXX_X,
          "psi_sxz_x": L.VI_PSI_SXZ_X, "psi_vx_z": L.VI_PSI_VX_Z, "psi_vz_z": L.VI_PSI_VZ_Z,
          "psi_sxz_z": L.VI_PSI_SXZ_Z, "psi_szz_z": L.VI_PSI_SZZ_Z}
SOURCE_KINDS = {"force_z": L.VI_SRC_FORCE_Z, "explosive": L.VI_SRC_EXPLOSIVE,
                "force_x": L.VI_SRC_FORCE_X}
COMPONENTS = {"vx": 0, "vz": 1, "sxx": 2, "szz": 3, "sxz": 4}


FREE_SURFACE_MODES = {"none": 0, "image": 1, "imaging": 1, "afda": 2, "hermitian": 3}


def _as_int32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())


def free_surface_mode(mode) -> int:
    """False/None/"none" -> 0, True/"image" -> 1 (stress imaging), "afda" -> 2 (H-AFDA),
    "hermitian" -> 3 (H-AFDA with the Hermitian formula #3 for d vz/dz at z = h)."""
    if isinstance(mode, str):
        if mode not in FREE_SURFACE_MODES:
            raise ValueError(f"unknown free surface {mode!r}: none, image, afda or hermitian")
        return FREE_SURFACE_MODES[mode]
    return int(bool(mode))


def afda_weights(h, counts=None):
    """H-AFDA one-sided d/dz weights / h as the library computes them, shape (13, 8): rows
    A0..A3 (d sxz/dz at z = k h), B0..B2 (d szz/dz at (k+1/2) h), C1..C3 (d vz/dz at k h),
    D0..D2 (d vx/dz at (k+1/2) h); column m multiplies grid row 4 + m (FreeSurfaceStencil,
    Stencil.jl:52-105). Host-only; needs no GPU."""
    out = np.zeros((13, 8))
    cp = None
    if counts is not None:
        cp = (ctypes.c_int * 13)(*[int(c) for c in counts])
    L.check(L.lib().vi_afda_weights(float(h), cp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def afda_hermitian_weights(h, points=5):
    """Formula #3 as the library computes it (_stencil_hermitian, Stencil.jl:119-189):
    [alpha, beta_0 / h, ..., beta_7 / h] with d vz/dz(h) = alpha d vz/dz(0) +
    sum_m beta_m vz((m+1/2) h), m < points. Host-only; needs no GPU."""
    out = np.zeros(9)
    L.check(L.lib().vi_afda_hermitian_weights(float(h), int(points),
                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def fd_weights(h, order=8):
    """Staggered first-derivative weights / h of order 2, 4, 6 or 8 as the library computes
    them (Stencil.jl:280-320), padded to 4 with zeros. Host-only; needs no GPU."""
    out = np.zeros(4)
    L.check(L.lib().vi_fd_weights(float(h), int(order),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def afda_default_points():
    out = (ctypes.c_int * 13)()
    L.check(L.lib().vi_afda_default_points(out))
    return list(out)


def observed_traces(dobs, nrec, nt, dtype):
    """Validate observed traces for a gradient of ``nt`` steps: ``dobs`` must be
    (nrec, nt_obs) with nt_obs >= nt (receiver-major, the layout ``traces()`` returns).
    Returns ``nt`` and the first ``nt`` columns as a contiguous array of ``dtype``. A short,
    transposed or flat array raises ValueError instead of being reinterpreted."""
    d = np.asarray(dobs)
    if d.ndim != 2:
        raise ValueError(f"dobs must be a 2-D (nrec, nt) array, got shape {d.shape}")
    if d.shape[0] != nrec:
        raise ValueError(f"dobs has {d.shape[0]} rows but the simulation has {nrec} receivers "
                         "(expected (nrec, nt))")
    nt = d.shape[1] if nt is None else int(nt)
    if nt < 1:
        raise ValueError("nt must be >= 1")
    if d.shape[1] < nt:
        raise ValueError(f"dobs holds {d.shape[1]} time samples, fewer than nt = {nt}")
    return nt, np.ascontiguousarray(d[:, :nt], dtype=dtype)


class Simulation:
    """One simulation on the GPU.

    ``precision`` is "float32" (the reference's GPU precision, utils.jl:4) or "float64".
    """

    def __init__(self, nx, nz, dx, dz, dt, precision="float32", nrelax=0, fd_order=(8, 8)):
        self.nx, self.nz, self.dx, self.dz, self.dt = int(nx), int(nz), float(dx), float(dz), float(dt)
        self.dtype = np.dtype(np.float32 if precision in ("float32", np.float32, "f32") else np.float64)
        self.nrelax = int(nrelax)
        self.fd_order = tuple(int(o) for o in fd_order)  # (x, z): 2, 4, 6 or 8
        g = L.GridDesc(self.nx, self.nz, self.dx, self.dz, self.dt,
                       L.VI_F32 if self.dtype == np.float32 else L.VI_F64, self.nrelax,
                       *self.fd_order)
        h = ctypes.c_void_p()
        L.check(L.lib().vi_create(ctypes.byref(g), ctypes.byref(h)))
        self._h = h
        self.cpml_thickness = 0
        self.vti = False
        self.nrec = 0
        self.rec_nt = 0
        self.model_host = None  # the model arrays as last uploaded (None: a shared window)
        self.taun = None

    # -- helpers ---------------------------------------------------------------------
    def _arr(self, a, shape=None):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype))
        if shape is not None and a.shape != shape:
            if a.size != int(np.prod(shape)):
                raise ValueError(f"expected shape {shape}, got {a.shape}")
            a = a.reshape(shape)
        return a

    def _ptr(self, a):
        return a.ctypes.data_as(ctypes.c_void_p)

    @property
    def handle(self):
        return self._h

    # -- configuration ---------------------------------------------------------------
    def set_model(self, rho, lam, mu, tau=None):
        shp = (self.nz, self.nx)
        arrs = [self._arr(np.broadcast_to(x, shp), shp) for x in (rho, lam, mu)]
        t = None
        if self.nrelax:
            if tau is None:
                raise ValueError("tau is required when nrelax > 0")
            t = self._arr(np.broadcast_to(tau, shp), shp)
        L.check(L.lib().vi_set_model(self._h, *(self._ptr(a) for a in arrs),
                                     self._ptr(t) if t is not None else None))
        self.vti = False
        self.model_host = dict(zip(("rho", "lam", "mu"), arrs), tau=t)

    def set_model_vti(self, rho, c11, c13, c33, c55, tau=None):
        """VTI medium: density and stiffnesses C11, C13, C33, C55 (VTIModel, Models.jl:27-33)."""
        shp = (self.nz, self.nx)
        arrs = [self._arr(np.broadcast_to(x, shp), shp) for x in (rho, c11, c13, c33, c55)]
        t = None
        if self.nrelax:
            if tau is None:
                raise ValueError("tau is required when nrelax > 0")
            t = self._arr(np.broadcast_to(tau, shp), shp)
        L.check(L.lib().vi_set_model_vti(self._h, *(self._ptr(a) for a in arrs),
                                         self._ptr(t) if t is not None else None))
        self.vti = True
        self.model_host = dict(zip(("rho", "c11", "c13", "c33", "c55"), arrs), tau=t)

    def set_model_dict(self, m):
        """Model from a dict: rho, lam, mu (isotropic) or rho, c11, c13, c33, c55 (VTI),
        plus tau when nrelax > 0."""
        if "c11" in m:
            self.set_model_vti(m["rho"], m["c11"], m["c13"], m["c33"], m["c55"], m.get("tau"))
        else:
            self.set_model(m["rho"], m["lam"], m["mu"], m.get("tau"))

    def set_model_window(self, model, x0):
        """Read the model from columns [x0, x0 + nx) of a shared device Model."""
        L.check(L.lib().vi_set_model_window(self._h, model.handle, int(x0)))
        self._model_ref = model
        self.vti = bool(getattr(model, "vti", False))
        self.model_host = None

    def device_bytes(self):
        b = ctypes.c_size_t()
        L.check(L.lib().vi_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def is_finite(self):
        """False once the wavefield holds a NaN or an infinity (an unstable step or model)."""
        f = ctypes.c_int()
        L.check(L.lib().vi_check_finite(self._h, ctypes.byref(f)))
        return bool(f.value)

    def trim(self):
        """Free the workspaces the handle keeps for its next gradient or rates call."""
        L.check(L.lib().vi_trim(self._h))

    def gradient_bytes(self, nt, checkpoint_every=0):
        """Temporary device bytes of one gradient(…, nt, checkpoint_every) call."""
        b = ctypes.c_size_t()
        L.check(L.lib().vi_gradient_bytes(self._h, int(nt), int(checkpoint_every), ctypes.byref(b)))
        return b.value

    def checkpoint_for_memory(self, nt, max_bytes):
        """The largest checkpoint interval whose gradient of ``nt`` steps needs at most
        ``max_bytes`` of device memory (0 = ceil(sqrt(nt)) when none does); see
        ``largest_checkpoint_interval``."""
        return largest_checkpoint_interval(nt, lambda K: self.gradient_bytes(nt, K), max_bytes)

    def set_relaxation(self, taun):
        t = np.ascontiguousarray(np.asarray(taun, dtype=np.float64).ravel())
        if t.size != self.nrelax:
            raise ValueError(f"expected {self.nrelax} relaxation times")
        L.check(L.lib().vi_set_relaxation(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.taun = t.copy()

    def model_differs(self, **arrays):
        """Whether any given model array (rho, lam, mu, c11, c13, c33, c55, tau; None = not
        given) differs from what was last uploaded, in this simulation's precision. A model
        read from a shared device Model always counts as different."""
        if self.model_host is None:
            return True
        shp = (self.nz, self.nx)
        for k, v in arrays.items():
            if v is None:
                continue
            cur = self.model_host.get(k)
            if cur is None:
                return True
            new = np.broadcast_to(np.asarray(v, dtype=self.dtype), shp)
            if not np.array_equal(cur, new):
                return True
        return False

    def set_cpml(self, thickness=0, d0=0.0, alpha_max=0.0, kappa_max=1.0, power=2.0,
                 left=True, right=True, top=False, bottom=True):
        if not thickness:
            L.check(L.lib().vi_set_cpml(self._h, None))
            self.cpml_thickness = 0
            return
        c = L.CpmlDesc(int(thickness), float(d0), float(alpha_max), float(kappa_max),
                       float(power), int(left), int(right), int(top), int(bottom))
        L.check(L.lib().vi_set_cpml(self._h, ctypes.byref(c)))
        self.cpml_thickness = int(thickness)

    def set_free_surface(self, mode=True):
        """False/"none": Dirichlet top; True/"image": stress imaging (FreeSurfaceBC,
        CPML.jl:366-393); "afda": the H-AFDA one-sided surface (free-surface-kernels.jl);
        "hermitian": H-AFDA with formula #3 on row 5 (free-surface-kernels.jl:154-214)."""
        L.check(L.lib().vi_set_free_surface(self._h, free_surface_mode(mode)))

    def set_afda_points(self, counts=None):
        """Points of the 13 H-AFDA stencils (None: the default); see ``afda_weights``."""
        cp = None
        if counts is not None:
            cp = (ctypes.c_int * 13)(*[int(c) for c in counts])
        L.check(L.lib().vi_set_afda_points(self._h, cp))

    def set_afda_hermitian(self, points=5):
        """Half points of formula #3 (2..8); see ``afda_hermitian_weights``."""
        L.check(L.lib().vi_set_afda_hermitian(self._h, int(points)))

    def set_sources(self, ix, iz, kind, wavelets):
        ix, iz = _as_int32(ix), _as_int32(iz)
        kinds = _as_int32([SOURCE_KINDS.get(k, k) if isinstance(k, str) else k
                           for k in np.atleast_1d(kind)])
        n = ix.size
        w = self._arr(wavelets).reshape(n, -1) if n else self._arr(np.zeros((0, 0)))
        IP = ctypes.POINTER(ctypes.c_int)
        L.check(L.lib().vi_set_sources(self._h, n, ix.ctypes.data_as(IP), iz.ctypes.data_as(IP),
                                       kinds.ctypes.data_as(IP), w.shape[1] if n else 0,
                                       self._ptr(w)))

    def set_receivers(self, ix, iz, comp, nt_max):
        ix, iz = _as_int32(ix), _as_int32(iz)
        comps = _as_int32([COMPONENTS.get(c, c) if isinstance(c, str) else c
                           for c in np.broadcast_to(np.asarray(comp, dtype=object), ix.shape)])
        IP = ctypes.POINTER(ctypes.c_int)
        L.check(L.lib().vi_set_receivers(self._h, ix.size, ix.ctypes.data_as(IP),
                                         iz.ctypes.data_as(IP), comps.ctypes.data_as(IP),
                                         int(nt_max)))
        self.nrec, self.rec_nt = ix.size, int(nt_max)

    # -- state -------------------------------------------------------------------------
    def _shape(self, fid):
        if fid <= L.VI_SXZ:
            return (self.nz, self.nx)
        if fid <= L.VI_MXZ:
            return (self.nrelax, self.nz, self.nx)
        T2 = 2 * self.cpml_thickness
        return (self.nz, T2) if fid < L.VI_PSI_VX_Z else (T2, self.nx)

    def get(self, field):
        fid = FIELDS[field] if isinstance(field, str) else int(field)
        out = np.empty(self._shape(fid), dtype=self.dtype)
        L.check(L.lib().vi_get_state(self._h, fid, self._ptr(out)))
        return out

    def set(self, field, values):
        fid = FIELDS[field] if isinstance(field, str) else int(field)
        a = self._arr(values, self._shape(fid))
        L.check(L.lib().vi_set_state(self._h, fid, self._ptr(a)))

    def reset(self):
        L.check(L.lib().vi_reset(self._h))

    # -- stepping ----------------------------------------------------------------------
    def step(self, nsteps=1):
        L.check(L.lib().vi_step(self._h, int(nsteps)))

    def sync(self):
        L.check(L.lib().vi_sync(self._h))

    @property
    def step_index(self):
        return int(L.lib().vi_step_index(self._h))

    def rates(self, src_values=None, memory=True):
        """Right-hand side at the current state (the state is not changed): dv (2, nz, nx) =
        (vx, vz), ds (3, nz, nx) = (sxx, szz, sxz) and, when ``memory`` and nrelax > 0,
        dm (3, L, nz, nx) (else None). ``src_values``: one sample per source in the order of
        set_sources (None: no source term). One half step of ``step`` is x + dt * rate."""
        dv = np.empty((2, self.nz, self.nx), dtype=self.dtype)
        ds = np.empty((3, self.nz, self.nx), dtype=self.dtype)
        dm = np.empty((3, self.nrelax, self.nz, self.nx), dtype=self.dtype) \
            if memory and self.nrelax else None
        sv = None
        if src_values is not None:
            sv = np.ascontiguousarray(np.asarray(src_values, dtype=self.dtype).ravel())
        L.check(L.lib().vi_rates(self._h, self._ptr(sv) if sv is not None else None,
                                 self._ptr(dv), self._ptr(ds),
                                 self._ptr(dm) if dm is not None else None))
        return dv, ds, dm

    def traces(self, nt=None):
        nt = min(self.step_index, self.rec_nt) if nt is None else int(nt)
        out = np.empty((self.nrec, nt), dtype=self.dtype)
        L.check(L.lib().vi_get_traces(self._h, self._ptr(out), nt))
        return out

    # -- gradient ----------------------------------------------------------------------
    def gradient(self, dobs, nt=None, checkpoint_every=0):
        """L2 trace misfit J and its gradient with respect to the model arrays by the discrete
        adjoint of ``nt`` steps from a zero state. ``dobs`` is (nrec, nt). Leaves the
        simulation reset. Returns (J, g) with g = {"rho", "lam", "mu"[, "tau"]} for an
        isotropic model and {"rho", "c11", "c13", "c33", "c55"[, "tau"]} for a VTI model."""
        nt, d = observed_traces(dobs, self.nrec, nt, self.dtype)
        shp = (self.nz, self.nx)
        keys = ("rho", "c11", "c13", "c33", "c55", "tau") if self.vti else ("rho", "lam", "mu", "tau")
        g = {k: np.empty(shp, dtype=self.dtype) for k in keys}  # every element is written
        J = ctypes.c_double()
        ptrs = [self._ptr(g[k]) for k in keys[:-1]] + [self._ptr(g["tau"]) if self.nrelax else None]
        fn = L.lib().vi_gradient_vti if self.vti else L.lib().vi_gradient
        L.check(fn(self._h, nt, self._ptr(d), int(checkpoint_every), ctypes.byref(J), *ptrs))
        if not self.nrelax:
            g.pop("tau")
        return J.value, g

    # -- tiling ------------------------------------------------------------------------
    def set_tiling(self, rows_per_wave):
        L.check(L.lib().vi_set_tiling(self._h, int(rows_per_wave)))

    @property
    def tiling(self):
        return int(L.lib().vi_get_tiling(self._h))

    def autotune(self, trial_steps=20):
        """Pick the fastest kernel variant for this configuration; the state is unchanged.
        Returns (chosen rows per wave, {rows: us per step})."""
        chosen = ctypes.c_int()
        us = (ctypes.c_double * 4)()
        L.check(L.lib().vi_autotune(self._h, int(trial_steps), ctypes.byref(chosen), us))
        return chosen.value, dict(zip((1, 2, 4, 8), list(us)))

    # -- timing ------------------------------------------------------------------------
    def profile(self, every=1):
        """Time the launches of every ``every``-th step with dispatch events (0 = off)."""
        L.check(L.lib().vi_profile_enable(self._h, int(every)))

    def profile_read(self, kernel):
        ms, n = ctypes.c_double(), ctypes.c_long()
        L.check(L.lib().vi_profile_read(self._h, int(kernel), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def algorithmic_bytes(self, kernel):
        b = ctypes.c_double()
        L.check(L.lib().vi_algorithmic_bytes(self._h, int(kernel), ctypes.byref(b)))
        return b.value

    # -- lifetime ----------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            L.lib().vi_destroy(self._h)
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


class GradientJob:
    """The gradient of one Simulation (as Simulation.gradient) computed as a job that
    advances in chunks without blocking: ``advance(k)`` enqueues at most k steps of work,
    ``idle()`` tells whether the device has finished what was enqueued, ``finished`` whether
    everything is enqueued, ``end()`` returns (J, g). The simulation belongs to the job until
    end() or cancel()."""

    def __init__(self, sim, dobs, nt=None, checkpoint_every=0):
        nt, d = observed_traces(dobs, sim.nrec, nt, sim.dtype)
        self.sim, self.nt = sim, nt
        h = ctypes.c_void_p()
        L.check(L.lib().vi_gradient_begin(sim.handle, nt, sim._ptr(d), int(checkpoint_every),
                                          ctypes.byref(h)))
        self._h = h
        self.finished = False

    def advance(self, max_steps):
        fin = ctypes.c_int()
        L.check(L.lib().vi_gradient_advance(self._h, int(max_steps), ctypes.byref(fin)))
        self.finished = bool(fin.value)
        return self.finished

    def idle(self):
        r = L.lib().vi_gradient_idle(self._h)
        if r < 0:
            L.check(r)
        return r == 1

    def end(self):
        sim = self.sim
        shp = (sim.nz, sim.nx)
        keys = (("rho", "c13", "c55", "tau", "c11", "c33") if sim.vti
                else ("rho", "lam", "mu", "tau"))
        g = {k: np.zeros(shp, dtype=sim.dtype) for k in keys if k != "tau" or sim.nrelax}
        ptr = lambda k: sim._ptr(g[k]) if k in g else None
        J = ctypes.c_double()
        order = keys + (None,) * (6 - len(keys))
        h, self._h = self._h, None
        L.check(L.lib().vi_gradient_end(h, ctypes.byref(J), *(ptr(k) for k in order)))
        return J.value, g

    def cancel(self):
        if self._h:
            h, self._h = self._h, None
            L.check(L.lib().vi_gradient_cancel(h))


def largest_checkpoint_interval(nt, bytes_of, max_bytes):
    """The largest checkpoint interval K whose gradient needs at most ``max_bytes`` according
    to ``bytes_of(K)``, or 0 (= ceil(sqrt(nt)) checkpoints) when none of the intervals
    ceil(nt / s), s = 1 .. ceil(sqrt(nt)), fits. A gradient with interval K recomputes
    about nt - K forward steps (every segment but the last), so the largest interval that fits
    is the fastest; the gradient itself does not depend on K."""
    nt = int(nt)
    smax = int(np.ceil(np.sqrt(nt)))
    for s in range(1, smax + 1):
        K = -(-nt // s)
        if bytes_of(K) <= max_bytes:
            return K
    return 0


def plan_bytes(nx, nz, precision="float32", nrelax=0, vti=False, cpml_thickness=0, nsrc=0,
               src_nt=0, nrec=0, rec_nt=0, grad_nt=0, checkpoint_every=0):
    """Upper bound of the device bytes a shot holds (plus one gradient call when
    grad_nt > 0), computed by the library without a device."""
    dt = np.dtype(np.float32 if precision in ("float32", np.float32, "f32") else np.float64)
    g = L.GridDesc(int(nx), int(nz), 1.0, 1.0, 1.0, L.VI_F32 if dt == np.float32 else L.VI_F64,
                   int(nrelax), 8, 8)
    b = ctypes.c_size_t()
    L.check(L.lib().vi_plan_bytes(ctypes.byref(g), int(bool(vti)), int(cpml_thickness), int(nsrc),
                                  int(src_nt), int(nrec), int(rec_nt), int(grad_nt),
                                  int(checkpoint_every), ctypes.byref(b)))
    return b.value
