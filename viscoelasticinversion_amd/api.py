"""Host-side mirror of ViscoelasticInversion's solver API, backed by the HIP library.

Names, keyword arguments, defaults and error behaviour follow the reference:
  _ricker, RickerSource            src/sciml_solver/Source.jl:3-58
  CPMLConfig                       src/sciml_solver/CPML.jl:28-63
  FreeSurfaceBC                    src/sciml_solver/CPML.jl:366-370
  DirichletBC / get_bc             src/sciml_solver/BoundaryConditions.jl:3-11, 88-100
  IsotropicModel                   src/sciml_solver/Models.jl:3-25
  K, dK_dω, get_τ, get_Q, QModel   src/sciml_solver/Models.jl:65-107
  Parameters                       src/sciml_solver/Parameters.jl:31-86
  make_problem, solve_problem      src/sciml_solver/ViscoelasticProblem.jl:273-309
  stress_eq, velocity_eq           src/sciml_solver/ViscoelasticProblem.jl:62-174
  Fields, leapfrog_step!           src/visco-kernels.jl:87-113
The time integrator is the fused staggered leapfrog of visco-kernels.jl (the adaptive
SciML Tsit5 path of the reference is replaced by it; ``solver`` only accepts "leapfrog").
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .stepper import Simulation, free_surface_mode


# ---------------------------------------------------------------------------- source
def _ricker(f, t):
    """Ricker wavelet (1 - 2(pi f t)^2) exp(-(pi f t)^2)   (Source.jl:3-5)."""
    a = (np.pi * f * np.asarray(t, dtype=np.float64)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


class RickerSource:
    """Point Ricker source (Source.jl:7-33).

    ``kind`` is "force_z" (the reference: added to dvz before buoyancy), "force_x" or
    "explosive" (added to the sxx and szz rates). The active window ``tspan`` is the open
    interval (tc - 2/f, tc + 2/f) unless given; |r| there is below 6e-16 at the ends.
    Grid indices are 0-based (ix, iz).
    """

    def __init__(self, freq, tc, nx, nz, coef=1.0, kind="force_z", tspan=None, rtol=1e-6):
        if freq <= 0:
            raise ValueError("freq must be positive")
        self.freq, self.tc = float(freq), float(tc)
        self.grid = (int(nx), int(nz))
        self.coefs = float(coef)
        self.kind = kind
        self.tspan = tuple(tspan) if tspan is not None else (tc - 2.0 / freq, tc + 2.0 / freq)

    def value(self, t):
        """Source sample c * r(t - tc) inside the open window, else 0 (Source.jl:42-58)."""
        t = np.asarray(t, dtype=np.float64)
        inside = (self.tspan[0] < t) & (t < self.tspan[1])
        return np.where(inside, self.coefs * _ricker(self.freq, t - self.tc), 0.0)

    def __call__(self, du, t):
        """Add the sample at time t to du[nz, nx] (inject_source_kernel!, Source.jl:35-40)."""
        du[self.grid[1], self.grid[0]] += self.value(t)
        return None

    def samples(self, nt, dt, t0=0.0):
        """Samples used by step k: force sources at t0 + k dt, explosive at t0 + (k+1/2) dt."""
        off = 0.5 if self.kind == "explosive" else 0.0
        return self.value(t0 + (np.arange(nt) + off) * dt)


# ---------------------------------------------------------------------------- CPML
@dataclass
class CPMLConfig:
    """CPML configuration (CPML.jl:28-63)."""
    thickness: int = 10
    d0: Optional[float] = None
    alpha_max: float = 0.0
    kappa_max: float = 1.0
    power: float = 2.0
    left: bool = True
    right: bool = True
    top: bool = False
    bottom: bool = True
    vmax: Optional[float] = None
    dx: Optional[float] = None
    reflection_coef: float = 1e-5

    def __post_init__(self):
        if self.d0 is None:
            if self.vmax is None or self.dx is None:
                raise ValueError("Must provide either d0 or (vmax and dx) for automatic "
                                 "damping calculation")
            # Collino & Tsogka (2001), CPML.jl:60
            self.d0 = -(self.power + 1) * self.vmax * math.log(self.reflection_coef) / (
                2 * self.thickness * self.dx)


@dataclass
class FreeSurfaceBC:
    """Stress-imaging free surface on row 4 (CPML.jl:366-393). ``depth`` is kept for API
    compatibility; the 8th-order stencil always uses 4 image rows."""
    depth: int = 1


class DirichletBC:
    """Fields outside the padded interior stay at their initial value (BoundaryConditions.jl:3-11)."""

    def __call__(self, du):
        return None


def get_bc(bc_type):
    """BoundaryConditions.jl:88-100. Only :dirichlet is implemented; periodic and Neumann are
    outside this library's scope."""
    key = str(bc_type).lstrip(":")
    if key == "dirichlet":
        return DirichletBC()
    if key in ("neumann", "periodic", "absorbing"):
        raise NotImplementedError(f"boundary condition {key!r} is not implemented; use "
                                  "'dirichlet' with an optional CPMLConfig")
    raise ValueError(f"Unknown boundary condition type: {bc_type}")


# ---------------------------------------------------------------------------- models
class IsotropicModel:
    """Isotropic model from density and Lamé parameters, scalars or (Nz, Nx) arrays
    (Models.jl:3-25; fields b, λ, μ)."""

    def __init__(self, rho, lam, mu, Nx, Nz):
        shp = (int(Nz), int(Nx))
        self.rho = np.array(np.broadcast_to(rho, shp), dtype=np.float64)
        self.lam = np.array(np.broadcast_to(lam, shp), dtype=np.float64)
        self.mu = np.array(np.broadcast_to(mu, shp), dtype=np.float64)

    @property
    def b(self):
        return 1.0 / self.rho

    @classmethod
    def from_velocities(cls, vp, vs, rho, Nx, Nz):
        rho = np.broadcast_to(rho, (Nz, Nx)).astype(np.float64)
        mu = rho * np.asarray(vs, dtype=np.float64) ** 2
        lam = rho * np.asarray(vp, dtype=np.float64) ** 2 - 2.0 * mu
        return cls(rho, lam, mu, Nx, Nz)

    def vmax(self):
        return float(np.sqrt((self.lam + 2 * self.mu) / self.rho).max())


def thomsen_to_stiffness(vp0, vs0, rho, eps, delta):
    """Thomsen parameters of a VTI medium to stiffnesses (Pa): C33 = rho vp0^2,
    C55 = rho vs0^2, C11 = C33 (1 + 2 eps) and, from
    delta = ((C13 + C55)^2 - (C33 - C55)^2) / (2 C33 (C33 - C55)),
    C13 = sqrt((C33 - C55) (C33 (1 + 2 delta) - C55)) - C55."""
    vp0, vs0, rho, eps, delta = (np.asarray(a, dtype=np.float64)
                                 for a in (vp0, vs0, rho, eps, delta))
    c33 = rho * vp0 ** 2
    c55 = rho * vs0 ** 2
    c11 = c33 * (1.0 + 2.0 * eps)
    c13 = np.sqrt((c33 - c55) * (c33 * (1.0 + 2.0 * delta) - c55)) - c55
    return c11, c13, c33, c55


class VTIModel:
    """Vertically transverse isotropic model from density and the stiffnesses C11, C13,
    C33, C55, scalars or (Nz, Nx) arrays (Models.jl:27-33; fields b, C11, C13, C33, C55).
    The stress update is the TTI form of ViscoelasticProblem.jl:21-35 with C15 = C35 = 0."""

    def __init__(self, rho, C11, C13, C33, C55, Nx, Nz):
        shp = (int(Nz), int(Nx))
        f = lambda a: np.array(np.broadcast_to(a, shp), dtype=np.float64)
        self.rho, self.C11, self.C13, self.C33, self.C55 = f(rho), f(C11), f(C13), f(C33), f(C55)

    @property
    def b(self):
        return 1.0 / self.rho

    @classmethod
    def from_thomsen(cls, vp0, vs0, rho, eps, delta, Nx, Nz):
        c11, c13, c33, c55 = thomsen_to_stiffness(vp0, vs0, rho, eps, delta)
        return cls(rho, c11, c13, c33, c55, Nx, Nz)

    def vmax(self):
        return float(np.sqrt(np.maximum(self.C11, self.C33) / self.rho).max())


def K(w, taun):
    """K(ω) = Σ ωτₙ/(1+τₙ²ω²)   (Models.jl:65)."""
    w = np.asarray(w, dtype=np.float64)[..., None]
    t = np.asarray(taun, dtype=np.float64)
    return np.sum(w * t / (1.0 + t ** 2 * w ** 2), axis=-1)


def dK_dw(w, taun):
    """dK/dω   (Models.jl:66)."""
    w = np.asarray(w, dtype=np.float64)[..., None]
    t = np.asarray(taun, dtype=np.float64)
    return np.sum(t * (1.0 - t ** 2 * w ** 2) / (1.0 + t ** 2 * w ** 2) ** 2, axis=-1)


def get_tau(Q0, Kbar, N):
    """τ = N/(K̄ Q₀)   (Models.jl:67)."""
    return N / (Kbar * Q0)


def get_Q(w, tau, taun):
    """Q(ω) = N/(τ K(ω))   (Models.jl:68)."""
    return len(np.atleast_1d(taun)) / (tau * K(w, taun))


def Q_model_J2(taun, w1, w2):
    """∫_{ω1}^{ω2} (dK/dω)² dω   (Models.jl:70-74)."""
    from scipy.integrate import quad
    val, _ = quad(lambda w: float(dK_dw(w, taun)) ** 2, w1, w2, limit=200)
    return val


def find_optimal_taun(w1, w2, N):
    """Nelder-Mead minimisation of Q_model_J2 from τₙ = 1/linspace(ω1, ω2, N)
    (Models.jl:77-87). The objective is scale-degenerate, so the result depends on the
    optimiser's path; callers that need reproducible τₙ pass them explicitly."""
    from scipy.optimize import minimize
    t0 = 1.0 / np.linspace(w1, w2, N)
    res = minimize(lambda t: Q_model_J2(t, w1, w2), t0, method="Nelder-Mead",
                   options=dict(maxiter=10 ** 6, maxfev=10 ** 6))
    return np.sort(res.x)


class QModel:
    """Nearly-constant-Q SLS model (Models.jl:56-107): τₙ (N mechanisms) and the field τ."""

    def __init__(self, f1, f2, Nx, Nz, Q0=100, N=5, taun=None):
        w1, w2 = 2 * np.pi * f1, 2 * np.pi * f2
        self.f1, self.f2, self.N = float(f1), float(f2), int(N)
        self.taun = np.sort(np.asarray(taun, dtype=np.float64)) if taun is not None else \
            find_optimal_taun(w1, w2, N)
        self.K = float(np.mean(K(np.linspace(w1, w2, 100), self.taun)))
        self.tau = np.full((int(Nz), int(Nx)), get_tau(Q0, self.K, self.N))

    def update(self, Q):
        """τ = get_τ.(Q, K̄, N)   (Models.jl:104-107)."""
        self.tau = get_tau(np.asarray(Q, dtype=np.float64), self.K, self.N)

    def Q(self, w):
        return get_Q(w, float(np.mean(self.tau)), self.taun)


# ---------------------------------------------------------------------------- parameters
@dataclass
class Parameters:
    """Solver configuration (Parameters.jl:31-86), keyword-constructed like the reference.

    Differences: ``bc`` defaults to "dirichlet" (the reference's periodic BC is defective
    and out of scope), ``device`` must be "mi355x", ``precision`` selects FP32 (default,
    the reference's GPU precision) or FP64, ``qmodel`` adds SLS attenuation, ``receivers``
    is an optional (ix, iz, comp) triple.
    """
    Nx: int = 64
    Nz: int = 64
    dx: float = 10.0
    dz: float = 10.0
    dt: float = 0.001
    fd_order_x: int = 8
    fd_order_z: int = 8
    bc: str = "dirichlet"
    source: Optional[object] = None
    device: str = "mi355x"
    model: Optional[object] = None  # IsotropicModel or VTIModel
    cpml: Optional[CPMLConfig] = None
    # False, True / "image" (stress imaging), "afda" (H-AFDA) or "hermitian" (H-AFDA with
    # formula #3 on the first row below the surface)
    free_surface: object = False
    precision: str = "float32"
    qmodel: Optional[QModel] = None
    receivers: Optional[tuple] = None

    def __post_init__(self):
        if self.fd_order_x not in (2, 4, 6, 8) or self.fd_order_z not in (2, 4, 6, 8):
            raise NotImplementedError("staggered operators of order 2, 4, 6 or 8 are implemented")
        if str(self.device).lower() not in ("mi355x", "gpu", "hip"):
            raise ValueError(f"unsupported device {self.device!r}: this library runs on MI355X")
        self.bc_obj = get_bc(self.bc)
        if self.source is None:
            self.source = RickerSource(40.0, 0.2, self.Nx // 2, self.Nz // 2)  # Parameters.jl:59
        if self.model is None:
            self.model = IsotropicModel(2000.0, 1e9, 1e9, self.Nx, self.Nz)  # Parameters.jl:63
        self.free_surface_bc = FreeSurfaceBC() if self.free_surface else None
        self.sources = self.source if isinstance(self.source, (list, tuple)) else [self.source]

    @property
    def nrelax(self):
        return self.qmodel.N if self.qmodel is not None else 0

    def build(self) -> Simulation:
        """Create and configure a device simulation for these parameters."""
        sim = Simulation(self.Nx, self.Nz, self.dx, self.dz, self.dt, self.precision,
                         self.nrelax, (self.fd_order_x, self.fd_order_z))
        m = self.model
        tau = self.qmodel.tau if self.qmodel is not None else None
        if isinstance(m, VTIModel):
            sim.set_model_vti(m.rho, m.C11, m.C13, m.C33, m.C55, tau)
        else:
            sim.set_model(m.rho, m.lam, m.mu, tau)
        if self.qmodel is not None:
            sim.set_relaxation(self.qmodel.taun)
        if self.cpml is not None:
            c = self.cpml
            sim.set_cpml(c.thickness, c.d0, c.alpha_max, c.kappa_max, c.power, c.left, c.right,
                         c.top, c.bottom)
        sim.set_free_surface(self.free_surface)
        return sim


# ---------------------------------------------------------------------------- problem
@dataclass
class Problem:
    """What make_problem returns: initial state u0 = (s0, v0) with s0 (3, Nz, Nx) holding
    [σxx, σzz, σxz] and v0 (2, Nz, Nx) holding [vx, vz], the time span and the parameters."""
    u0: tuple
    tspan: tuple
    p: Parameters


def make_problem(parameters, s0=None, v0=None, tspan=(0.0, 2.0)):
    """ViscoelasticProblem.jl:273-292 (zero initial state by default)."""
    shp = (parameters.Nz, parameters.Nx)
    s0 = np.zeros((3,) + shp) if s0 is None else np.asarray(s0, dtype=np.float64)
    v0 = np.zeros((2,) + shp) if v0 is None else np.asarray(v0, dtype=np.float64)
    if s0.shape != (3,) + shp or v0.shape != (2,) + shp:
        raise ValueError("s0 must be (3, Nz, Nx) and v0 (2, Nz, Nx)")
    return Problem((s0, v0), tuple(tspan), parameters)


@dataclass
class Solution:
    t: np.ndarray
    s: np.ndarray            # (3, Nz, Nx) final stress
    v: np.ndarray            # (2, Nz, Nx) final velocity
    traces: Optional[np.ndarray] = None   # (nrec, nt)
    retcode: str = "Success"
    saved: list = field(default_factory=list)


def solve_problem(problem, solver="leapfrog", saveat=None, **kwargs):
    """Integrate from tspan[0] to tspan[1] with nt = round((t1 - t0)/dt) fused leapfrog
    steps on the GPU (ViscoelasticProblem.jl:307-309). ``saveat`` (a step interval) stores
    (t, s, v) snapshots."""
    if solver not in ("leapfrog", None):
        raise NotImplementedError("only the fused leapfrog integrator is available")
    p = problem.p
    t0, t1 = problem.tspan
    nt = int(round((t1 - t0) / p.dt))
    sim = p.build()
    try:
        s0, v0 = problem.u0
        for k, name in enumerate(("sxx", "szz", "sxz")):
            sim.set(name, s0[k])
        for k, name in enumerate(("vx", "vz")):
            sim.set(name, v0[k])
        if p.sources:
            ix = [s.grid[0] for s in p.sources]
            iz = [s.grid[1] for s in p.sources]
            kinds = [getattr(s, "kind", "force_z") for s in p.sources]
            w = np.stack([s.samples(nt, p.dt, t0) for s in p.sources])
            sim.set_sources(ix, iz, kinds, w)
        if p.receivers is not None:
            rix, riz, rc = p.receivers
            sim.set_receivers(rix, riz, rc, nt)
        saved = []
        every = int(saveat) if saveat else 0
        done = 0
        while done < nt:
            n = nt - done if not every else min(every, nt - done)
            sim.step(n)
            done += n
            if every:
                saved.append((t0 + done * p.dt,
                              np.stack([sim.get(f) for f in ("sxx", "szz", "sxz")]),
                              np.stack([sim.get(f) for f in ("vx", "vz")])))
        s = np.stack([sim.get(f) for f in ("sxx", "szz", "sxz")])
        v = np.stack([sim.get(f) for f in ("vx", "vz")])
        tr = sim.traces(nt) if p.receivers is not None else None
        return Solution(t=t0 + np.arange(1, nt + 1) * p.dt, s=s, v=v, traces=tr, saved=saved)
    finally:
        sim.close()


# ---------------------------------------------------------------------------- stability
# sum of |weights| x h of the staggered operator of each order (8: 1225/1024, -245/3072,
# 49/5120, -5/7168): the largest magnitude of its symbol
FD_WEIGHT_SUMS = {2: 1.0, 4: 9 / 8 + 1 / 24, 6: 75 / 64 + 25 / 384 + 3 / 640,
                  8: 1225 / 1024 + 245 / 3072 + 49 / 5120 + 5 / 7168}
FD_WEIGHT_SUM = FD_WEIGHT_SUMS[8]


def stable_dt(model, dx, dz, fd_order=(8, 8)):
    """Largest stable time step of the staggered leapfrog for a model dict (rho and lam, mu or
    c11, c33, and tau when viscoelastic) and operator orders (x, z): the von Neumann bound
    dt <= 1 / (vmax * sqrt((Sx/dx)^2 + (Sz/dz)^2)), S = sum|c| of the order's weights, with
    vmax the largest unrelaxed P-wave speed, sqrt(M (1 + tau) / rho), M = lam + 2 mu
    (isotropic) or max(C11, C33)."""
    rho = np.asarray(model["rho"], np.float64)
    if "c11" in model:
        m = np.maximum(np.asarray(model["c11"], np.float64), np.asarray(model["c33"], np.float64))
    else:
        m = np.asarray(model["lam"], np.float64) + 2 * np.asarray(model["mu"], np.float64)
    if model.get("tau") is not None:
        m = m * (1 + np.asarray(model["tau"], np.float64))
    vmax = float(np.sqrt(m / rho).max())
    sx, sz = (FD_WEIGHT_SUMS[int(o)] for o in fd_order)
    return 1.0 / (vmax * math.sqrt((sx / dx) ** 2 + (sz / dz) ** 2))


# ---------------------------------------------------------------------------- right-hand side
def _rhs_simulation(p):
    """The device simulation that evaluates the right-hand side for the Parameters ``p``:
    created on first use and kept on ``p`` (model, CPML, free surface and source positions of
    ``p``; CPML memory zero, as the reference's RHS never advances it)."""
    sim = getattr(p, "_rhs_sim", None)
    if sim is None:
        sim = p.build()
        if p.sources:
            sim.set_sources([s.grid[0] for s in p.sources], [s.grid[1] for s in p.sources],
                            [getattr(s, "kind", "force_z") for s in p.sources],
                            np.zeros((len(p.sources), 1)))
        p._rhs_sim = sim
    return sim


def _apply_images(s):
    """Stress imaging on the state (the reference's apply_free_surface!(s), CPML.jl:382-393,
    for the 8th-order stencil's 4 image rows above the surface row 4): szz[4-m] = -szz[4+m],
    sxz[4-m] = -sxz[3+m] on the interior columns."""
    nx = s.shape[-1]
    for m in range(1, 5):
        s[1, 4 - m, 4:nx - 4] = -s[1, 4 + m, 4:nx - 4]
        s[2, 4 - m, 4:nx - 4] = -s[2, 3 + m, 4:nx - 4]


def _rates(s, v, p, t, M):
    sim = _rhs_simulation(p)
    s = np.asarray(s)
    v = np.asarray(v)
    if s.shape != (3, p.Nz, p.Nx) or v.shape != (2, p.Nz, p.Nx):
        raise ValueError("s must be (3, Nz, Nx) and v (2, Nz, Nx)")
    for k, name in enumerate(("sxx", "szz", "sxz")):
        sim.set(name, s[k])
    for k, name in enumerate(("vx", "vz")):
        sim.set(name, v[k])
    if sim.nrelax:
        m = np.zeros((3, sim.nrelax, p.Nz, p.Nx)) if M is None else np.asarray(M)
        for k, name in enumerate(("mxx", "mzz", "mxz")):
            sim.set(name, m[k])
    vals = [float(src.value(t)) for src in p.sources] if p.sources else None
    return sim.rates(vals, memory=True)


def stress_eq(ds, s, v, p, t, M=None, dM=None):
    """stress_eq!(ds, s, v, p, t) (ViscoelasticProblem.jl:62-102): ds (3, Nz, Nx) <- the
    stress rates (sxx, szz, sxz) of the state s (3, Nz, Nx), v (2, Nz, Nx) at time t, on the
    GPU. ``M`` (3, L, Nz, Nx) are the memory variables (None: zero) and ``dM`` receives their
    rates when given. Explosive sources add their sample at t."""
    _, r, dm = _rates(s, v, p, t, M)
    ds[...] = r
    if dM is not None and dm is not None:
        dM[...] = dm
    return None


def velocity_eq(dv, s, v, p, t, M=None):
    """velocity_eq!(dv, s, v, p, t) (ViscoelasticProblem.jl:110-174): dv (2, Nz, Nx) <- the
    velocity rates (vx, vz) at time t, force sources added before the buoyancy. With the
    stress-imaging free surface the images are first written into ``s``, as the reference's
    apply_free_surface!(s) does inside velocity_eq!."""
    if free_surface_mode(p.free_surface) == 1:
        _apply_images(s)
    r, _, _ = _rates(s, v, p, t, M)
    dv[...] = r
    return None


# ---------------------------------------------------------------------------- fused step
@dataclass
class Fields:
    """Fields(s, M, v) of visco-kernels.jl:102-106, held on the device by ``sim``."""
    sim: Simulation


@dataclass
class StepParameters:
    """The fused step's Parameters(λs, μs, τs, τns) of visco-kernels.jl:108-113 (arrays
    (Nz, Nx); τns the L relaxation times). ``rho`` is the density the velocity kernel reads
    (visco-kernels.jl:10-11); None keeps the density already loaded."""
    lam: object
    mu: object
    tau: object = None
    taun: object = None
    rho: object = None


def apply_parameters(sim, parameters):
    """Load the model of ``parameters`` into ``sim`` where it differs from what the
    simulation holds: a StepParameters (λs, μs, τs, τns[, ρ]) or a Parameters (its
    IsotropicModel / VTIModel and QModel). Unchanged arrays are not uploaded again."""
    if isinstance(parameters, StepParameters):
        p = parameters
        if sim.vti:
            raise ValueError("this simulation has a VTI model: pass a Parameters with a VTIModel")
        if sim.nrelax and p.tau is None and (sim.model_host is None or sim.model_host.get("tau") is None):
            raise ValueError("τs are required when the simulation has memory variables")
        if sim.model_differs(rho=p.rho, lam=p.lam, mu=p.mu, tau=p.tau if sim.nrelax else None):
            if p.rho is None and (sim.model_host is None or sim.model_host.get("rho") is None):
                raise ValueError("rho is required: the simulation holds no density to keep")
            rho = p.rho if p.rho is not None else sim.model_host["rho"]
            tau = p.tau if p.tau is not None else (sim.model_host or {}).get("tau")
            sim.set_model(rho, p.lam, p.mu, tau if sim.nrelax else None)
        taun = p.taun
    elif isinstance(parameters, Parameters):
        p = parameters
        if (p.Nx, p.Nz) != (sim.nx, sim.nz):
            raise ValueError("parameters grid differs from the simulation's grid")
        if p.nrelax != sim.nrelax:
            raise ValueError("parameters have another number of SLS mechanisms than the simulation")
        m = p.model
        tau = p.qmodel.tau if p.qmodel is not None else None
        if isinstance(m, VTIModel):
            if not sim.vti or sim.model_differs(rho=m.rho, c11=m.C11, c13=m.C13, c33=m.C33,
                                                c55=m.C55, tau=tau):
                sim.set_model_vti(m.rho, m.C11, m.C13, m.C33, m.C55, tau)
        elif sim.vti or sim.model_differs(rho=m.rho, lam=m.lam, mu=m.mu, tau=tau):
            sim.set_model(m.rho, m.lam, m.mu, tau)
        taun = p.qmodel.taun if p.qmodel is not None else None
    else:
        raise TypeError("parameters must be a StepParameters or a Parameters")
    if sim.nrelax and taun is not None:
        t = np.asarray(taun, dtype=np.float64).ravel()
        if sim.taun is None or not np.array_equal(sim.taun, t):
            sim.set_relaxation(t)


def leapfrog_step(fields, parameters=None, dt=None, nsteps=1):
    """leapfrog_step!(fields, parameters, dt) (visco-kernels.jl:87-100): advance the device
    state by full steps (velocity kernel, then stress kernel) with the model of
    ``parameters`` (StepParameters or Parameters; loaded where it differs from the
    simulation's, see ``apply_parameters``; None keeps the loaded model). ``dt`` must equal
    the simulation's dt when given."""
    sim = fields.sim if isinstance(fields, Fields) else fields
    if dt is not None and not math.isclose(dt, sim.dt, rel_tol=0, abs_tol=0):
        raise ValueError("dt differs from the simulation's dt")
    if parameters is not None:
        apply_parameters(sim, parameters)
    sim.step(nsteps)
    sim.sync()
    return None


# ---------------------------------------------------------------------------- gradients
def gradient_to_velocity_params(grads, rho, lam, mu, tau=None, Q=None):
    """Chain rule from the gradient with respect to (rho, lam, mu, tau) to (vp, vs, rho, Q)
    with lam = rho (vp^2 - 2 vs^2), mu = rho vs^2 and tau = L/(Kbar Q) (Models.jl:67), i.e.
    d tau / d Q = -tau / Q."""
    rho = np.asarray(rho, np.float64)
    vp = np.sqrt((np.asarray(lam) + 2 * np.asarray(mu)) / rho)
    vs = np.sqrt(np.asarray(mu) / rho)
    gl, gm, gr = (np.asarray(grads[k], np.float64) for k in ("lam", "mu", "rho"))
    out = {"vp": gl * 2 * rho * vp,
           "vs": gl * (-4 * rho * vs) + gm * 2 * rho * vs,
           "rho": gr + gl * (vp ** 2 - 2 * vs ** 2) + gm * vs ** 2}
    if tau is not None and Q is not None and "tau" in grads:
        out["Q"] = np.asarray(grads["tau"], np.float64) * (-np.asarray(tau) / np.asarray(Q))
    return out


def gradient_to_thomsen_params(grads, rho, vp0, vs0, eps, delta, Q=None, tau=None):
    """Chain rule from the gradient with respect to (rho, C11, C13, C33, C55[, tau]) of a VTI
    model to (vp0, vs0, rho, eps, delta[, Q]) with C33 = rho vp0^2, C55 = rho vs0^2,
    C11 = C33 (1 + 2 eps), C13 = sqrt((C33 - C55)(C33 (1 + 2 delta) - C55)) - C55
    (thomsen_to_stiffness) and tau = L/(Kbar Q)."""
    rho, vp0, vs0, eps, delta = (np.asarray(a, np.float64) for a in (rho, vp0, vs0, eps, delta))
    g11, g13, g33, g55, gr = (np.asarray(grads[k], np.float64)
                              for k in ("c11", "c13", "c33", "c55", "rho"))
    c11, c13, c33, c55 = thomsen_to_stiffness(vp0, vs0, rho, eps, delta)
    A = c33 - c55
    B = c33 * (1.0 + 2.0 * delta) - c55
    S = np.sqrt(A * B)
    d13_d33 = (B + A * (1.0 + 2.0 * delta)) / (2.0 * S)
    d13_d55 = -(A + B) / (2.0 * S) - 1.0
    out = {"vp0": (g11 * (1.0 + 2.0 * eps) + g33 + g13 * d13_d33) * 2.0 * rho * vp0,
           "vs0": (g55 + g13 * d13_d55) * 2.0 * rho * vs0,
           "rho": gr + (g11 * c11 + g13 * c13 + g33 * c33 + g55 * c55) / rho,
           "eps": g11 * 2.0 * c33,
           "delta": g13 * A * c33 / S}
    if Q is not None and tau is not None and "tau" in grads:
        out["Q"] = np.asarray(grads["tau"], np.float64) * (-np.asarray(tau) / np.asarray(Q))
    return out


def misfit_and_pullback(parameters, dobs, tspan=(0.0, 1.0), checkpoint_every=0):
    """Forward solve and L2 trace misfit with a pullback for the model, the analogue of a
    ChainRulesCore rrule for J(model): returns (J, pullback) where pullback(dJ) gives
    dJ * dJ/d(vp, vs, rho[, Q]) as arrays (Nz, Nx) for an IsotropicModel, and
    dJ * dJ/d(rho, C11, C13, C33, C55[, Q]) for a VTIModel (gradient_to_thomsen_params
    chains those to Thomsen parameters). The gradient comes from the discrete adjoint on the
    GPU with sqrt(nt) checkpointing."""
    p = parameters
    nt = int(round((tspan[1] - tspan[0]) / p.dt))
    if p.receivers is None:
        raise ValueError("parameters.receivers is required for a trace misfit")
    sim = p.build()
    try:
        ix = [s.grid[0] for s in p.sources]
        iz = [s.grid[1] for s in p.sources]
        kinds = [getattr(s, "kind", "force_z") for s in p.sources]
        sim.set_sources(ix, iz, kinds, np.stack([s.samples(nt, p.dt, tspan[0]) for s in p.sources]))
        rix, riz, rc = p.receivers
        sim.set_receivers(rix, riz, rc, nt)
        J, g = sim.gradient(dobs, nt, checkpoint_every)
    finally:
        sim.close()
    m = p.model
    q = p.qmodel
    Qarr = None
    if q is not None:
        Qarr = q.N / (q.K * q.tau)
    if isinstance(m, VTIModel):
        gv = {k: np.asarray(g[k], np.float64) for k in ("rho", "c11", "c13", "c33", "c55")}
        if q is not None:
            gv["Q"] = np.asarray(g["tau"], np.float64) * (-q.tau / Qarr)
    else:
        gv = gradient_to_velocity_params(g, m.rho, m.lam, m.mu, q.tau if q is not None else None,
                                         Qarr)

    def pullback(dJ=1.0):
        return {k: dJ * v for k, v in gv.items()}

    return J, pullback
