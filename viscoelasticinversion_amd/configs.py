"""Seeded synthetic workloads.

* ``config_small_elastic``: 256 x 256, dx = dz = 5 m, homogeneous vp 2000, vs 1150,
  rho 2000, dt 0.5 ms, 15 Hz Ricker explosive source at the centre, Dirichlet ring, 1000
  steps, 64 receivers on row 50.
* ``config_layered_visco``: 1024 x 512, dx = 5 m, 8 horizontal layers with vp ~ U[1800, 4500],
  vs = vp/1.8, Gardner rho = 310 vp^0.25, each with i.i.d. 2 % multiplicative noise per cell;
  Q = 50 with L = 3 SLS mechanisms for 2-40 Hz; 20-cell CPML on the sides and bottom; free
  surface on top; 12 Hz Ricker force_z source 10 m below the surface; 201 receivers 10 m
  below the surface; dt 0.4 ms; 4000 steps.

Each builder returns a plain dict so the same description drives the GPU stepper and the
CPU oracle.
"""
from __future__ import annotations

import numpy as np

from .api import K, _ricker, get_tau

# SLS relaxation times for Q0 = 50, L = 3, 2-40 Hz: Nelder-Mead minimum of the flatness
# objective of Models.jl:70-87 from the start 1/linspace(w1, w2, 3), computed once with
# api.find_optimal_taun and stored so that runs do not depend on the optimiser path.
TAUN_2_40HZ_L3 = (0.0013110983134659838, 0.012486195325463249, 0.100672534509276)

JS = 4  # surface row of the free-surface scheme


def ricker_samples(freq, tc, nt, dt, offset=0.0):
    t = (np.arange(nt) + offset) * dt
    inside = (tc - 2.0 / freq < t) & (t < tc + 2.0 / freq)
    return np.where(inside, _ricker(freq, t - tc), 0.0)


def config_small_elastic(n=256, nt=1000):
    nx = nz = n
    dx = 5.0
    dt = 5e-4
    vp, vs, rho = 2000.0, 1150.0, 2000.0
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    f = 15.0
    tc = 1.0 / f
    rec_ix = np.linspace(8, nx - 9, 64).round().astype(np.int32)
    return dict(
        name=f"elastic_{nx}x{nz}", nx=nx, nz=nz, dx=dx, dz=dx, dt=dt, nt=nt, nrelax=0,
        taun=None,
        model=dict(rho=np.full((nz, nx), rho), lam=np.full((nz, nx), lam),
                   mu=np.full((nz, nx), mu)),
        free_surface=False, cpml=None,
        src_ix=np.array([nx // 2], np.int32), src_iz=np.array([nz // 2], np.int32),
        src_kind=np.array([1], np.int32),
        wavelets=ricker_samples(f, tc, nt, dt, offset=0.5)[None, :],
        rec_ix=rec_ix, rec_iz=np.full(64, 50, np.int32), rec_comp=np.full(64, 3, np.int32),
    )


def layered_model(nx, nz, seed=0, nlayers=8, noise=0.02, q0=50.0, taun=TAUN_2_40HZ_L3,
                  f1=2.0, f2=40.0):
    rng = np.random.default_rng(seed)
    vp_l = rng.uniform(1800.0, 4500.0, nlayers)
    layer = np.minimum((np.arange(nz) * nlayers) // nz, nlayers - 1)
    vp = np.repeat(vp_l[layer][:, None], nx, axis=1)
    vs = vp / 1.8
    rho = 310.0 * vp ** 0.25
    vp = vp * (1.0 + noise * rng.standard_normal((nz, nx)))
    vs = vs * (1.0 + noise * rng.standard_normal((nz, nx)))
    rho = rho * (1.0 + noise * rng.standard_normal((nz, nx)))
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2.0 * mu
    L = len(taun)
    kbar = float(np.mean(K(np.linspace(2 * np.pi * f1, 2 * np.pi * f2, 100), taun)))
    tau = np.full((nz, nx), get_tau(q0, kbar, L))
    return dict(rho=rho, lam=lam, mu=mu, tau=tau), float(vp.max() * np.sqrt(1 + tau.max()))


def config_layered_visco(nx=1024, nz=512, nt=4000, seed=0):
    dx = 5.0
    dt = 4e-4
    model, vmax = layered_model(nx, nz, seed=seed)
    T = 20
    power, R = 2.0, 1e-5
    d0 = -(power + 1) * vmax * np.log(R) / (2 * T * dx)
    f = 12.0
    tc = 1.0 / f
    depth = JS + int(round(10.0 / dx))
    nrec = 201
    rec_ix = np.linspace(T, nx - T - 1, nrec).round().astype(np.int32)
    return dict(
        name=f"visco_L3_{nx}x{nz}", nx=nx, nz=nz, dx=dx, dz=dx, dt=dt, nt=nt, nrelax=3,
        taun=np.array(TAUN_2_40HZ_L3), model=model, free_surface=True,
        cpml=dict(thickness=T, d0=d0, alpha_max=0.0, kappa_max=1.0, power=power, left=True,
                  right=True, top=False, bottom=True),
        src_ix=np.array([nx // 2], np.int32), src_iz=np.array([depth], np.int32),
        src_kind=np.array([0], np.int32), wavelets=ricker_samples(f, tc, nt, dt)[None, :],
        rec_ix=rec_ix, rec_iz=np.full(nrec, depth, np.int32),
        rec_comp=np.full(nrec, 1, np.int32),
    )


def build_simulation(cfg, precision="float32"):
    """Create a configured device Simulation for a config dict."""
    from .stepper import Simulation
    sim = Simulation(cfg["nx"], cfg["nz"], cfg["dx"], cfg["dz"], cfg["dt"], precision,
                     cfg["nrelax"], cfg.get("fd_order", (8, 8)))
    sim.set_model_dict(cfg["model"])
    if cfg["nrelax"]:
        sim.set_relaxation(cfg["taun"])
    if cfg["cpml"]:
        sim.set_cpml(**cfg["cpml"])
    sim.set_free_surface(cfg["free_surface"])
    if cfg.get("afda_points") is not None:
        sim.set_afda_points(cfg["afda_points"])
    if cfg.get("hermitian_points") is not None:
        sim.set_afda_hermitian(cfg["hermitian_points"])
    sim.set_sources(cfg["src_ix"], cfg["src_iz"], cfg["src_kind"], cfg["wavelets"])
    sim.set_receivers(cfg["rec_ix"], cfg["rec_iz"], cfg["rec_comp"], cfg["nt"])
    return sim


def config_multishot(nshots=64, nz=512, wmin=512, wmax=1536, nt=3000, seed=0, model_nx=None,
                     rec_every=8):
    """Batched multi-shot forward: ``nshots`` windows of width U{wmin..wmax} x nz cut from
    one layered viscoelastic model (the single-shot model type), each with its own Ricker
    force source at a seeded position along the surface and receivers every ``rec_every``
    cells 10 m below the surface; L = 3, CPML + free surface, FP32 by default."""
    rng = np.random.default_rng(seed)
    widths = rng.integers(wmin, wmax + 1, nshots)
    W = int(model_nx) if model_nx else max(4 * wmax, 4096)
    model, vmax = layered_model(W, nz, seed=seed)
    dx, dt = 5.0, 4e-4
    T = 20
    power, R = 2.0, 1e-5
    d0 = -(power + 1) * vmax * np.log(R) / (2 * T * dx)
    f = 12.0
    depth = JS + int(round(10.0 / dx))
    wav = ricker_samples(f, 1.0 / f, nt, dt)
    shots = []
    for s in range(nshots):
        w = int(widths[s])
        sx = int(rng.integers(T + 8, W - T - 8))
        x0 = int(np.clip(sx - w // 2, 0, W - w))
        rix = np.arange(T, w - T, rec_every, dtype=np.int32)
        shots.append(dict(nx=w, x0=x0, src_ix=np.array([sx - x0], np.int32),
                          src_iz=np.array([depth], np.int32), src_kind=np.array([0], np.int32),
                          wavelets=wav[None, :], rec_ix=rix,
                          rec_iz=np.full(len(rix), depth, np.int32),
                          rec_comp=np.full(len(rix), 1, np.int32)))
    return dict(name=f"multishot_{nshots}x{nz}", nz=nz, model_nx=W, dx=dx, dz=dx, dt=dt, nt=nt,
                nrelax=3, taun=np.array(TAUN_2_40HZ_L3), model=model, free_surface=True,
                cpml=dict(thickness=T, d0=d0, alpha_max=0.0, kappa_max=1.0, power=power,
                          left=True, right=True, top=False, bottom=True),
                shots=shots)


def shot_config(mcfg, s):
    """The single-shot config dict of shot ``s`` of a multi-shot config (model window
    copied out), for running it alone or on the oracle."""
    sh = mcfg["shots"][s]
    x0, w = sh["x0"], sh["nx"]
    model = {k: np.ascontiguousarray(v[:, x0:x0 + w]) for k, v in mcfg["model"].items()}
    return dict(name=f"shot{s}", nx=w, nz=mcfg["nz"], dx=mcfg["dx"], dz=mcfg["dz"],
                dt=mcfg["dt"], nt=mcfg["nt"], nrelax=mcfg["nrelax"], taun=mcfg["taun"],
                model=model, free_surface=mcfg["free_surface"], cpml=mcfg["cpml"],
                src_ix=sh["src_ix"], src_iz=sh["src_iz"], src_kind=sh["src_kind"],
                wavelets=sh["wavelets"], rec_ix=sh["rec_ix"], rec_iz=sh["rec_iz"],
                rec_comp=sh["rec_comp"])


def build_multishot(mcfg, precision="float32"):
    """Upload the shared model once and create one Simulation per shot reading its window.
    Returns (model, sims)."""
    from .batch import Model
    from .stepper import Simulation
    m = mcfg["model"]
    model = Model(m["rho"], m["lam"], m["mu"], m.get("tau"), precision)
    sims = []
    for sh in mcfg["shots"]:
        sim = Simulation(sh["nx"], mcfg["nz"], mcfg["dx"], mcfg["dz"], mcfg["dt"], precision,
                         mcfg["nrelax"])
        sim.set_model_window(model, sh["x0"])
        if mcfg["nrelax"]:
            sim.set_relaxation(mcfg["taun"])
        if mcfg["cpml"]:
            sim.set_cpml(**mcfg["cpml"])
        sim.set_free_surface(mcfg["free_surface"])
        sim.set_sources(sh["src_ix"], sh["src_iz"], sh["src_kind"], sh["wavelets"])
        sim.set_receivers(sh["rec_ix"], sh["rec_iz"], sh["rec_comp"], mcfg["nt"])
        sims.append(sim)
    return model, sims


def moduli_from_velocity(vp, vs, rho, Q, kbar, L=3):
    """(vp, vs, rho, Q) -> model dict (rho, lam, mu, tau) with lam = rho (vp^2 - 2 vs^2),
    mu = rho vs^2 and tau = L / (Kbar Q) (Models.jl:67)."""
    mu = rho * vs ** 2
    return dict(rho=rho, lam=rho * vp ** 2 - 2.0 * mu, mu=mu, tau=get_tau(Q, kbar, L),
                vp=vp, vs=vs, Q=Q)


def adjoint_models(nx=1600, nz=600, seed=0, nlayers=6, nanom=6, smooth_cells=40.0, L=3,
                   taun=TAUN_2_40HZ_L3, f1=2.0, f2=40.0):
    """True and starting models of the gradient workload: a smooth layered background
    (vp 1800-4000 m/s with tanh interfaces, vs = vp/1.8, Gardner rho) plus ``nanom`` seeded
    Gaussian anomalies of +-10 % in vp and vs with radii U[20, 60] cells; Q ~ U[30, 100] per
    layer. The starting model is the true model smoothed by a ``smooth_cells``-cell Gaussian.
    Returns (true, start, kbar); true and start are dicts of rho, lam, mu, tau, vp, vs, Q."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float64)
    depths = np.sort(rng.uniform(0.1, 0.95, nlayers - 1)) * nz
    vp_l = np.sort(rng.uniform(1800.0, 4000.0, nlayers))
    q_l = rng.uniform(30.0, 100.0, nlayers)
    vp1 = np.full(nz, vp_l[0])
    q1 = np.full(nz, q_l[0])
    for k, dz in enumerate(depths):
        w = 0.5 * (1 + np.tanh((z - dz) / 4.0))
        vp1 = vp1 + (vp_l[k + 1] - vp_l[k]) * w
        q1 = q1 + (q_l[k + 1] - q_l[k]) * w
    vp = np.repeat(vp1[:, None], nx, axis=1)
    vs = vp / 1.8
    Q = np.repeat(q1[:, None], nx, axis=1)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz))
    for _ in range(nanom):
        cx, cz = rng.uniform(0.1, 0.9) * nx, rng.uniform(0.2, 0.9) * nz
        r = rng.uniform(20.0, 60.0)
        amp = rng.choice([-0.1, 0.1])
        g = 1 + amp * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * r * r))
        vp = vp * g
        vs = vs * g
    kbar = float(np.mean(K(np.linspace(2 * np.pi * f1, 2 * np.pi * f2, 100), taun)))

    def pack(vp, vs, Q):
        return moduli_from_velocity(vp, vs, 310.0 * vp ** 0.25, Q, kbar, L)

    true = pack(vp, vs, Q)
    sm = smooth_cells
    start = pack(gaussian_filter(vp, sm, mode="nearest"), gaussian_filter(vs, sm, mode="nearest"),
                 gaussian_filter(Q, sm, mode="nearest"))
    return true, start, kbar


def config_adjoint(nx=1600, nz=600, nshots=16, nrec=301, nt=3000, seed=0, smooth_cells=40.0):
    """Gradient workload: nx x nz (default 1600 x 600), dx 5 m, L = 3, CPML 20 + free
    surface, ``nshots`` 12 Hz Ricker force sources evenly along the surface 10 m deep,
    ``nrec`` receivers (vz) 10 m deep, ``nt`` steps at dt 0.4 ms; L2 trace misfit of the
    starting model against data of the true model."""
    true, start, kbar = adjoint_models(nx, nz, seed, smooth_cells=smooth_cells)
    dx, dt, T = 5.0, 4e-4, 20
    vmax = float(true["vp"].max() * np.sqrt(1 + true["tau"].max()))
    d0 = -(2.0 + 1) * vmax * np.log(1e-5) / (2 * T * dx)
    depth = JS + int(round(10.0 / dx))
    src_x = np.linspace(T + 10, nx - T - 11, nshots).round().astype(np.int32)
    rec_x = np.linspace(T, nx - T - 1, nrec).round().astype(np.int32)
    wav = ricker_samples(12.0, 1.0 / 12.0, nt, dt)
    return dict(name=f"adjoint_{nx}x{nz}", nx=nx, nz=nz, dx=dx, dz=dx, dt=dt, nt=nt, nrelax=3,
                taun=np.array(TAUN_2_40HZ_L3), kbar=kbar, true=true, start=start,
                free_surface=True,
                cpml=dict(thickness=T, d0=d0, alpha_max=0.0, kappa_max=1.0, power=2.0,
                          left=True, right=True, top=False, bottom=True),
                src_x=src_x, src_z=depth, wavelet=wav, rec_ix=rec_x,
                rec_iz=np.full(nrec, depth, np.int32), rec_comp=np.full(nrec, 1, np.int32))


def adjoint_shot(acfg, s, model):
    """Single-shot config dict of shot ``s`` of a gradient workload in ``model`` (a dict with
    rho, lam, mu, tau, or the name of one of the workload's models)."""
    m = acfg[model] if isinstance(model, str) else model
    return dict(name=f"{acfg['name']}_shot{s}", nx=acfg["nx"], nz=acfg["nz"], dx=acfg["dx"],
                dz=acfg["dz"], dt=acfg["dt"], nt=acfg["nt"], nrelax=acfg["nrelax"],
                taun=acfg["taun"], model={k: m[k] for k in ("rho", "lam", "mu", "tau")},
                free_surface=acfg["free_surface"], cpml=acfg["cpml"],
                src_ix=np.array([acfg["src_x"][s]], np.int32),
                src_iz=np.array([acfg["src_z"]], np.int32), src_kind=np.array([0], np.int32),
                wavelets=acfg["wavelet"][None, :], rec_ix=acfg["rec_ix"], rec_iz=acfg["rec_iz"],
                rec_comp=acfg["rec_comp"])


def observed_data(acfg, precision="float32", model="true"):
    """Traces [nrec, nt] of every shot in ``model`` (default: the true model)."""
    out = []
    for s in range(len(acfg["src_x"])):
        sim = build_simulation(adjoint_shot(acfg, s, model), precision)
        sim.step(acfg["nt"])
        out.append(sim.traces(acfg["nt"]))
        sim.close()
    return out


def multishot_misfit(acfg, model, dobs, precision="float32"):
    """Sum over shots of 1/2 |d(model) - dobs|^2."""
    syn = observed_data(acfg, precision, model)
    return float(sum(0.5 * np.sum((a.astype(np.float64) - b) ** 2) for a, b in zip(syn, dobs)))


def multishot_gradient(acfg, dobs, precision="float32", model="start", checkpoint_every=0,
                       concurrent=1, chunk=100, sims=None):
    """Misfit and its gradient with respect to (vp, vs, rho, Q), summed over the shots,
    from the discrete adjoint of each shot.

    ``concurrent`` shots are computed at once, each as a chunked gradient job on a simulation
    (and HIP stream) of its own whose source is moved from shot to shot; ``chunk`` is the
    number of steps a job enqueues per turn. The launches of two jobs fill each other's
    launch gaps: at 1600 x 600, 3000 steps, two jobs gave 10-12 % more shots per second than
    one, four fewer (measured). The result does not depend on ``concurrent``: every shot's
    gradient is computed alone and the sums run in shot order. ``sims``: simulations of this
    workload's model to run the jobs on instead (one job each; their sources are moved and
    they stay open, with their gradient workspaces)."""
    import time
    from .api import gradient_to_velocity_params
    from .stepper import GradientJob
    m = acfg[model] if isinstance(model, str) else model
    nshots, nt = len(acfg["src_x"]), acfg["nt"]
    own = sims is None
    if own:
        n = max(1, min(int(concurrent), nshots))
        sims = [build_simulation(adjoint_shot(acfg, s, m), precision) for s in range(n)]
    res = [None] * nshots
    queue, active = list(range(nshots)), {}
    try:
        while queue or active:
            for k, sim in enumerate(sims):
                if k not in active and queue:
                    s = queue.pop(0)
                    c = adjoint_shot(acfg, s, m)
                    sim.set_sources(c["src_ix"], c["src_iz"], c["src_kind"], c["wavelets"])
                    active[k] = (s, GradientJob(sim, dobs[s], nt, checkpoint_every))
            for k in list(active):
                s, job = active[k]
                if not job.finished:
                    job.advance(chunk)
                elif job.idle():
                    res[s] = job.end()
                    del active[k]
            if active:
                time.sleep(2e-4)  # the jobs' work is enqueued; wait for the device
    finally:
        for _, job in active.values():
            job.cancel()
        if own:
            for sim in sims:
                sim.close()
    J = 0.0
    total = None
    for Js, g in res:
        gv = gradient_to_velocity_params(g, m["rho"], m["lam"], m["mu"], m["tau"], m["Q"])
        J += Js
        total = gv if total is None else {k: total[k] + gv[k] for k in gv}
    return J, total


# grid sizes (nx, nz) of the serving workload: 256^2, 512 x 1024, 1024 x 2048, 2048^2
SERVING_GRIDS = ((256, 256), (512, 1024), (1024, 2048), (2048, 2048))


def serving_job_config(nx, nz, nt, seed, vti=False, eps=0.0, delta=0.0, rec_every=8,
                       dtype=np.float32):
    """One shot of the serving workload: seeded 8-layer model (as the single-shot workload),
    isotropic or VTI with the given Thomsen eps, delta; L = 3, Q = 50, CPML 20 + free surface,
    12 Hz Ricker force source 10 m deep at a seeded x, receivers every ``rec_every`` columns
    10 m deep; dx 5 m, dt 0.4 ms. Model arrays are stored in ``dtype``."""
    from .api import thomsen_to_stiffness
    rng = np.random.default_rng(seed + 7919)
    model, vmax = layered_model(nx, nz, seed=seed)
    if vti:
        vp = np.sqrt((model["lam"] + 2 * model["mu"]) / model["rho"])
        vs = np.sqrt(model["mu"] / model["rho"])
        c11, c13, c33, c55 = thomsen_to_stiffness(vp, vs, model["rho"], eps, delta)
        model = dict(rho=model["rho"], c11=c11, c13=c13, c33=c33, c55=c55, tau=model["tau"])
        vmax *= np.sqrt(1 + 2 * max(eps, 0.0))
    model = {k: np.ascontiguousarray(v, dtype=dtype) for k, v in model.items()}
    dx, dt, T = 5.0, 4e-4, 20
    d0 = -(2.0 + 1) * vmax * np.log(1e-5) / (2 * T * dx)
    depth = JS + int(round(10.0 / dx))
    rec_ix = np.arange(T, nx - T, rec_every, dtype=np.int32)
    return dict(name=f"serve_{nx}x{nz}", nx=nx, nz=nz, dx=dx, dz=dx, dt=dt, nt=int(nt), nrelax=3,
                taun=np.array(TAUN_2_40HZ_L3), model=model, free_surface=True,
                cpml=dict(thickness=T, d0=d0, alpha_max=0.0, kappa_max=1.0, power=2.0,
                          left=True, right=True, top=False, bottom=True),
                src_ix=np.array([int(rng.integers(T + 10, nx - T - 10))], np.int32),
                src_iz=np.array([depth], np.int32), src_kind=np.array([0], np.int32),
                wavelets=ricker_samples(12.0, 1.0 / 12.0, int(nt), dt)[None, :],
                rec_ix=rec_ix, rec_iz=np.full(len(rec_ix), depth, np.int32),
                rec_comp=np.full(len(rec_ix), 1, np.int32))


def serving_workload(njobs, rate, seed=0, grids=SERVING_GRIDS, steps=(1000, 4000),
                     gradient_fraction=0.3, vti_fraction=0.5):
    """Seeded stream of shot jobs: Poisson arrivals at ``rate`` per second, 30 % gradient
    and 70 % forward, grid drawn uniformly from ``grids``, isotropic or VTI
    (eps ~ U[0, 0.25], delta ~ U[-0.1, 0.15]), L = 3, FP32, steps ~ U[steps]. Gradient jobs
    carry zero observed data, so their misfit is 1/2 |d|^2 of the synthetic traces."""
    from .serve import Job
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, njobs))
    jobs = []
    for k in range(njobs):
        nx, nz = grids[int(rng.integers(len(grids)))]
        nt = int(rng.integers(steps[0], steps[1] + 1))
        vti = bool(rng.random() < vti_fraction)
        eps, delta = float(rng.uniform(0.0, 0.25)), float(rng.uniform(-0.1, 0.15))
        kind = "gradient" if rng.random() < gradient_fraction else "forward"
        cfg = serving_job_config(nx, nz, nt, seed=seed * 100003 + k, vti=vti, eps=eps, delta=delta)
        dobs = np.zeros((len(cfg["rec_ix"]), nt), np.float32) if kind == "gradient" else None
        jobs.append(Job(id=k, kind=kind, cfg=cfg, arrival=float(t[k]), dobs=dobs))
    return jobs
