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
                     cfg["nrelax"])
    m = cfg["model"]
    sim.set_model(m["rho"], m["lam"], m["mu"], m.get("tau"))
    if cfg["nrelax"]:
        sim.set_relaxation(cfg["taun"])
    if cfg["cpml"]:
        sim.set_cpml(**cfg["cpml"])
    sim.set_free_surface(cfg["free_surface"])
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
