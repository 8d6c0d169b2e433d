"""The rrule analogue (misfit_and_pullback) through the reference-style API, and the chain
rules to velocities / Thomsen parameters, against central finite differences of the misfit
(FP64). Parameters, IsotropicModel, QModel, CPMLConfig and RickerSource are the
host mirror of the reference's constructors; the forward map is the GPU stepper."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(model, qmodel, receivers):
    from viscoelasticinversion_amd import CPMLConfig, Parameters, RickerSource
    return Parameters(Nx=90, Nz=70, dx=5.0, dz=5.0, dt=4e-4, precision="float64", model=model,
                      qmodel=qmodel, free_surface=True,
                      cpml=CPMLConfig(thickness=10, vmax=3500.0, dx=5.0, top=False),
                      source=RickerSource(25.0, 0.04, 45, 8), receivers=receivers)


def _misfit(p, dobs, tspan):
    from viscoelasticinversion_amd import make_problem, solve_problem
    tr = solve_problem(make_problem(p, tspan=tspan)).traces
    return 0.5 * float(np.sum((tr - dobs) ** 2))


def test_pullback_isotropic_vp_vs_rho_q():
    from viscoelasticinversion_amd import IsotropicModel, QModel, misfit_and_pullback
    from viscoelasticinversion_amd.configs import TAUN_2_40HZ_L3
    rng = np.random.default_rng(3)
    nx, nz = 90, 70
    vp = 2500.0 + 300.0 * rng.random((nz, nx))
    vs = vp / 1.8
    rho = 2000.0 + 200.0 * rng.random((nz, nx))
    Q = 40.0 + 20.0 * rng.random((nz, nx))
    rec = (np.arange(12, 80, 6), np.full(12, 6), np.full(12, 1))
    tspan = (0.0, 0.06)

    def build(vp, vs, rho, Q):
        q = QModel(2.0, 40.0, nx, nz, Q0=50, N=3, taun=TAUN_2_40HZ_L3)
        q.update(Q)
        return _params(IsotropicModel.from_velocities(vp, vs, rho, nx, nz), q, rec)

    from viscoelasticinversion_amd import make_problem, solve_problem
    dobs = solve_problem(make_problem(build(vp * 1.02, vs, rho, Q), tspan=tspan)).traces
    print("[case] pullback isotropic", flush=True)
    J, pb = misfit_and_pullback(build(vp, vs, rho, Q), dobs, tspan)
    assert abs(J - _misfit(build(vp, vs, rho, Q), dobs, tspan)) <= 1e-12 * J
    g = pb(1.0)
    assert set(g) == {"vp", "vs", "rho", "Q"}
    base = dict(vp=vp, vs=vs, rho=rho, Q=Q)
    for key in ("vp", "vs", "rho", "Q"):
        dm = np.zeros((nz, nx))
        dm[4:-4, 4:-4] = rng.standard_normal((nz - 8, nx - 8))
        eps = 1e-6
        plus = dict(base); plus[key] = base[key] * (1 + eps * dm)
        minus = dict(base); minus[key] = base[key] * (1 - eps * dm)
        fd = (_misfit(build(**plus), dobs, tspan) - _misfit(build(**minus), dobs, tspan)) / (2 * eps)
        ad = float(np.sum(g[key] * base[key] * dm))
        print(f"  {key}: pullback {ad:.9e} fd {fd:.9e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-14 * J, (key, ad, fd)


def test_thomsen_chain_rule():
    from viscoelasticinversion_amd import gradient_to_thomsen_params, thomsen_to_stiffness
    from viscoelasticinversion_amd.configs import build_simulation
    from helpers import small_cfg
    cfg = small_cfg(nx=72, nz=56, nrelax=0, cpml=True, fs=True, seed=61, nt=100)
    rng = np.random.default_rng(62)
    nz, nx = cfg["nz"], cfg["nx"]
    vp0 = 2200.0 + 400.0 * rng.random((nz, nx))
    vs0 = vp0 / (1.7 + 0.2 * rng.random((nz, nx)))
    rho = 1900.0 + 300.0 * rng.random((nz, nx))
    eps = 0.25 * rng.random((nz, nx))
    delta = -0.1 + 0.25 * rng.random((nz, nx))
    th = dict(vp0=vp0, vs0=vs0, rho=rho, eps=eps, delta=delta)

    def model(t):
        c11, c13, c33, c55 = thomsen_to_stiffness(t["vp0"], t["vs0"], t["rho"], t["eps"], t["delta"])
        return dict(rho=t["rho"], c11=c11, c13=c13, c33=c33, c55=c55)

    def run(t):
        sim = build_simulation(dict(cfg, model=model(t)), "float64")
        sim.step(cfg["nt"])
        tr = sim.traces(cfg["nt"])
        sim.close()
        return tr

    dobs = run(dict(th, vp0=vp0 * 1.02, eps=eps * 0.9))
    print("[case] thomsen chain rule", flush=True)
    sim = build_simulation(dict(cfg, model=model(th)), "float64")
    J, g = sim.gradient(dobs, cfg["nt"])
    sim.close()
    gt = gradient_to_thomsen_params(g, rho, vp0, vs0, eps, delta)
    for key in ("vp0", "vs0", "rho", "eps", "delta"):
        dm = np.zeros((nz, nx))
        dm[4:-4, 4:-4] = rng.standard_normal((nz - 8, nx - 8))
        h = 1e-6
        scale = th[key] if key != "delta" else np.ones((nz, nx)) * 0.1
        plus = dict(th); plus[key] = th[key] + h * scale * dm
        minus = dict(th); minus[key] = th[key] - h * scale * dm
        fd = (0.5 * np.sum((run(plus) - dobs) ** 2) - 0.5 * np.sum((run(minus) - dobs) ** 2)) / (2 * h)
        ad = float(np.sum(gt[key] * scale * dm))
        print(f"  {key}: chain {ad:.9e} fd {fd:.9e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-14 * J, (key, ad, fd)
