"""Drop-in boundary behaviour on the GPU:

* leapfrog_step!(fields, parameters, dt) takes the model from ``parameters``
  (visco-kernels.jl:87-98 reads λs, μs, τs, τns from it): stepping with new parameters equals
  a simulation built with that model, and the old model is not used;
* a handle that a gradient job owns is refused by the batch entry points (status VI_ESTATE).
"""
import numpy as np
import pytest

from helpers import small_cfg

pytestmark = pytest.mark.gpu


def _stepped(cfg, nt, prec="float64"):
    from viscoelasticinversion_amd.configs import build_simulation
    sim = build_simulation(cfg, prec)
    sim.step(nt)
    out = {k: sim.get(k) for k in ("vx", "vz", "sxx", "szz", "sxz", "mxx")}
    out["traces"] = sim.traces(nt)
    sim.close()
    return out


def test_leapfrog_step_uses_the_given_parameters():
    from viscoelasticinversion_amd.api import Fields, StepParameters, leapfrog_step
    from viscoelasticinversion_amd.configs import build_simulation
    cfg_a = small_cfg(nx=90, nz=60, nrelax=3, seed=3, nt=30)
    cfg_b = small_cfg(nx=90, nz=60, nrelax=3, seed=4, nt=30)  # another model, same sources
    mb = cfg_b["model"]
    print("[case] leapfrog_step with changed parameters", flush=True)
    sim = build_simulation(cfg_a, "float64")
    fields = Fields(sim)
    p = StepParameters(lam=mb["lam"], mu=mb["mu"], tau=mb["tau"], taun=cfg_b["taun"] * 1.0,
                       rho=mb["rho"])
    for _ in range(30):
        leapfrog_step(fields, p, cfg_a["dt"])
    got = {k: sim.get(k) for k in ("vx", "vz", "sxx", "szz", "sxz", "mxx")}
    got["traces"] = sim.traces(30)
    sim.close()
    want = _stepped(cfg_b, 30)
    old = _stepped(cfg_a, 30)
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    assert not np.array_equal(got["vx"], old["vx"])


def test_leapfrog_step_parameters_object_and_rho_kept():
    """A Parameters (IsotropicModel + QModel) switches the model too; a StepParameters
    without rho keeps the loaded density."""
    from viscoelasticinversion_amd.api import (Fields, IsotropicModel, Parameters, QModel,
                                               StepParameters, leapfrog_step)
    from viscoelasticinversion_amd.configs import build_simulation
    cfg_a = small_cfg(nx=80, nz=50, nrelax=2, cpml=False, fs=False, seed=5, nt=20)
    cfg_b = small_cfg(nx=80, nz=50, nrelax=2, cpml=False, fs=False, seed=6, nt=20)
    mb = cfg_b["model"]
    q = QModel(2.0, 40.0, 80, 50, Q0=50, N=2, taun=cfg_b["taun"])
    q.tau = mb["tau"]
    p = Parameters(Nx=80, Nz=50, dx=cfg_a["dx"], dz=cfg_a["dz"], dt=cfg_a["dt"],
                   model=IsotropicModel(mb["rho"], mb["lam"], mb["mu"], 80, 50), qmodel=q,
                   precision="float64")
    sim = build_simulation(cfg_a, "float64")
    leapfrog_step(Fields(sim), p, cfg_a["dt"], nsteps=20)
    got = sim.get("vz")
    sim.close()
    assert np.array_equal(got, _stepped(cfg_b, 20)["vz"])
    # density kept: lam, mu, tau of b with rho of a
    cfg_c = dict(cfg_b)
    cfg_c["model"] = dict(mb, rho=cfg_a["model"]["rho"])
    sim = build_simulation(cfg_a, "float64")
    leapfrog_step(Fields(sim), StepParameters(lam=mb["lam"], mu=mb["mu"], tau=mb["tau"],
                                              taun=cfg_b["taun"]), nsteps=20)
    got = sim.get("vz")
    sim.close()
    assert np.array_equal(got, _stepped(cfg_c, 20)["vz"])


def test_batch_refuses_a_handle_owned_by_a_gradient_job():
    from viscoelasticinversion_amd import Batch, _lib
    from viscoelasticinversion_amd.configs import build_simulation
    from viscoelasticinversion_amd.stepper import GradientJob
    cfg = small_cfg(nx=70, nz=50, nrelax=1, seed=9, nt=30)
    a = build_simulation(cfg, "float32")
    other = build_simulation(cfg, "float32")
    print("[case] batch with a gradient-owned member", flush=True)
    b = Batch([a, other])  # members set before the job starts
    job = GradientJob(a, np.zeros((len(cfg["rec_ix"]), 30), np.float32), 30)
    with pytest.raises(_lib.ViError) as e1:
        b.step(1)
    assert e1.value.code == _lib.VI_ESTATE
    with pytest.raises(_lib.ViError) as e2:
        b.set([a, other])
    assert e2.value.code == _lib.VI_ESTATE
    with pytest.raises(_lib.ViError) as e3:
        Batch([other, a])
    assert e3.value.code == _lib.VI_ESTATE
    job.cancel()
    b.set([a, other])  # free again
    b.step(2)
    b.sync()
    b.close()
    a.close()
    other.close()
