"""Serving loop on the GPU: a stream of forward and gradient jobs of different grid sizes,
isotropic and VTI, arriving faster than they are served under an HBM budget that forces
queueing. Every job's result must be the same as the job run alone with the plain stepper
(bit for bit: the batched launch and the adjoint are deterministic), the budget must hold,
and jobs must overlap on the device."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_gradients", [1, 2])
def test_served_results_equal_alone_under_budget(max_gradients):
    from viscoelasticinversion_amd.configs import serving_workload
    from viscoelasticinversion_amd.serve import Server, job_bytes, run_serial
    grids = ((96, 64), (160, 96), (200, 120), (120, 150))
    jobs = serving_workload(16, rate=400.0, seed=5, grids=grids, steps=(60, 180))
    kinds = [j.kind for j in jobs]
    print(f"[case] serve 16 jobs: {kinds.count('gradient')} gradient, "
          f"max_gradients {max_gradients}", flush=True)
    assert "gradient" in kinds and "forward" in kinds
    need = [job_bytes(j) for j in jobs]
    budget = int(2.5 * max(need))
    srv = Server(hbm_budget=budget, block_steps=25, max_gradients=max_gradients,
                 max_head_wait=0.05)
    events = []
    res = srv.run(jobs, on_event=lambda w, j, t: events.append((t, w, j)))
    assert all(r.error is None for r in res), [r.error for r in res]
    assert srv.peak_bytes <= budget
    # jobs overlapped on the device, and queueing happened (not all admitted on arrival)
    live, most = 0, 0
    for t, w, j in sorted(events, key=lambda e: (e[0], e[1] == "admit")):
        live += 1 if w == "admit" else -1
        most = max(most, live)
    assert most >= 2
    assert any(r.admitted - r.arrival > 0.01 for r in res)
    alone = run_serial(jobs)
    for r, a in zip(res, alone):
        assert r.id == a.id
        if r.kind == "forward":
            assert np.array_equal(r.traces, a.traces), r.id
            assert np.abs(r.traces).max() > 0
        else:
            assert r.misfit == a.misfit, r.id
            for k in a.gradient:
                assert np.array_equal(r.gradient[k], a.gradient[k]), (r.id, k)
    print(f"  peak {srv.peak_bytes / 1e6:.1f} MB of {budget / 1e6:.1f} MB, "
          f"max concurrent {most}", flush=True)


@pytest.mark.parametrize("vti", [False, True])
def test_plan_bounds_device_bytes(vti):
    from viscoelasticinversion_amd.configs import build_simulation, serving_job_config
    from viscoelasticinversion_amd.serve import Job, job_bytes
    cfg = serving_job_config(300, 200, 120, seed=9, vti=vti, eps=0.1, delta=0.05)
    sim = build_simulation(cfg, "float32")
    sim.step(3)
    sim.sync()
    base = sim.device_bytes()
    grad = sim.gradient_bytes(120)
    sim.close()
    fwd_plan = job_bytes(Job(0, "forward", cfg))
    grad_plan = job_bytes(Job(1, "gradient", cfg))
    print(f"[case] plan vti={vti}: device {base} plan {fwd_plan}; gradient {grad} "
          f"plan {grad_plan - fwd_plan}", flush=True)
    assert base <= fwd_plan <= base + 64 * 1024
    assert grad_plan - fwd_plan == grad


def test_blown_up_jobs_fail_alone():
    """A forward job and a gradient job whose dt exceeds the stability limit (the pre-check
    switched off) blow up: each fails alone with an error, every other job's result equals the
    job run alone."""
    from viscoelasticinversion_amd.api import stable_dt
    from viscoelasticinversion_amd.configs import serving_job_config
    from viscoelasticinversion_amd.serve import Job, Server, run_serial
    jobs = []
    for k, (kind, bad) in enumerate((("forward", False), ("forward", True), ("gradient", False),
                                     ("gradient", True), ("forward", False))):
        cfg = serving_job_config(96 + 16 * k, 72, 300, seed=40 + k)
        if bad:
            cfg["dt"] = 1.6 * stable_dt(cfg["model"], cfg["dx"], cfg["dz"])
        dobs = np.zeros((len(cfg["rec_ix"]), 300), np.float32) if kind == "gradient" else None
        jobs.append(Job(id=k, kind=kind, cfg=cfg, arrival=0.01 * k, dobs=dobs))
    print("[case] serve 5 jobs, 2 unstable", flush=True)
    res = Server(block_steps=25, check_cfl=False).run(jobs)
    alone = run_serial(jobs, check_cfl=False)
    for r, a in zip(res, alone):
        if r.id in (1, 3):
            assert r.error and "blew up" in r.error, r.error
            assert a.error and "blew up" in a.error, a.error
            continue
        assert r.error is None and a.error is None
        if r.kind == "forward":
            assert np.array_equal(r.traces, a.traces)
        else:
            assert r.misfit == a.misfit and np.isfinite(r.misfit)


def test_unstable_jobs_refused_on_arrival():
    from viscoelasticinversion_amd.api import stable_dt
    from viscoelasticinversion_amd.configs import serving_job_config
    from viscoelasticinversion_amd.serve import Job, Server
    good = serving_job_config(96, 72, 60, seed=1)
    bad = serving_job_config(96, 72, 60, seed=2)
    bad["dt"] = 1.05 * stable_dt(bad["model"], bad["dx"], bad["dz"])
    res = Server(block_steps=20).run([Job(0, "forward", good), Job(1, "forward", bad, 0.001)])
    assert res[0].error is None and np.abs(res[0].traces).max() > 0
    assert res[1].error and "stability limit" in res[1].error and res[1].traces is None
