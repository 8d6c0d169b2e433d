"""Host-side parts of the serving loop (no device needed): the admission policy, the
device-memory planner that admission uses, the workload generator and the summary."""
import numpy as np

from viscoelasticinversion_amd.serve import JobResult, select_admissions, summarize
from viscoelasticinversion_amd.stepper import plan_bytes


def test_admission_fifo_within_budget():
    w = [(0, "forward", 5, 0.0), (1, "gradient", 50, 0.0), (2, "forward", 10, 0.0),
         (3, "forward", 4, 0.0)]
    assert select_admissions(w, 0, 60, 0, 2, 0, 10) == ([0, 1, 3], [])


def test_admission_backfill_until_the_blocked_job_has_waited():
    w = [(1, "forward", 60, 0.0), (2, "forward", 15, 0.1), (3, "forward", 4, 0.2)]
    # job 1 does not fit next to 50 used bytes; the small ones overtake it for a while ...
    assert select_admissions(w, 50, 100, 0, 1, 0, 10, now=0.5, max_head_wait=1.0) == ([2, 3], [])
    # ... but not once it has waited max_head_wait
    assert select_admissions(w, 50, 100, 0, 1, 0, 10, now=1.0, max_head_wait=1.0) == ([], [])


def test_admission_limits_and_rejection():
    w = [(0, "gradient", 1, 0.0), (1, "gradient", 1, 0.0), (2, "forward", 1, 0.0),
         (3, "forward", 1000, 0.0), (4, "forward", 1, 0.0)]
    picks, rej = select_admissions(w, 0, 100, 1, 2, 0, 1)
    assert picks == [0, 2] and rej == [3]  # one free gradient slot, one free batch slot


def test_plan_bytes_components():
    base = dict(nx=300, nz=200, precision="float32", nrelax=3, cpml_thickness=20, nsrc=1,
                src_nt=1000, nrec=40, rec_nt=1000)
    iso = plan_bytes(**base)
    vti = plan_bytes(**base, vti=True)
    ldx = (300 + 72 + 63) // 64 * 64
    pb = ldx * (200 + 48) * 4
    assert vti - iso == 2 * pb  # C11 and C33 planes
    assert plan_bytes(**dict(base, nrelax=0)) == iso - (9 + 1) * pb  # M planes and tau
    assert plan_bytes(**dict(base, precision="float64")) > 1.9 * iso
    g1 = plan_bytes(**base, grad_nt=1000) - iso
    long = dict(base, rec_nt=4000, src_nt=4000)
    g4 = plan_bytes(**long, grad_nt=4000) - plan_bytes(**long)
    # sqrt(nt) checkpointing: about 2 sqrt(nt) states, so 4x the steps ~ 2x the memory
    assert 1.6 < g4 / g1 < 2.4
    # checkpoint every step: nt states plus one slot per step
    assert plan_bytes(**base, grad_nt=1000, checkpoint_every=1) > 10 * g1


def test_serving_workload_stream():
    from viscoelasticinversion_amd.configs import serving_workload
    grids = ((64, 48), (96, 40))
    jobs = serving_workload(300, rate=50.0, seed=3, grids=grids, steps=(10, 40))
    kinds = np.array([j.kind == "gradient" for j in jobs])
    assert abs(kinds.mean() - 0.3) < 0.1
    gaps = np.diff([0.0] + [j.arrival for j in jobs])
    assert np.all(gaps > 0) and abs(gaps.mean() * 50.0 - 1) < 0.2
    assert {(j.cfg["nx"], j.cfg["nz"]) for j in jobs} == set(grids)
    vti = [j for j in jobs if "c11" in j.cfg["model"]]
    assert 0.35 < len(vti) / len(jobs) < 0.65
    for j in jobs:
        assert 10 <= j.cfg["nt"] <= 40 and j.cfg["nrelax"] == 3
        if j.kind == "gradient":
            assert j.dobs.shape == (len(j.cfg["rec_ix"]), j.cfg["nt"])
    again = serving_workload(300, rate=50.0, seed=3, grids=grids, steps=(10, 40))
    assert [j.arrival for j in again] == [j.arrival for j in jobs]


def test_summary():
    rs = [JobResult(k, "forward", arrival=float(k), done=float(k) + 1 + k % 3, cells=100,
                    steps=10) for k in range(10)]
    s = summarize(rs)
    assert s["jobs"] == 10 and s["wall_s"] == 11.0
    assert abs(s["shots_per_s"] - 10 / 11) < 1e-3
    assert s["latency_p50_s"] == 2.0 and s["latency_p99_s"] > 2.9


def test_stability_check_and_checkpoint_policy():
    from viscoelasticinversion_amd.api import stable_dt
    from viscoelasticinversion_amd.configs import serving_job_config
    from viscoelasticinversion_amd.serve import (Job, gradient_checkpoint, job_bytes,
                                                 stability_error)
    cfg = serving_job_config(256, 256, 1000, seed=3, vti=True, eps=0.25, delta=0.1)
    job = Job(0, "gradient", cfg)
    assert stability_error(job) is None
    lim = stable_dt(cfg["model"], cfg["dx"], cfg["dz"])
    fast = dict(cfg, dt=1.01 * lim)
    assert "stability limit" in stability_error(Job(1, "forward", fast))
    full = job_bytes(Job(0, "gradient", cfg, checkpoint_every=1000))
    assert full > job_bytes(job)  # storing every step needs more than sqrt(nt) checkpoints
    assert gradient_checkpoint(job, 4 * full) == 1000
    # just below: two segments of 500 steps (half the forward recomputed) fit
    half = job_bytes(Job(0, "gradient", cfg, checkpoint_every=500))
    assert half < full
    assert gradient_checkpoint(job, 4 * full - 8) == 500
    # no interval ceil(nt / s) fits: ceil(sqrt(nt)) checkpoints
    small = min(job_bytes(Job(0, "gradient", cfg, checkpoint_every=-(-1000 // s)))
                for s in range(1, 33))
    assert gradient_checkpoint(job, 4 * small - 8) == 0
    assert gradient_checkpoint(Job(0, "forward", cfg), 10 ** 15) == 0


def test_largest_checkpoint_interval():
    """The largest interval ceil(nt / s) whose bytes fit, against a brute-force search over
    every interval of that form; 0 when none fits."""
    from viscoelasticinversion_amd.stepper import largest_checkpoint_interval
    for nt in (1, 2, 7, 100, 1000, 3000, 4001):
        def cost(K, nt=nt):
            return -(-nt // K) * 300 + K * 41  # checkpoints + slots
        cands = sorted({-(-nt // s) for s in range(1, int(np.ceil(np.sqrt(nt))) + 1)})
        for budget in (0, 350, 2000, 20000, 10 ** 9):
            fit = [K for K in cands if cost(K) <= budget]
            assert largest_checkpoint_interval(nt, cost, budget) == (max(fit) if fit else 0)
