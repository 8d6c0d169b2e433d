"""Serving loop: continuous batching of forward and gradient shot jobs on one GPU.

Jobs (one shot each: grid, isotropic or VTI model, sources, receivers, number of steps)
arrive over time. The server keeps every admitted job's model and wavefield resident on the
device and

* admits arrived jobs in arrival order while their device memory fits an HBM budget (the
  library plans each job's bytes before it is created; a job that does not fit keeps its
  place and reserves its bytes, later jobs may only use what is left besides that
  reservation, so large jobs are not starved);
* advances all admitted forward jobs together in one batched launch per half step (the
  batched stepper: shots of any size, isotropic and VTI mixed, one launch), in blocks of
  ``block_steps`` steps replayed from a HIP graph, cut short when a job reaches its last
  step; jobs that arrive meanwhile join at the next block and finished jobs leave;
* runs gradient jobs (forward with checkpoints, recomputation, adjoint) as chunked jobs on
  their own streams: between two batched blocks the loop enqueues the next chunk of every
  gradient whose previous chunk has finished, so gradients overlap the batched forward
  work on the device. All device calls come from the one thread that runs the loop.

Every job's result is the one it gets when run alone (the batched launch advances each shot
exactly as its own launch does, and the adjoint is deterministic). The reference has no
serving layer; its single-shot entry point is solve_problem / leapfrog_step!
(src/sciml_solver/ViscoelasticProblem.jl:307-309, src/visco-kernels.jl:87-100), which
``run_serial`` uses one job at a time as the baseline.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Job:
    """One shot job. ``cfg`` is a shot configuration (nx, nz, dx, dz, dt, nt, nrelax, taun,
    model, free_surface, cpml, sources, receivers); ``arrival`` in seconds from the start of
    the stream; gradient jobs carry observed traces ``dobs`` (nrec, nt)."""
    id: int
    kind: str  # "forward" or "gradient"
    cfg: dict
    arrival: float = 0.0
    dobs: Optional[np.ndarray] = None
    checkpoint_every: int = 0


@dataclass
class JobResult:
    id: int
    kind: str
    arrival: float
    admitted: float = float("nan")
    done: float = float("nan")
    bytes: int = 0
    traces: Optional[np.ndarray] = None
    misfit: Optional[float] = None
    gradient: Optional[dict] = None
    error: Optional[str] = None
    cells: int = 0
    steps: int = 0

    @property
    def latency(self):
        return self.done - self.arrival


def job_bytes(job: Job, precision="float32") -> int:
    """Device bytes the job holds while admitted (planned by the library)."""
    from .stepper import plan_bytes
    c = job.cfg
    cp = c.get("cpml") or {}
    return plan_bytes(c["nx"], c["nz"], precision, c["nrelax"], "c11" in c["model"],
                      int(cp.get("thickness", 0)), len(c["src_ix"]),
                      np.asarray(c["wavelets"]).shape[-1], len(c["rec_ix"]), c["nt"],
                      c["nt"] if job.kind == "gradient" else 0, job.checkpoint_every)


def stability_error(job):
    """Why the job's time step cannot be stable (None when it can): dt above the von Neumann
    bound of the staggered leapfrog (the job's operator orders) for its fastest unrelaxed P
    wave."""
    from .api import stable_dt
    c = job.cfg
    order = tuple(c.get("fd_order", (8, 8)))
    lim = stable_dt(c["model"], c["dx"], c["dz"], order)
    if c["dt"] > lim:
        return (f"dt = {c['dt']:.4g} s exceeds the stability limit {lim:.4g} s of this model "
                f"and grid (staggered leapfrog of orders {order[0]}, {order[1]})")
    return None


def gradient_checkpoint(job, budget, precision="float32", fraction=0.25):
    """Checkpoint interval of a gradient job that does not set one: the largest interval
    whose plan needs at most ``fraction`` of the HBM budget (nt when every step can be
    recorded in the forward pass: nothing recomputed, ~25 % faster than ceil(sqrt(nt))
    checkpoints; fewer, longer segments otherwise: less to recompute), else 0
    (ceil(sqrt(nt)) checkpoints). The gradient does not depend on it."""
    from .stepper import largest_checkpoint_interval
    if job.kind != "gradient" or job.checkpoint_every:
        return job.checkpoint_every
    nt = job.cfg["nt"]
    return largest_checkpoint_interval(
        nt, lambda K: job_bytes(Job(job.id, job.kind, job.cfg, checkpoint_every=K), precision),
        fraction * budget)


def _with_checkpoint(job, K):
    if K == job.checkpoint_every:
        return job
    return Job(job.id, job.kind, job.cfg, job.arrival, job.dobs, K)


def _build(cfg, precision):
    from .configs import build_simulation
    return build_simulation(cfg, precision)


def _run_gradient(job, precision):
    sim = _build(job.cfg, precision)
    try:
        J, g = sim.gradient(job.dobs, job.cfg["nt"], job.checkpoint_every)
    finally:
        sim.close()
    if not np.isfinite(J):
        raise FloatingPointError("the wavefield blew up: non-finite misfit")
    return J, g


def _start_gradient(job, precision):
    from .stepper import GradientJob
    sim = _build(job.cfg, precision)
    try:
        return sim, GradientJob(sim, job.dobs, job.cfg["nt"], job.checkpoint_every)
    except Exception:
        sim.close()
        raise


def summarize(results, t_end=None):
    """shots/s over the stream (jobs / time from the first arrival to the last completion)
    and latency percentiles (completion - arrival)."""
    ok = [r for r in results if r.error is None]
    if not ok:
        return {"jobs": 0}
    lat = np.array([r.latency for r in ok])
    t0 = min(r.arrival for r in ok)
    t1 = max(r.done for r in ok) if t_end is None else t_end
    cells = sum(r.cells * r.steps for r in ok)
    return {"jobs": len(ok), "errors": len(results) - len(ok),
            "wall_s": round(t1 - t0, 4), "shots_per_s": round(len(ok) / (t1 - t0), 3),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "latency_mean_s": round(float(lat.mean()), 4),
            "forward_equivalent_GPt_per_s": round(cells / (t1 - t0) / 1e9, 3)}


def select_admissions(waiting, used, budget, grads_running, max_gradients, fwd_running,
                      max_forward, now=0.0, max_head_wait=1.0):
    """Admission policy. ``waiting``: (id, kind, bytes, arrival) of arrived jobs in arrival
    order. Jobs are taken in order while they fit ``budget`` next to ``used``. When a job
    does not fit, later jobs that fit may go ahead of it (backfill) until it has waited
    ``max_head_wait`` seconds; from then on nothing overtakes it, so the memory drains until
    it fits and a large job is not starved by a stream of small ones. Gradient jobs also need
    a free gradient slot and forward jobs a free batch slot; a job larger than the whole budget is
    rejected. Returns (ids to admit, ids rejected)."""
    picks, rejected = [], []
    g, f = grads_running, fwd_running
    for jid, kind, b, arrival in waiting:
        if b > budget:
            rejected.append(jid)
            continue
        if (kind == "gradient" and g >= max_gradients) or (kind == "forward" and f >= max_forward):
            continue
        if used + b > budget:
            if now - arrival >= max_head_wait:
                break  # the oldest blocked job has waited long enough: no overtaking
            continue
        picks.append(jid)
        used += b
        if kind == "gradient":
            g += 1
        else:
            f += 1
    return picks, rejected


class Server:
    """Continuous-batching shot server on the current GPU.

    ``hbm_budget``: device bytes the admitted jobs may hold (default 85 % of the free
    memory when the server is created); ``block_steps``: steps per batched block (graph
    length) and per gradient chunk; ``max_gradients``: gradient jobs in flight at once;
    ``max_forward``: forward jobs in one batch (16: fewer jobs advancing together finish
    sooner and keep a smaller working set beside the gradient; on the 200-job C5 stream at
    200/s, 8 / 16 / 32 / 64 / 128 gave 3.28 / 3.38 / 3.11 / 3.12 / 3.03 shots/s against 3.08
    for the plain stepper one job at a time); ``max_head_wait``: seconds a job that does
    not fit may be overtaken by later, smaller jobs; ``check_cfl``: refuse jobs whose dt
    exceeds the stability limit when they arrive; ``store_fraction``: gradient jobs whose
    every-step plan needs at most this fraction of the budget record every step instead of
    checkpointing (see gradient_checkpoint); ``max_gradient_chunk``:
    a gradient's steps per chunk start at ``block_steps`` and double each time the job is
    found idle, up to this many (``block_steps`` keeps them fixed: gradients then run dry
    between the loop's turns, which leaves the forward jobs more of the GPU, lower median
    latency and lower throughput); ``tiling``: rows per wave of the forward batch's tiles
    (None: 1 below 2^21 cells in the batch, else 2; fixed 8-row tiles measured within 1 % of
    that on the C5 stream). A forward job whose wavefield stops being finite is failed
    alone after the block in which it happened; so is a gradient job whose misfit is not
    finite."""

    def __init__(self, precision="float32", hbm_budget=None, block_steps=100, max_gradients=1,
                 max_forward=16, graph=True, max_head_wait=1.0, keep_outputs=True,
                 check_cfl=True, store_fraction=0.25, max_gradient_chunk=1 << 16, tiling=None):
        from .batch import mem_info
        self.max_chunk = max(int(max_gradient_chunk), int(block_steps))
        self.tiling = tiling  # rows per wave of the forward batch (None: by its cell count)
        self.precision = precision
        if hbm_budget is None:
            free, _ = mem_info()
            hbm_budget = int(0.85 * free)
        self.budget = int(hbm_budget)
        self.block = int(block_steps)
        self.max_gradients = int(max_gradients)
        self.max_forward = int(max_forward)
        self.graph = graph
        self.max_head_wait = float(max_head_wait)
        self.keep_outputs = keep_outputs
        self.check_cfl = check_cfl
        self.store_fraction = float(store_fraction)
        self.peak_bytes = 0

    def run(self, jobs, on_event=None):
        """Serve ``jobs`` in real time (arrival times relative to the call) and return their
        JobResults in job order."""
        from .batch import Batch
        jobs = sorted((_with_checkpoint(j, gradient_checkpoint(j, self.budget, self.precision,
                                                               self.store_fraction))
                       for j in jobs), key=lambda j: j.arrival)
        res = {j.id: JobResult(j.id, j.kind, j.arrival, cells=j.cfg["nx"] * j.cfg["nz"],
                               steps=j.cfg["nt"]) for j in jobs}
        need = {j.id: job_bytes(j, self.precision) for j in jobs}
        unstable = {j.id: stability_error(j) for j in jobs} if self.check_cfl else {}
        pending = deque(jobs)
        waiting = deque()
        active = {}  # forward: id -> [job, sim, remaining steps]
        grads = {}  # gradient: id -> (sim, GradientJob)
        # steps per gradient chunk: doubled each time a job is found idle, so that a chunk
        # outlasts the loop's turn (a batched block) and the gradient does not run dry
        # between two turns
        chunk = {}
        used = 0
        batch = Batch([])
        if self.graph:
            batch.set_graph(self.block)
        members = None
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0

        def note(what, job_id):
            if on_event:
                on_event(what, job_id, clock())

        def arrivals():
            now = clock()
            while pending and pending[0].arrival <= now:
                job = pending.popleft()
                if unstable.get(job.id):
                    res[job.id].error = unstable[job.id]
                    res[job.id].done = clock()
                    continue
                waiting.append(job)

        def release(jid):
            nonlocal used
            used -= need[jid]
            res[jid].done = clock()
            note("done", jid)

        # Finished jobs whose simulation is closed later. Freeing device memory waits for all
        # work queued on the device, so a simulation is closed only while no gradient work is
        # queued; closing it beside a long gradient chunk would hold the loop (and the forward
        # jobs) until that chunk had run. Its bytes stay accounted until it is closed.
        closing = []  # (job id, simulation)

        def finish(jid, sim):
            res[jid].done = clock()
            note("done", jid)
            closing.append((jid, sim))

        def close_finished():
            nonlocal used
            if closing and all(gj.idle() for _, gj in grads.values()):
                for jid, sim in closing:
                    sim.close()
                    used -= need[jid]
                closing.clear()

        def progress_gradients():
            """Enqueue the next chunk of every gradient whose last chunk has finished; collect
            the finished ones."""
            for jid in list(grads):
                sim, gj = grads[jid]
                if not gj.idle():
                    continue
                r = res[jid]
                try:
                    if gj.finished:
                        del grads[jid]
                        chunk.pop(jid, None)
                        r.misfit, r.gradient = gj.end()
                        if not np.isfinite(r.misfit):
                            r.error = "the wavefield blew up: non-finite misfit"
                            r.gradient = None
                        if not self.keep_outputs:
                            r.gradient = None
                        finish(jid, sim)
                    else:
                        chunk[jid] = min(2 * chunk[jid], self.max_chunk)
                        gj.advance(chunk[jid])
                except Exception as e:  # noqa: BLE001 - reported per job
                    grads.pop(jid, None)
                    chunk.pop(jid, None)
                    gj.cancel()
                    sim.close()
                    r.error = str(e)
                    release(jid)

        def admit():
            nonlocal used
            picks, rejected = select_admissions(
                [(j.id, j.kind, need[j.id], j.arrival) for j in waiting], used, self.budget,
                len(grads), self.max_gradients, len(active), self.max_forward, clock(),
                self.max_head_wait)
            by_id = {j.id: j for j in waiting}
            for jid in rejected:
                waiting.remove(by_id[jid])
                res[jid].error = f"needs {need[jid]} bytes, budget {self.budget}"
                res[jid].done = clock()
            for jid in picks:
                job = by_id[jid]
                waiting.remove(job)
                used += need[jid]
                self.peak_bytes = max(self.peak_bytes, used)
                res[jid].admitted = clock()
                res[jid].bytes = need[jid]
                note("admit", jid)
                if job.kind == "gradient":
                    grads[jid] = _start_gradient(job, self.precision)
                    chunk[jid] = self.block
                    grads[jid][1].advance(self.block)
                else:
                    active[jid] = [job, _build(job.cfg, self.precision), job.cfg["nt"]]

        try:
            while pending or waiting or active or grads or closing:
                close_finished()
                arrivals()
                if not active:
                    admit()
                if active:
                    ids = sorted(active)
                    if ids != members:
                        batch.set([active[k][1] for k in ids])
                        members = ids
                    cells = sum(active[k][0].cfg["nx"] * active[k][0].cfg["nz"] for k in ids)
                    rows = self.tiling or (1 if cells < (1 << 21) else 2)
                    if rows != batch.tiling:
                        batch.set_tiling(rows)
                    n = min(self.block, min(active[k][2] for k in ids))
                    batch.step(n)
                    # while the block runs: gradient chunks, arrivals, and admissions (the new
                    # jobs' simulations are built on the host meanwhile and join the next block)
                    progress_gradients()
                    arrivals()
                    admit()
                    batch.sync()
                    finished = []
                    for k in ids:
                        active[k][2] -= n
                        if not active[k][1].is_finite():  # failed alone, the others go on
                            job, sim, left = active.pop(k)
                            res[k].error = (f"the wavefield blew up (non-finite values by step "
                                            f"{job.cfg['nt'] - left})")
                            finished.append((job, sim, left))
                        elif active[k][2] == 0:
                            finished.append(active.pop(k))
                    if finished:
                        # leave the batch before the handles are released
                        members = sorted(active)
                        batch.set([active[k][1] for k in members])
                        for job, sim, _ in finished:
                            r = res[job.id]
                            if r.error is None:
                                r.traces = sim.traces(job.cfg["nt"])
                            if not self.keep_outputs:
                                r.traces = None
                            finish(job.id, sim)
                elif grads:
                    progress_gradients()
                    time.sleep(0.0005)
                elif closing:
                    continue  # no gradient in flight: closed at the top of the loop
                else:
                    nxt = pending[0].arrival - clock() if pending else 0.01
                    time.sleep(min(max(nxt, 0.0), 0.05))
        finally:
            for jid, sim in closing:
                sim.close()
                used -= need[jid]
            for a in active.values():
                a[1].close()
            for sim, gj in grads.values():
                gj.cancel()
                sim.close()
            batch.close()
        if used != 0:
            raise RuntimeError(f"serving loop ended with {used} bytes still accounted")
        return [res[j.id] for j in sorted(jobs, key=lambda j: j.id)]


def run_serial(jobs, precision="float32", keep_outputs=True, on_event=None, hbm_budget=None,
               check_cfl=True, store_fraction=0.25):
    """Baseline: the plain single-shot stepper, one job at a time in arrival order (a job
    starts when it has arrived and the previous one is done), with the server's checks and
    gradient checkpoint policy, so that the two differ only in how jobs share the GPU."""
    from .batch import mem_info
    if hbm_budget is None:
        hbm_budget = int(0.85 * mem_info()[0])
    jobs = sorted((_with_checkpoint(j, gradient_checkpoint(j, hbm_budget, precision,
                                                           store_fraction))
                   for j in jobs), key=lambda j: j.arrival)
    out = {}
    t0 = time.perf_counter()
    for job in jobs:
        wait = job.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        r = JobResult(job.id, job.kind, job.arrival, cells=job.cfg["nx"] * job.cfg["nz"],
                      steps=job.cfg["nt"])
        r.admitted = time.perf_counter() - t0
        try:
            err = stability_error(job) if check_cfl else None
            if err:
                raise FloatingPointError(err)
            if job.kind == "gradient":
                r.misfit, r.gradient = _run_gradient(job, precision)
            else:
                sim = _build(job.cfg, precision)
                try:
                    sim.step(job.cfg["nt"])
                    if not sim.is_finite():
                        raise FloatingPointError("the wavefield blew up (non-finite values)")
                    r.traces = sim.traces(job.cfg["nt"])
                finally:
                    sim.close()
        except Exception as e:  # noqa: BLE001 - reported per job
            r.error = str(e)
        r.done = time.perf_counter() - t0
        if not keep_outputs:
            r.traces = r.gradient = None
        out[job.id] = r
        if on_event:
            on_event("done", job.id, r.done)
    return [out[k] for k in sorted(out)]
