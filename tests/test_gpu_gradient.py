"""Gradient of the L2 trace misfit by the hand-written discrete adjoint, checked against
central finite differences of the forward map (FP64) at random grid points and along random
directions, for elastic and viscoelastic (L = 1, 3) models with CPML and free surface (stress
imaging, H-AFDA, and H-AFDA with the Hermitian formula #3 on row 5).
The forward map itself is pinned to the CPU oracle by test_gpu_parity.py."""
import numpy as np
import pytest

from helpers import small_cfg

pytestmark = pytest.mark.gpu


def _sim(cfg, model, prec="float64"):
    from viscoelasticinversion_amd.configs import build_simulation
    c = dict(cfg)
    c["model"] = model
    return build_simulation(c, prec)


def _misfit(cfg, model, dobs, nt):
    sim = _sim(cfg, model)
    sim.step(nt)
    tr = sim.traces(nt)
    sim.close()
    return 0.5 * float(np.sum((tr - dobs) ** 2))


def _setup(nrelax, fs, cpml, seed, nt=120):
    """Small heterogeneous case whose data depend on the boundaries: receivers on the rows
    just below the free surface (row 4 is the surface), inside the CPML strips and in the
    interior, on every component; a shallow extra source."""
    cfg = small_cfg(nx=72, nz=56, nrelax=nrelax, cpml=cpml, fs=fs, seed=seed, nt=nt)
    rng = np.random.default_rng(seed + 100)
    nx, nz = cfg["nx"], cfg["nz"]
    rix = np.concatenate([rng.integers(6, nx - 6, 8), rng.integers(4, 8, 4),
                          rng.integers(nx - 8, nx - 4, 4), rng.integers(6, nx - 6, 8)])
    riz = np.concatenate([rng.integers(4, 7, 8), rng.integers(5, nz - 5, 8),
                          rng.integers(nz - 8, nz - 4, 4), rng.integers(8, nz - 8, 4)])
    cfg["rec_ix"] = rix.astype(np.int32)
    cfg["rec_iz"] = riz.astype(np.int32)
    cfg["rec_comp"] = (np.arange(len(rix)) % 5).astype(np.int32)
    # observed data: the same experiment in a perturbed model
    true = {k: v * (1 + 0.03 * rng.standard_normal(v.shape)) for k, v in cfg["model"].items()}
    sim = _sim(cfg, true)
    sim.step(cfg["nt"])
    dobs = sim.traces(cfg["nt"])
    sim.close()
    return cfg, dobs


@pytest.mark.parametrize("nrelax,fs,cpml", [(0, False, False), (0, True, True), (1, True, True),
                                            (3, True, True), (0, "afda", False),
                                            (3, "afda", True), (0, "hermitian", False),
                                            (3, "hermitian", True)])
def test_gradient_matches_finite_differences(nrelax, fs, cpml):
    cfg, dobs = _setup(nrelax, fs, cpml, seed=11 + nrelax)
    nt = cfg["nt"]
    print(f"[case] gradient FD L={nrelax} fs={fs} cpml={cpml}", flush=True)
    sim = _sim(cfg, cfg["model"])
    J, g = sim.gradient(dobs, nt, checkpoint_every=9)
    sim.close()
    assert J > 0
    assert abs(J - _misfit(cfg, cfg["model"], dobs, nt)) <= 1e-12 * J
    rng = np.random.default_rng(5)
    keys = ["rho", "lam", "mu"] + (["tau"] if nrelax else [])
    nz, nx = cfg["nz"], cfg["nx"]
    # 1) random directions over the whole interior
    for key in keys:
        dm = np.zeros((nz, nx))
        dm[4:-4, 4:-4] = rng.standard_normal((nz - 8, nx - 8))
        base = cfg["model"][key]
        eps = 1e-6
        mp = dict(cfg["model"]); mp[key] = base * (1 + eps * dm)
        mm = dict(cfg["model"]); mm[key] = base * (1 - eps * dm)
        fd = (_misfit(cfg, mp, dobs, nt) - _misfit(cfg, mm, dobs, nt)) / (2 * eps)
        ad = float(np.sum(g[key] * base * dm))
        print(f"  direction {key}: adjoint {ad:.9e} fd {fd:.9e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-14 * J, (key, ad, fd)
    # 2) single interior grid points with the largest sensitivities + random ones
    for key in keys:
        base = cfg["model"][key]
        gi = np.abs(g[key][4:-4, 4:-4] * base[4:-4, 4:-4])
        flat = np.argsort(gi.ravel())[::-1][:3].tolist() + rng.integers(0, gi.size, 2).tolist()
        for f in flat:
            j, i = np.unravel_index(f, gi.shape)
            j, i = j + 4, i + 4
            eps = 1e-5
            mp = dict(cfg["model"]); mp[key] = base.copy(); mp[key][j, i] *= 1 + eps
            mm = dict(cfg["model"]); mm[key] = base.copy(); mm[key][j, i] *= 1 - eps
            fd = (_misfit(cfg, mp, dobs, nt) - _misfit(cfg, mm, dobs, nt)) / (2 * eps * base[j, i])
            ad = float(g[key][j, i])
            assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-3 * float(np.abs(g[key]).max())), \
                (key, j, i, ad, fd)


@pytest.mark.parametrize("nrelax,fs", [(0, True), (3, True), (3, "afda"), (3, "hermitian")])
def test_adjoint_dot_product(nrelax, fs):
    """<J dm, dd> = <dm, J^T dd> for the Jacobian J of the receiver traces with respect to the
    model: J^T dd is the gradient with the residual set to dd (observed data d - dd), J dm
    the central difference of the traces along dm (FP64), for random dm and dd."""
    cfg, _ = _setup(nrelax, fs, True, seed=31 + nrelax)
    nt = cfg["nt"]
    rng = np.random.default_rng(77)
    sim = _sim(cfg, cfg["model"])
    sim.step(nt)
    d0 = sim.traces(nt)
    sim.close()
    dd = rng.standard_normal(d0.shape) * np.abs(d0).max()
    print(f"[case] adjoint dot product L={nrelax} fs={fs}", flush=True)
    sim = _sim(cfg, cfg["model"])
    _, g = sim.gradient(d0 - dd, nt, checkpoint_every=0)
    sim.close()
    nz, nx = cfg["nz"], cfg["nx"]
    for key in ["rho", "lam", "mu"] + (["tau"] if nrelax else []):
        dm = np.zeros((nz, nx))
        dm[4:-4, 4:-4] = rng.standard_normal((nz - 8, nx - 8))
        base = cfg["model"][key]
        eps = 1e-6
        tr = []
        for sgn in (1, -1):
            m = dict(cfg["model"]); m[key] = base * (1 + sgn * eps * dm)
            s_ = _sim(cfg, m)
            s_.step(nt)
            tr.append(s_.traces(nt))
            s_.close()
        jdm = (tr[0] - tr[1]) / (2 * eps)          # J (base * dm)
        lhs = float(np.sum(jdm * dd))
        rhs = float(np.sum(g[key] * base * dm))
        assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs)), (key, lhs, rhs)


def test_gradient_tau_to_zero_is_the_elastic_gradient():
    """With tau = 0 the SLS terms vanish from the stress update (X1 = X tau = 0), so the
    viscoelastic step is the elastic one and its gradient with respect to rho, lam and mu is
    the elastic gradient (the memory variables still evolve but feed nothing back)."""
    cfg0, dobs = _setup(0, True, True, seed=21)
    cfg3 = small_cfg(nx=72, nz=56, nrelax=3, cpml=True, fs=True, seed=21, nt=cfg0["nt"])
    for k in ("src_ix", "src_iz", "src_kind", "wavelets", "rec_ix", "rec_iz", "rec_comp"):
        cfg3[k] = cfg0[k]
    m3 = dict(cfg0["model"])
    m3["tau"] = np.zeros_like(m3["rho"])
    print("[case] gradient tau -> 0", flush=True)
    s0 = _sim(cfg0, cfg0["model"])
    J0, g0 = s0.gradient(dobs, cfg0["nt"], checkpoint_every=0)
    s0.close()
    s3 = _sim(cfg3, m3)
    J3, g3 = s3.gradient(dobs, cfg0["nt"], checkpoint_every=0)
    s3.close()
    assert abs(J3 - J0) <= 1e-13 * J0
    for k in ("rho", "lam", "mu"):
        assert np.linalg.norm(g3[k] - g0[k]) <= 1e-12 * np.linalg.norm(g0[k]), k
    assert np.abs(g3["tau"]).max() > 0  # the data are still sensitive to attenuation


def test_gradient_checkpoint_interval_invariant():
    """The gradient does not depend on the checkpoint interval (1, 7, sqrt(nt), and nt or more:
    one segment, recorded during the forward pass without recomputation)."""
    cfg, dobs = _setup(2, True, True, seed=3)
    nt = cfg["nt"]
    print("[case] gradient checkpoint intervals", flush=True)
    res = []
    for K in (1, 7, 0, nt, nt + 5):
        sim = _sim(cfg, cfg["model"])
        res.append(sim.gradient(dobs, nt, checkpoint_every=K))
        sim.close()
    for J, g in res[1:]:
        assert J == res[0][0]
        for k in g:
            assert np.array_equal(g[k], res[0][1][k]), k


def test_gradient_independent_of_tiling():
    """The gradient does not depend on the handle's tiling (the recording launches use tiles
    of at most two rows per wavefront), and the handle keeps its tiling afterwards."""
    cfg, dobs = _setup(1, "afda", True, seed=9, nt=60)
    print("[case] gradient tilings", flush=True)
    res = []
    for rows in (1, 2, 4, 8):
        sim = _sim(cfg, cfg["model"])
        sim.set_tiling(rows)
        res.append(sim.gradient(dobs, 60, checkpoint_every=9))
        assert sim.tiling == rows
        sim.close()
    for J, g in res[1:]:
        assert J == res[0][0]
        for k in g:
            assert np.array_equal(g[k], res[0][1][k]), k


def test_gradient_fp32_close_to_fp64():
    cfg, dobs = _setup(3, True, True, seed=7)
    nt = cfg["nt"]
    print("[case] gradient fp32 vs fp64", flush=True)
    s64 = _sim(cfg, cfg["model"], "float64")
    J64, g64 = s64.gradient(dobs, nt)
    s64.close()
    s32 = _sim(cfg, cfg["model"], "float32")
    J32, g32 = s32.gradient(dobs.astype(np.float32), nt)
    s32.close()
    assert abs(J32 - J64) <= 1e-4 * J64
    for k in g64:
        num = np.linalg.norm(g32[k] - g64[k])
        den = np.linalg.norm(g64[k])
        print(f"  {k}: rel L2 {num / den:.2e}", flush=True)
        assert num <= 1e-3 * den, k


def test_workload_gradient_vp_vs_rho_q_at_random_points():
    """The gradient workload (smooth layered background with Gaussian anomalies, Q per layer,
    L = 3, CPML + free surface, several shots, surface receivers) at reduced size in FP64:
    dJ/d(vp, vs, rho, Q) from the adjoint and the chain rule against central differences of
    the multi-shot misfit at 20 random grid points per parameter (drawn among the cells whose
    sensitivity |g m| is at least 1 % of the largest, so that every check is non-trivial)."""
    from viscoelasticinversion_amd.configs import (config_adjoint, moduli_from_velocity,
                                                   multishot_gradient, multishot_misfit,
                                                   observed_data)
    print("[case] workload gradient 200x100 3 shots FP64", flush=True)
    acfg = config_adjoint(nx=200, nz=100, nshots=3, nrec=41, nt=500, seed=3, smooth_cells=8)
    dobs = observed_data(acfg, "float64")
    st = acfg["start"]
    J, g = multishot_gradient(acfg, dobs, "float64", "start", checkpoint_every=0)
    assert abs(J - multishot_misfit(acfg, st, dobs, "float64")) <= 1e-12 * J
    rng = np.random.default_rng(7)
    nz, nx = acfg["nz"], acfg["nx"]
    keys = ("vp", "vs", "rho", "Q")
    for key in keys:
        sens = np.abs(g[key] * st[key])
        sens[:5, :] = 0
        sens[-4:, :] = 0
        sens[:, :4] = 0
        sens[:, -4:] = 0
        cand = np.argwhere(sens >= 0.01 * sens.max())
        pts = cand[rng.choice(len(cand), 20, replace=False)]
        worst = 0.0
        for j, i in pts:
            eps = 1e-4 * st[key][j, i]

            def J_at(sign):
                p = {k: st[k].copy() for k in keys}
                p[key][j, i] += sign * eps
                m = moduli_from_velocity(p["vp"], p["vs"], p["rho"], p["Q"], acfg["kbar"], 3)
                return multishot_misfit(acfg, m, dobs, "float64")

            fd = (J_at(1) - J_at(-1)) / (2 * eps)
            ad = float(g[key][j, i])
            err = abs(ad - fd) / abs(fd)
            worst = max(worst, err)
            assert err <= 1e-5, (key, int(j), int(i), ad, fd)
        print(f"  {key}: 20 points, worst relative difference {worst:.2e}", flush=True)


@pytest.mark.parametrize("concurrent,K", [(2, 0), (3, 40), (5, 200)])
def test_multishot_gradient_concurrent_jobs_equal_one_at_a_time(concurrent, K):
    """Shots computed as concurrent gradient jobs (each on a simulation and stream of its own,
    sources moved from shot to shot) give the one-at-a-time misfit and gradients bit for bit,
    for any number of jobs (more jobs than shots included) and checkpoint interval."""
    from viscoelasticinversion_amd.configs import (config_adjoint, multishot_gradient,
                                                   observed_data)
    print(f"[case] multishot gradient, {concurrent} concurrent jobs, K={K}", flush=True)
    acfg = config_adjoint(nx=200, nz=100, nshots=4, nrec=41, nt=200, seed=5, smooth_cells=8)
    dobs = observed_data(acfg, "float32")
    J1, g1 = multishot_gradient(acfg, dobs, "float32", "start", checkpoint_every=K)
    Jc, gc = multishot_gradient(acfg, dobs, "float32", "start", checkpoint_every=K,
                                concurrent=concurrent, chunk=37)
    assert Jc == J1
    for k in g1:
        assert np.array_equal(gc[k], g1[k]), k


@pytest.mark.parametrize("vti,K", [(False, 8), (True, 8), (False, 70)])
def test_chunked_gradient_job_equals_gradient(vti, K):
    """The gradient as a chunked job (begin / advance by a few steps / end) is bit-identical
    to the one-call gradient (K = 70: every step recorded in the forward pass), and the handle
    is locked while the job owns it."""
    from viscoelasticinversion_amd.configs import build_simulation
    from viscoelasticinversion_amd.stepper import GradientJob
    from viscoelasticinversion_amd import _lib
    cfg = small_cfg(nx=80, nz=60, nrelax=3, seed=51, nt=70, vti=vti)
    sim = build_simulation(cfg, "float32")
    sim.step(70)
    dobs = sim.traces(70) * 0.9
    sim.reset()
    print(f"[case] chunked gradient vti={vti} K={K}", flush=True)
    J, g = sim.gradient(dobs, 70, checkpoint_every=8)
    job = GradientJob(sim, dobs, 70, checkpoint_every=K)
    with pytest.raises(_lib.ViError):
        sim.step(1)
    while not job.finished:
        job.advance(5)
    J2, g2 = job.end()
    assert J2 == J
    for k in g:
        assert np.array_equal(g[k], g2[k]), k
    sim.step(2)  # the handle is free again
    sim.close()


def test_chunked_gradient_job_after_the_observed_data_are_freed():
    """begin copies the observed data before it returns: the caller's buffer (here a fresh
    float64 array that the job converts and drops) can be freed and its memory reused while
    the job runs, and the result still equals the one-call gradient bit for bit."""
    import gc
    from viscoelasticinversion_amd.configs import build_simulation
    from viscoelasticinversion_amd.stepper import GradientJob
    cfg = small_cfg(nx=80, nz=60, nrelax=2, seed=52, nt=60)
    sim = build_simulation(cfg, "float32")
    sim.step(60)
    dobs = sim.traces(60).astype(np.float64) * 0.95
    sim.reset()
    print("[case] chunked gradient with the observed data freed after begin", flush=True)
    J, g = sim.gradient(dobs, 60, checkpoint_every=6)
    job = GradientJob(sim, dobs.copy(), 60, checkpoint_every=6)
    gc.collect()
    junk = [np.full(dobs.shape, np.nan, np.float32) for _ in range(8)]  # reuse freed memory
    while not job.finished:
        job.advance(7)
    J2, g2 = job.end()
    del junk
    assert J2 == J
    for k in g:
        assert np.array_equal(g[k], g2[k]), k
    sim.close()
