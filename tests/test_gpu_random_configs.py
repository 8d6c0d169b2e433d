"""Seeded random configurations against the CPU oracle: grid sizes (including non-multiples of
the tile), L = 0..6, free surface none / imaging / H-AFDA, CPML on or off (with alpha and
kappa), isotropic or VTI, every tiling, FP32 or FP64, random source kinds and receivers on
every component. Each case is a fixed seed, so a failure names a reproducible configuration.
Bar: relative L2 <= 1e-5 (FP32) / 1e-12 (FP64)."""
import numpy as np
import pytest

from helpers import compare, run_gpu, run_oracle, small_cfg

pytestmark = pytest.mark.gpu


def _random_cfg(seed):
    rng = np.random.default_rng(1000 + seed)
    nx = int(rng.integers(16, 260))
    nz = int(rng.integers(17, 150))
    nrelax = int(rng.integers(0, 7))
    fs = [False, True, "afda"][int(rng.integers(0, 3))]
    cpml = bool(rng.integers(0, 2)) and min(nx, nz) >= 20
    vti = bool(rng.integers(0, 2))
    nt = int(rng.integers(10, 60))
    cfg = small_cfg(nx=nx, nz=nz, nrelax=nrelax, cpml=cpml, fs=fs, seed=seed, nt=nt, vti=vti)
    if cpml:
        cfg["cpml"]["thickness"] = int(rng.integers(2, min(nx, nz) // 2 - 1))
    nsrc = int(rng.integers(0, 6))
    cfg["src_ix"] = rng.integers(4, nx - 4, nsrc).astype(np.int32)
    cfg["src_iz"] = rng.integers(4, nz - 4, nsrc).astype(np.int32)
    cfg["src_kind"] = rng.integers(0, 3, nsrc).astype(np.int32)
    t = np.arange(nt + 5) * cfg["dt"]
    cfg["wavelets"] = np.stack([1e3 * (k + 1) * np.sin(2 * np.pi * 25 * t) *
                                np.exp(-((t - 0.003) / 0.002) ** 2) for k in range(nsrc)]) \
        if nsrc else np.zeros((0, nt + 5))
    nrec = int(rng.integers(0, 12))
    cfg["rec_ix"] = rng.integers(4, nx - 4, nrec).astype(np.int32)
    cfg["rec_iz"] = rng.integers(4, nz - 4, nrec).astype(np.int32)
    cfg["rec_comp"] = rng.integers(0, 5, nrec).astype(np.int32)
    rows = [1, 2, 4, 8][int(rng.integers(0, 4))]
    dtype = [np.float32, np.float64][int(rng.integers(0, 2))]
    return cfg, rows, dtype


@pytest.mark.parametrize("seed", range(96))
def test_random_configuration(seed):
    cfg, rows, dtype = _random_cfg(seed)
    prec = "float64" if dtype == np.float64 else "float32"
    print(f"[case] seed {seed}: {cfg['nx']}x{cfg['nz']} L={cfg['nrelax']} "
          f"fs={cfg['free_surface']} cpml={bool(cfg['cpml'])} vti={'c11' in cfg['model']} "
          f"rows={rows} {prec} nt={cfg['nt']} nsrc={len(cfg['src_ix'])} "
          f"nrec={len(cfg['rec_ix'])}", flush=True)
    ref = run_oracle(cfg, dtype)
    gpu = run_gpu(cfg, prec, tiling=rows)
    if not len(cfg["rec_ix"]):
        ref.pop("traces")
        gpu.pop("traces")
    compare(gpu, ref, dtype, f"random seed {seed}")


@pytest.mark.parametrize("seed", range(6))
def test_random_batch_equals_alone(seed):
    """A random mix of the configurations above (different sizes, surfaces, L, VTI and
    precision) in one batch, with a random cache block and graph block: every shot equals the
    same shot stepped alone, bit for bit."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_simulation
    rng = np.random.default_rng(500 + seed)
    picks = [_random_cfg(int(s)) for s in rng.integers(0, 1000, 5)]
    nt = 24
    sims, cfgs = [], []
    for cfg, _, dtype in picks:
        cfg["nt"] = nt
        prec = "float64" if dtype == np.float64 else "float32"
        sims.append(build_simulation(cfg, prec))
        cfgs.append((cfg, prec))
    block = int(rng.choice([0, 1, 1 << 20, 224 << 20]))
    graph = int(rng.choice([0, 5, 10]))
    print(f"[case] random batch {seed}: block {block}, graph {graph}", flush=True)
    b = Batch(sims)
    b.set_cache_block(block)
    b.set_graph(graph)
    b.step(11)
    b.step(13)
    b.sync()
    for (cfg, prec), sim in zip(cfgs, sims):
        got = {f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz")}
        alone = run_gpu(cfg, prec)
        for f in got:
            assert np.array_equal(got[f], alone[f]), f
        if len(cfg["rec_ix"]):
            assert np.array_equal(sim.traces(nt), alone["traces"])
    b.close()
    for s in sims:
        s.close()
