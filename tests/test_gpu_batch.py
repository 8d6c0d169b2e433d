"""Batched multi-shot stepping: every shot advanced in the shared launch equals the same
shot stepped alone and the CPU oracle run of that shot (bit-identical in practice; bar
relative L2 <= 1e-5 FP32 / 1e-12 FP64)."""
import numpy as np
import pytest

from helpers import compare, run_gpu, run_oracle, small_cfg

pytestmark = pytest.mark.gpu


def _small_multishot(nshots=5, nt=40, precision_seed=0):
    from viscoelasticinversion_amd.configs import config_multishot
    return config_multishot(nshots=nshots, nz=72, wmin=60, wmax=230, nt=nt, seed=precision_seed,
                            model_nx=700, rec_every=9)


def _collect(sim, nt, nrelax):
    out = {k: sim.get(k) for k in ("vx", "vz", "sxx", "szz", "sxz")}
    if nrelax:
        for k in ("mxx", "mzz", "mxz"):
            out[k] = sim.get(k)
    out["traces"] = sim.traces(nt)
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_equals_oracle_and_alone(dtype):
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_multishot, shot_config
    prec = "float32" if dtype == np.float32 else "float64"
    mcfg = _small_multishot()
    nt = mcfg["nt"]
    print(f"[case] batch of {len(mcfg['shots'])} shots {prec}", flush=True)
    model, sims = build_multishot(mcfg, prec)
    b = Batch(sims)
    b.step(nt)
    b.sync()
    for s, sim in enumerate(sims):
        got = _collect(sim, nt, mcfg["nrelax"])
        cfg = shot_config(mcfg, s)
        ref = run_oracle(cfg, dtype)
        compare(got, {k: ref[k] for k in got}, dtype, f"batch shot {s} vs oracle")
        alone = run_gpu(cfg, prec)
        for k in got:
            assert np.array_equal(got[k], alone[k]), (s, k)
    b.close()
    for sim in sims:
        sim.close()
    model.close()


@pytest.mark.parametrize("streams", [1, 2, 3, 4])
@pytest.mark.parametrize("block", [0, 1, 3 << 20])
def test_cache_blocked_batch_equals_alone(block, streams):
    """Cache blocking over shots (groups of at most `block` / `streams` bytes of working set
    run each block of steps in turn; 1 byte = one shot per group), on one stream or with
    consecutive groups taking 2 to 4 streams in turn, changes nothing in the results, eager
    or with graph blocks."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_multishot, shot_config
    mcfg = _small_multishot(nshots=6, nt=45)
    nt = mcfg["nt"]
    print(f"[case] cache-blocked batch, block {block} bytes, {streams} streams", flush=True)
    model, sims = build_multishot(mcfg, "float32")
    b = Batch(sims)
    b.set_streams(streams)
    b.set_cache_block(block)
    assert b.cache_block == block
    b.set_graph(10)
    b.step(17)
    b.step(28)
    b.sync()
    for s, sim in enumerate(sims):
        got = _collect(sim, nt, mcfg["nrelax"])
        alone = run_gpu(shot_config(mcfg, s), "float32")
        for k in got:
            assert np.array_equal(got[k], alone[k]), (s, k)
    b.close()
    for sim in sims:
        sim.close()
    model.close()


def test_set_streams_bounds():
    from viscoelasticinversion_amd import Batch, _lib
    print("[case] set_streams bounds", flush=True)
    b = Batch([])
    for n in (0, 5, -1):
        with pytest.raises(_lib.ViError):
            b.set_streams(n)
    for n in (1, 4, 2):
        b.set_streams(n)
    b.close()


@pytest.mark.parametrize("rows", [1, 2, 4, 8])
def test_batch_tilings(rows):
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_multishot, shot_config
    mcfg = _small_multishot(nshots=3, nt=25, precision_seed=2)
    print(f"[case] batch tiling {rows}", flush=True)
    model, sims = build_multishot(mcfg, "float32")
    b = Batch(sims)
    b.set_tiling(rows)
    b.step(25)
    for s, sim in enumerate(sims):
        got = _collect(sim, 25, 3)
        ref = run_oracle(shot_config(mcfg, s), np.float32)
        compare(got, {k: ref[k] for k in got}, np.float32, f"tiling {rows} shot {s}")
    b.close()


def test_graph_replays_match_eager():
    """hipGraph of 7 steps replayed several times, with eager host reads in between; every
    checkpoint is compared with the oracle at the same step."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_multishot, shot_config
    mcfg = _small_multishot(nshots=4, nt=40, precision_seed=3)
    print("[case] graph replays", flush=True)
    model, sims = build_multishot(mcfg, "float32")
    b = Batch(sims)
    b.set_graph(7)
    done = 0
    for chunk in (7, 14, 3, 16):
        b.step(chunk)
        done += chunk
        for s, sim in enumerate(sims):
            assert sim.step_index == done
            got = {k: sim.get(k) for k in ("vx", "szz", "sxz", "mxx")}
            got["traces"] = sim.traces(done)
            ref = run_oracle(shot_config(mcfg, s), np.float32, nt=done)
            compare(got, {k: ref[k][..., :done] if k == "traces" else ref[k] for k in got},
                    np.float32, f"graph step {done} shot {s}")
    b.close()


def test_batch_then_single_steps_and_membership_change():
    """Stepping a shot in a batch, then alone, then in a different batch is seamless."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_multishot, shot_config
    mcfg = _small_multishot(nshots=4, nt=30, precision_seed=4)
    print("[case] membership changes", flush=True)
    model, sims = build_multishot(mcfg, "float64")
    b = Batch(sims)
    b.step(10)
    sims[0].step(5)
    sims[1].step(5)
    b.set(sims[:2])
    b.step(15)
    b.set(sims[2:])
    b.step(20)
    for s, sim in enumerate(sims):
        assert sim.step_index == 30
        got = _collect(sim, 30, 3)
        ref = run_oracle(shot_config(mcfg, s), np.float64)
        compare(got, {k: ref[k] for k in got}, np.float64, f"membership shot {s}")
    b.close()


@pytest.mark.parametrize("graph", [0, 4])
def test_member_stepped_alone_between_batch_steps(graph):
    """A member advanced on its own between two batch steps of the same member set continues
    from its own step index (the batch rebuilds its descriptors), with and without graphs."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_multishot, shot_config
    mcfg = _small_multishot(nshots=3, nt=30, precision_seed=6)
    print(f"[case] member stepped alone between batch steps, graph={graph}", flush=True)
    model, sims = build_multishot(mcfg, "float32")
    b = Batch(sims)
    b.set_graph(graph)
    b.step(9)
    sims[1].step(4)
    b.step(9)
    sims[0].step(3)
    b.step(8)
    b.sync()
    for s, nt in enumerate((29, 30, 26)):
        assert sims[s].step_index == nt
        got = _collect(sims[s], nt, 3)
        ref = run_oracle(shot_config(mcfg, s), np.float32, nt=nt)
        compare(got, {k: ref[k] for k in got}, np.float32, f"alone-between shot {s}")
    b.close()
    for sim in sims:
        sim.close()
    model.close()


@pytest.mark.parametrize("block", [None, 1])
def test_batch_autotune_keeps_state(block):
    """The autotuner's trials (every tiling, and lockstep against the cache block when one
    is set; 1 byte = one shot per group) leave the state and step index as they were."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_multishot, shot_config
    mcfg = _small_multishot(nshots=3, nt=30, precision_seed=5)
    print(f"[case] batch autotune, block {block}", flush=True)
    model, sims = build_multishot(mcfg, "float32")
    b = Batch(sims)
    if block is not None:
        b.set_cache_block(block)
    b.step(12)
    chosen, table = b.autotune(trial_steps=4)
    assert chosen in (1, 2, 4, 8) and b.tiling == chosen
    if block is not None:
        assert b.cache_block in (0, block)
    b.step(18)
    for s, sim in enumerate(sims):
        got = _collect(sim, 30, 3)
        ref = run_oracle(shot_config(mcfg, s), np.float32)
        compare(got, {k: ref[k] for k in got}, np.float32, f"autotune shot {s}")
    b.close()


def test_cache_blocked_batch_mixed_classes_equals_alone():
    """Shots of different precisions, relaxation counts and sizes in one batch: each (precision,
    L) class is packed into its own cache-block groups (largest working set first), and every
    shot still equals itself stepped alone, bit for bit."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_simulation
    print("[case] cache-blocked batch, mixed precision and L", flush=True)
    specs = [("float32", 2, 90, 60), ("float64", 2, 70, 50), ("float32", 0, 130, 48),
             ("float32", 2, 160, 72), ("float64", 2, 100, 64), ("float32", 0, 60, 40),
             ("float32", 2, 75, 45)]
    cfgs = [small_cfg(nx=nx, nz=nz, nrelax=L, seed=30 + q, nt=30)
            for q, (_, L, nx, nz) in enumerate(specs)]
    sims = [build_simulation(c, p) for c, (p, _, _, _) in zip(cfgs, specs)]
    b = Batch(sims)
    b.set_streams(2)
    b.set_cache_block(4 << 20)  # about two shots per group and stream
    b.set_graph(10)
    b.step(17)
    b.step(13)
    b.sync()
    for q, (sim, c) in enumerate(zip(sims, cfgs)):
        got = _collect(sim, 30, c["nrelax"])
        alone = run_gpu(c, specs[q][0])
        for k in got:
            assert np.array_equal(got[k], alone[k]), (q, k)
    b.close()
    for sim in sims:
        sim.close()
