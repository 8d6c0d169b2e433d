"""Reuse of large device buffers: buffers of 64 MB or more that a simulation releases are
kept by the library and handed to the next allocation they fit. A gradient that runs in a
workspace left dirty by another job must equal the same gradient in fresh allocations, bit
for bit, and the kept bytes must be visible (and freeable) through the memory queries."""
import numpy as np
import pytest

from helpers import small_cfg

pytestmark = pytest.mark.gpu

NT = 120


def _case(seed, vti=False, fs=True):
    # 400 x 300 with every step recorded: the slot buffer is 120 steps x 8 planes x ~0.7 MB
    # (FP32), well above the 64 MB threshold
    cfg = small_cfg(nx=400, nz=300, nrelax=2, seed=seed, nt=NT, vti=vti, fs=fs)
    cfg["rec_ix"] = np.arange(20, 380, 9).astype(np.int32)
    cfg["rec_iz"] = np.full(len(cfg["rec_ix"]), 12, np.int32)
    cfg["rec_comp"] = (np.arange(len(cfg["rec_ix"])) % 5).astype(np.int32)
    return cfg


def _gradient(cfg, dobs, prec="float32"):
    from viscoelasticinversion_amd.configs import build_simulation
    sim = build_simulation(cfg, prec)
    try:
        return sim.gradient(dobs, NT, NT)
    finally:
        sim.close()


def _observed(cfg, prec="float32"):
    from viscoelasticinversion_amd.configs import build_simulation
    c = dict(cfg)
    c["model"] = {k: v * 1.02 for k, v in cfg["model"].items()}
    sim = build_simulation(c, prec)
    sim.step(NT)
    d = sim.traces(NT)
    sim.close()
    return d


@pytest.mark.parametrize("prec,vti,fs", [("float32", False, True), ("float64", True, "afda"),
                                         ("float32", True, "hermitian")])
def test_gradient_in_a_reused_dirty_workspace_equals_fresh(prec, vti, fs):
    from viscoelasticinversion_amd import cached_bytes, mem_info, release_cached
    print(f"[case] gradient in reused buffers: {prec}, vti {vti}, surface {fs}", flush=True)
    a, b = _case(11, vti, fs), _case(12, vti, fs)
    da, db = _observed(a, prec), _observed(b, prec)
    release_cached()
    assert cached_bytes() == 0
    J_fresh, g_fresh = _gradient(b, db, prec)  # fresh allocations
    kept = cached_bytes()
    assert kept >= 64 << 20
    free_with, _ = mem_info()
    release_cached()
    free_without, _ = mem_info()
    assert cached_bytes() == 0
    assert abs(free_with - free_without) < (64 << 20)  # kept bytes count as free
    _gradient(b, db, prec)
    _gradient(a, da, prec)  # another job writes its own values into the kept buffers
    assert cached_bytes() >= 64 << 20
    J, g = _gradient(b, db, prec)  # runs in the buffers the other job left behind
    assert J == J_fresh
    for k in g_fresh:
        assert np.array_equal(g[k], g_fresh[k]), k
    release_cached()
    assert cached_bytes() == 0


def test_forward_after_reuse_equals_fresh():
    from viscoelasticinversion_amd import cached_bytes, release_cached
    from viscoelasticinversion_amd.configs import build_simulation
    print("[case] forward in reused buffers", flush=True)
    cfg = small_cfg(nx=2100, nz=1100, nrelax=3, seed=13, nt=30)  # ~9 MB planes, 64 MB M
    release_cached()
    outs = []
    for rep in range(2):
        _gradient(_case(14), _observed(_case(14)))  # leaves dirty kept buffers
        sim = build_simulation(cfg, "float32")
        sim.step(30)
        outs.append({k: sim.get(k) for k in ("vx", "vz", "sxx", "szz", "sxz", "mxx")})
        outs[-1]["traces"] = sim.traces(30)
        sim.close()
        if rep == 0:
            release_cached()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    release_cached()
    assert cached_bytes() == 0


def test_failed_allocation_frees_kept_buffers_and_retries():
    """A gradient workspace that does not fit next to the kept buffers, but fits once they are
    freed: the allocation fails, the kept buffers of the device are freed, and the retry
    succeeds."""
    from viscoelasticinversion_amd import cached_bytes, mem_info, release_cached
    from viscoelasticinversion_amd.configs import build_simulation

    def make(nt):
        return build_simulation(small_cfg(nx=1024, nz=1024, nrelax=0, seed=21, nt=nt), "float32")

    print("[case] allocation retried after freeing kept buffers", flush=True)
    release_cached()
    sim = make(16)
    per_step = sim.gradient_bytes(2000, 2000) / 2000  # every step recorded
    sim.close()
    free0, total = mem_info()
    nt_a = int(min(total // 4, 60e9) / per_step)
    sim = make(nt_a)
    J_a, _ = sim.gradient(np.zeros((sim.nrec, nt_a), np.float32), nt_a, nt_a)
    sim.close()
    assert np.isfinite(J_a)
    kept = cached_bytes()
    assert kept >= 0.9 * nt_a * per_step
    free_all, _ = mem_info()
    raw = free_all - kept  # what the runtime reports free
    nt_b = int((raw + kept / 2) / per_step)
    sim = make(nt_b)
    need = sim.gradient_bytes(nt_b, nt_b)
    if not raw + 4e9 < need < free_all - 8e9:
        sim.close()
        pytest.skip(f"no window on this device: free {raw:.3g} + kept {kept:.3g}, need {need:.3g}")
    print(f"  kept {kept / 1e9:.1f} GB, free without them {raw / 1e9:.1f} GB, "
          f"gradient needs {need / 1e9:.1f} GB", flush=True)
    J_b, g_b = sim.gradient(np.zeros((sim.nrec, nt_b), np.float32), nt_b, nt_b)
    sim.close()
    assert np.isfinite(J_b) and all(np.isfinite(v).all() for v in g_b.values())
    assert cached_bytes() <= total / 3
    release_cached()
    assert cached_bytes() == 0
