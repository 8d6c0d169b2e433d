"""H-AFDA formula #3 (free surface "hermitian", mode 3) on the GPU: the HIP stepper equals the
CPU oracle bit for bit on identical seeded inputs (the paths share the operation order) on
every tiling, FP32 and FP64, elastic, viscoelastic and VTI, with CPML, for several Hermitian
point counts, and in batches that mix every surface mode. Its gradient is checked against
finite differences and by the dot-product test in test_gpu_gradient.py and test_gpu_vti.py."""
import numpy as np
import pytest

from helpers import run_gpu, run_oracle, small_cfg

pytestmark = pytest.mark.gpu


def _prec(dtype):
    return "float64" if dtype == np.float64 else "float32"


def _same(gpu, ref, label):
    for k in ref:
        assert np.array_equal(gpu[k], ref[k]), (label, k)


@pytest.mark.parametrize("rows", [1, 2, 4, 8])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_hermitian_every_tiling(rows, dtype):
    cfg = small_cfg(nx=150, nz=90, nrelax=3, nt=60, fs="hermitian")
    print(f"[case] hermitian tiling rows={rows} {_prec(dtype)}", flush=True)
    ref = run_oracle(cfg, dtype)
    gpu = run_gpu(cfg, _prec(dtype), tiling=rows)
    _same(gpu, ref, f"tiling {rows}")
    for k in ("vx", "vz", "szz", "sxz"):
        assert not np.any(gpu[k][:4]), k
    assert np.all(gpu["szz"][4, 4:-4] == 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nrelax,vti", [(0, False), (2, False), (0, True), (3, True)])
def test_hermitian_media(dtype, nrelax, vti):
    cfg = small_cfg(nx=100, nz=66, nrelax=nrelax, nt=70, fs="hermitian", vti=vti)
    print(f"[case] hermitian L={nrelax} vti={vti} {_prec(dtype)}", flush=True)
    _same(run_gpu(cfg, _prec(dtype)), run_oracle(cfg, dtype), f"L={nrelax} vti={vti}")


@pytest.mark.parametrize("points", [2, 3, 8])
def test_hermitian_points(points):
    cfg = small_cfg(nx=90, nz=60, nrelax=2, nt=40, fs="hermitian")
    cfg["hermitian_points"] = points
    cfg["afda_points"] = [2, 3, 4, 5, 2, 3, 4, 2, 3, 4, 2, 3, 4]
    print(f"[case] hermitian points {points}", flush=True)
    for dtype in (np.float64, np.float32):
        _same(run_gpu(cfg, _prec(dtype)), run_oracle(cfg, dtype), f"points {points}")


def test_every_surface_mode_in_one_batch():
    """Shots with formula #3, H-AFDA, imaging and no surface advance in one launch and each
    equals the same shot stepped alone and the oracle, bit for bit."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_simulation
    cfgs = [small_cfg(nx=70 + 23 * k, nz=50 + 7 * k, nrelax=2, nt=50, seed=k + 3, fs=fs)
            for k, fs in enumerate(("hermitian", "afda", True, False, "hermitian"))]
    print("[case] batch of 5 shots with every surface mode", flush=True)
    sims = [build_simulation(c, "float32") for c in cfgs]
    b = Batch(sims)
    b.step(30)
    b.step(20)
    b.sync()
    for k, (c, sim) in enumerate(zip(cfgs, sims)):
        got = {f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz", "mxx", "mzz", "mxz")}
        got["traces"] = sim.traces(50)
        alone = run_gpu(c, "float32")
        ref = run_oracle(c, np.float32)
        for f in got:
            assert np.array_equal(got[f], alone[f]), (k, f)
            assert np.array_equal(got[f], ref[f]), (k, f)
    b.close()
    for s in sims:
        s.close()


def test_hermitian_mode_switches_and_gradients():
    """Points set before the mode are used; the handle switches between modes 2 and 3, and the
    gradient follows the mode: mode 3, then mode 2, then mode 3 again gives the first mode-3
    gradient bit for bit (the adjoint's difference table is rebuilt on every switch)."""
    from viscoelasticinversion_amd.configs import build_simulation
    cfg = small_cfg(nx=80, nz=48, nrelax=1, nt=30, fs="hermitian")
    cfg["hermitian_points"] = 4
    ref = run_oracle(cfg, np.float64)
    print("[case] hermitian points before mode, switch to afda and back", flush=True)
    sim = build_simulation(dict(cfg, free_surface=False, hermitian_points=None), "float64")
    sim.set_afda_hermitian(4)
    sim.set_free_surface("hermitian")
    sim.step(30)
    _same({f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz")}, 
          {f: ref[f] for f in ("vx", "vz", "sxx", "szz", "sxz")}, "points before mode")
    dobs = sim.traces(30) * 1.01
    sim.reset()
    J3, g3 = sim.gradient(dobs, 30)
    assert np.isfinite(J3) and J3 > 0
    ref2 = run_oracle(dict(cfg, free_surface="afda"), np.float64)
    sim.reset()
    sim.set_free_surface("afda")
    sim.step(30)
    _same({f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz")},
          {f: ref2[f] for f in ("vx", "vz", "sxx", "szz", "sxz")}, "back to afda")
    sim.reset()
    J, g = sim.gradient(dobs, 30)
    assert np.isfinite(J) and J > 0
    assert any(not np.array_equal(g[k], g3[k]) for k in g)
    sim.reset()
    sim.set_free_surface("hermitian")
    J3b, g3b = sim.gradient(dobs, 30)
    assert J3b == J3
    for k in g3:
        assert np.array_equal(g3b[k], g3[k]), k
    sim.close()
