"""H-AFDA free surface on the GPU: the HIP stepper against the CPU oracle on identical seeded
inputs (bar: relative L2 <= 1e-5 FP32, <= 1e-12 FP64; the paths share the operation order, so
the results are identical in practice), every tiling variant, VTI, custom point counts, and
shots with different free surfaces mixed in one batched launch, and the H-AFDA gradient's
independence of the checkpoint interval."""
import numpy as np
import pytest

from helpers import compare, run_gpu, run_oracle, small_cfg

pytestmark = pytest.mark.gpu


def _prec(dtype):
    return "float64" if dtype == np.float64 else "float32"


@pytest.mark.parametrize("rows", [1, 2, 4, 8])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_afda_every_tiling(rows, dtype):
    cfg = small_cfg(nx=150, nz=90, nrelax=3, nt=60, fs="afda")
    print(f"[case] afda tiling rows={rows} {_prec(dtype)}", flush=True)
    ref = run_oracle(cfg, dtype)
    gpu = run_gpu(cfg, _prec(dtype), tiling=rows)
    errs = compare(gpu, ref, dtype, f"afda tiling {rows}")
    print(f"[parity] afda rows={rows}: worst rel L2 = {max(errs.values()):.3e}", flush=True)
    # nothing above the surface is written
    for k in ("vx", "vz", "szz", "sxz"):
        assert not np.any(gpu[k][:4]), k
    assert np.all(gpu["szz"][4, 4:-4] == 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nrelax", [0, 2])
def test_afda_elastic_and_visco(dtype, nrelax):
    cfg = small_cfg(nx=96, nz=64, nrelax=nrelax, nt=80, fs="afda")
    print(f"[case] afda L={nrelax} {_prec(dtype)}", flush=True)
    compare(run_gpu(cfg, _prec(dtype)), run_oracle(cfg, dtype), dtype, f"afda L={nrelax}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_afda_vti(dtype):
    cfg = small_cfg(nx=100, nz=70, nrelax=2, nt=60, fs="afda", vti=True)
    print(f"[case] afda vti {_prec(dtype)}", flush=True)
    compare(run_gpu(cfg, _prec(dtype)), run_oracle(cfg, dtype), dtype, "afda vti")


@pytest.mark.parametrize("counts", [[8] * 13, [1, 2, 3, 4, 2, 3, 4, 2, 3, 4, 2, 3, 4]])
def test_afda_custom_points(counts):
    cfg = small_cfg(nx=90, nz=60, nrelax=2, nt=40, fs="afda")
    cfg["afda_points"] = counts
    print(f"[case] afda points {counts}", flush=True)
    for dtype in (np.float64, np.float32):
        compare(run_gpu(cfg, _prec(dtype)), run_oracle(cfg, dtype), dtype, "afda points")


def test_mixed_free_surfaces_in_one_batch():
    """Shots with no surface, imaging and H-AFDA advance in one launch (the H-AFDA variant)
    and each equals the same shot stepped alone and the oracle."""
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_simulation
    cfgs = [small_cfg(nx=70 + 23 * k, nz=50 + 7 * k, nrelax=2, nt=50, seed=k + 3, fs=fs)
            for k, fs in enumerate(("afda", True, False, "afda"))]
    print("[case] batch of 4 shots with mixed free surfaces", flush=True)
    sims = [build_simulation(c, "float32") for c in cfgs]
    b = Batch(sims)
    b.step(30)
    b.step(20)
    b.sync()
    for k, (c, sim) in enumerate(zip(cfgs, sims)):
        got = {f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz", "mxx", "mzz", "mxz")}
        got["traces"] = sim.traces(50)
        alone = run_gpu(c, "float32")
        for f in got:
            assert np.array_equal(got[f], alone[f]), (k, f)
        ref = run_oracle(c, np.float32)
        compare(got, {f: ref[f] for f in got}, np.float32, f"mixed batch shot {k}")
    b.close()
    for s in sims:
        s.close()


def test_afda_gradient_checkpoint_invariant():
    """The H-AFDA gradient (finite differences in test_gpu_gradient.py) does not depend on
    the checkpoint interval: recording recompute (1, 7), sqrt(nt) and store-all agree."""
    from viscoelasticinversion_amd.configs import build_simulation
    cfg = small_cfg(nx=80, nz=50, nrelax=2, nt=45, fs="afda")
    sim = build_simulation(cfg, "float64")
    sim.step(45)
    dobs = sim.traces(45) * 1.01 + 1e-3
    sim.close()
    print("[case] afda gradient checkpoint intervals", flush=True)
    res = []
    for K in (1, 7, 0, 45):
        sim = build_simulation(cfg, "float64")
        res.append(sim.gradient(dobs, 45, checkpoint_every=K))
        sim.close()
    J0, g0 = res[0]
    assert J0 > 0
    for J, g in res[1:]:
        assert J == J0
        for k in g0:
            assert np.array_equal(g[k], g0[k]), k


def test_afda_points_before_mode_and_switch_back():
    """Point counts given before the surface mode are used once mode 2 is set; switching the
    same handle back to imaging restores the imaging result."""
    from viscoelasticinversion_amd.configs import build_simulation
    counts = [2, 3, 4, 5, 2, 3, 4, 2, 3, 4, 2, 3, 4]
    cfg = small_cfg(nx=80, nz=48, nrelax=1, nt=30, fs="afda")
    cfg["afda_points"] = counts
    ref = run_oracle(cfg, np.float64)
    c0 = dict(cfg, free_surface=False, afda_points=None)
    print("[case] afda points before mode", flush=True)
    sim = build_simulation(c0, "float64")
    sim.set_afda_points(counts)
    sim.set_free_surface("afda")
    sim.step(30)
    got = {f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz")}
    compare(got, {f: ref[f] for f in got}, np.float64, "points before mode")
    cimg = dict(cfg, free_surface=True, afda_points=None)
    ref_img = run_oracle(cimg, np.float64)
    sim.reset()
    sim.set_free_surface(True)
    sim.step(30)
    got = {f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz")}
    compare(got, {f: ref_img[f] for f in got}, np.float64, "back to imaging")
    sim.close()
