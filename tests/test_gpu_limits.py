"""Size limits and empty inputs on the GPU.

* A padded field plane must stay below 2 GiB (kernels address planes with 32-bit byte
  offsets): a larger grid is refused before anything is allocated.
* At the largest size that fits (23000 x 23000 FP32, 2.13 GB per plane), a source near the far
  corner gives the same field as the same source in a small grid until the waves reach a
  boundary (homogeneous medium: the arithmetic of every cell is the same), so addressing at
  offsets near 2^31 bytes is exact.
* No sources and no receivers: an arbitrary initial state evolves as in the oracle.
"""
import numpy as np
import pytest

from helpers import compare, run_oracle, small_cfg

pytestmark = pytest.mark.gpu


def test_plane_over_2gib_is_refused():
    from viscoelasticinversion_amd import Simulation, _lib
    print("[case] 30000 x 20000 FP32 refused", flush=True)
    with pytest.raises(_lib.ViError) as e:
        Simulation(30000, 20000, 5.0, 5.0, 4e-4, "float32", 0)
    assert "2 GiB" in str(e.value)


def _homogeneous(sim, nx, nz):
    mu = 2000.0 * 1150.0 ** 2
    sim.set_model(2000.0, 2000.0 * 2000.0 ** 2 - 2 * mu, mu)


def test_largest_plane_translation_invariance():
    from viscoelasticinversion_amd import Simulation
    nt, dt, h = 40, 4e-4, 5.0
    t = np.arange(nt + 1) * dt
    w = (1e6 * np.sin(2 * np.pi * 30 * t) * np.exp(-((t - 0.006) / 0.003) ** 2))[None, :]
    w = w.astype(np.float32)
    half = 60  # window half-width; the waves travel nt * dt * 2000 / h = 6.4 cells
    out = {}
    for (nx, nz, sx, sz) in ((200, 200, 100, 100), (23000, 23000, 23000 - 100, 23000 - 100)):
        print(f"[case] translation invariance {nx} x {nz}", flush=True)
        sim = Simulation(nx, nz, h, h, dt, "float32", 0)
        _homogeneous(sim, nx, nz)
        sim.set_sources([sx], [sz], [1], w)
        sim.step(nt)
        win = {}
        for f in ("vx", "vz", "sxx", "szz", "sxz"):
            a = sim.get(f)
            win[f] = np.array(a[sz - half:sz + half, sx - half:sx + half])
            del a
        sim.close()
        out[(nx, nz)] = win
    small, big = out[(200, 200)], out[(23000, 23000)]
    for f in small:
        assert np.abs(small[f]).max() > 0, f
        assert np.array_equal(small[f], big[f]), f


def test_no_sources_no_receivers():
    cfg = small_cfg(nx=90, nz=50, nrelax=2, nt=30)
    for k in ("src_ix", "src_iz", "src_kind", "rec_ix", "rec_iz", "rec_comp"):
        cfg[k] = np.zeros(0, np.int32)
    cfg["wavelets"] = np.zeros((0, 35))
    rng = np.random.default_rng(4)
    init = {0: rng.standard_normal((50, 90)), 3: rng.standard_normal((50, 90)) * 1e3}
    for a in init.values():
        a[:4] = 0; a[-4:] = 0; a[:, :4] = 0; a[:, -4:] = 0
    print("[case] no sources, no receivers", flush=True)
    from helpers import run_gpu
    for dtype, prec in ((np.float64, "float64"), (np.float32, "float32")):
        ref = run_oracle(cfg, dtype, init={k: v.astype(dtype) for k, v in init.items()})
        gpu = run_gpu(cfg, prec, init={k: v.astype(dtype) for k, v in init.items()})
        ref.pop("traces"); gpu.pop("traces")
        compare(gpu, ref, dtype, f"empty lists {prec}")
