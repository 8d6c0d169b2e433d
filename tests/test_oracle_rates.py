"""The oracle's right-hand side (the stress_eq! / velocity_eq! analogue,
ViscoelasticProblem.jl:62-174) is pinned by the step it belongs to: from any state, the
velocity half step equals v + dt * dv, and the stress half step from the new velocities equals
s + dt * ds and M + dt * dM, bit for bit, in FP32 and FP64, with every feature (CPML with
kappa and alpha, stress imaging, H-AFDA or formula #3 surface, VTI, L = 0 and 3, force and explosive
sources, two sources on one cell), on both oracle code paths (fused and multi-pass)."""
import numpy as np
import pytest

from helpers import make_oracle, small_cfg

FIELDS = ("vx", "vz", "sxx", "szz", "sxz")


def _state(o, L, T):
    st = {k: o.get(k) for k in range(5)}
    if L:
        for k in (5, 6, 7):
            st[k] = o.get(k)
    if T:
        for k in range(8, 16):
            st[k] = o.get(k)
    return st


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", ["image_L3", "afda_L3", "herm_L3", "none_L0",
                                  "vti_image_L0", "vti_afda_L3", "vti_herm_L0"])
def test_rates_reproduce_the_step(case, dtype, mode):
    fs = {"image": True, "afda": "afda", "herm": "hermitian", "none": False}[case.split("_")[-2]]
    L = int(case[-1])
    cfg = small_cfg(nx=70, nz=48, nrelax=L, cpml=True, fs=bool(fs), seed=len(case) + L,
                    nt=12, vti=case.startswith("vti"))
    cfg["free_surface"] = fs
    cfg["cpml"]["kappa_max"] = 1.8
    # a second source on the explosive cell
    cfg["src_ix"] = np.append(cfg["src_ix"], cfg["src_ix"][1])
    cfg["src_iz"] = np.append(cfg["src_iz"], cfg["src_iz"][1])
    cfg["src_kind"] = np.append(cfg["src_kind"], 0)
    cfg["wavelets"] = np.vstack([cfg["wavelets"], 0.5 * cfg["wavelets"][0]])
    rng = np.random.default_rng(3)
    o = make_oracle(cfg, dtype)
    nz, nx, T = cfg["nz"], cfg["nx"], cfg["cpml"]["thickness"]
    for k in range(5):
        o.set(k, rng.standard_normal((nz, nx)))
    if L:
        for k in (5, 6, 7):
            o.set(k, rng.standard_normal((L, nz, nx)))
    for k in range(8, 16):
        o.set(k, rng.standard_normal(o.get(k).shape))
    o.run(3, mode=mode)  # consistent images and surface row; step index 3
    n = o.step_index
    s0 = _state(o, L, T)
    sv = cfg["wavelets"][:, n].astype(dtype)
    dt = dtype(cfg["dt"])
    dv, _, _ = o.rates(sv)
    o.run(1, mode=mode)
    s1 = _state(o, L, T)
    for k, f in ((0, dv[0]), (1, dv[1])):
        assert np.array_equal(s1[k], s0[k] + dt * f), FIELDS[k]
    # stress half from the new velocities, the old stresses, memory variables and CPML memory
    o2 = make_oracle(cfg, dtype)
    for k, v in s0.items():
        o2.set(k, s1[k] if k in (0, 1) else v)
    _, ds, dm = o2.rates(sv)
    for k in range(3):
        assert np.array_equal(s1[2 + k], s0[2 + k] + dt * ds[k]), FIELDS[2 + k]
    if L:
        for k in range(3):
            assert np.array_equal(s1[5 + k], s0[5 + k] + dt * dm[k]), ("M", k)
    assert np.abs(dv).max() > 0 and np.abs(ds).max() > 0
    o.close()
    o2.close()


def test_rates_source_terms():
    """Force sources enter dv before the buoyancy, explosive ones ds; the state is unchanged."""
    cfg = small_cfg(nx=40, nz=40, nrelax=0, cpml=False, fs=False, nt=4)
    cfg["src_ix"] = np.array([20, 12], np.int32)
    cfg["src_iz"] = np.array([15, 25], np.int32)
    cfg["src_kind"] = np.array([0, 1], np.int32)
    cfg["wavelets"] = np.ones((2, 4))
    o = make_oracle(cfg, np.float64)
    dv, ds, dm = o.rates([3.0, 5.0])
    rho = cfg["model"]["rho"]
    assert dm is None
    assert dv[1, 15, 20] == (2.0 / (rho[15, 20] + rho[16, 20])) * 3.0
    assert np.count_nonzero(dv) == 1
    assert ds[0, 25, 12] == 5.0 and ds[1, 25, 12] == 5.0 and np.count_nonzero(ds) == 2
    assert all(np.count_nonzero(o.get(k)) == 0 for k in range(5))
    dv0, ds0, _ = o.rates(None)
    assert np.count_nonzero(dv0) == 0 and np.count_nonzero(ds0) == 0
    o.close()
