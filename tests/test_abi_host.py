"""CPU tests of the C ABI library (no compute calls) and of the host-side API logic.

Ported known-answer tests from the reference's own suites:
  tests/test_source.jl:8-64, 142-166   Ricker values and source windows
  tests/test_cpml.jl:7-94              CPMLConfig / CPMLCoefficients / FreeSurfaceBC
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vi_mi355x.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vi_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    from viscoelasticinversion_amd import _lib
    lib = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def _julia_ccalls():
    """(name, argument types) of every ccall in julia/MI355X.jl."""
    text = open(os.path.join(ROOT, "julia", "MI355X.jl")).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(vi_[a-z0-9_]+),\s*LIB\),\s*[A-Za-z{}]+,\s*\(", text):
        depth, j, cur, args = 1, m.end(), "", []
        while depth:
            c = text[j]
            depth += c in "({"
            depth -= c in ")}"
            if depth == 1 and c == ",":
                args.append(cur.strip())
                cur = ""
            elif depth:
                cur += c
            j += 1
        if cur.strip():
            args.append(cur.strip())
        calls.append((m.group(1), args))
    return calls


def test_julia_shim_binds_declared_symbols():
    """Julia is not in this image, so the shim's ccalls are checked statically: each names a
    function the header declares and the library exports, with its number of arguments."""
    from viscoelasticinversion_amd import _lib
    calls = _julia_ccalls()
    assert len(calls) >= 20
    declared = set(_declared_functions())
    lib = _lib.lib()
    for name, args in calls:
        assert name in declared and hasattr(lib, name), name
        assert len(args) == len(_lib.SIGNATURES[name][1]), (name, args)


def test_fd_coeffs_match_oracle(oracle_mod):
    from viscoelasticinversion_amd import _lib
    for h in (1.0, 5.0, 2.5, 0.1):
        out = (ctypes.c_double * 4)()
        _lib.lib().vi_fd_coeffs(h, out)
        assert np.array_equal(np.array(out[:]), oracle_mod.fd_coeffs(h))


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("axis", [0, 1])
def test_cpml_profile_matches_oracle_bitwise(oracle_mod, axis, precision):
    """The coefficients the kernels use (computed in the handle's precision) equal the
    oracle's in both precisions."""
    from viscoelasticinversion_amd import _lib
    n = 120
    c = _lib.CpmlDesc(20, 1234.5, 31.4, 1.7, 2.0, 1, 0, 1, 1)
    a, b, k = (np.zeros(n) for _ in range(3))
    P = ctypes.POINTER(ctypes.c_double)
    _lib.check(_lib.lib().vi_cpml_profile(n, ctypes.byref(c), 4e-4, axis, precision,
                                          a.ctypes.data_as(P), b.ctypes.data_as(P),
                                          k.ctypes.data_as(P)))
    lo, hi = (1, 0) if axis == 0 else (1, 1)
    dt = np.float32 if precision == 0 else np.float64
    ref = oracle_mod.cpml_axis(n, 20, 1234.5, 31.4, 1.7, 2.0, 4e-4, lo, hi, dtype=dt)
    assert np.array_equal(a, ref["a"].astype(np.float64))
    assert np.array_equal(b, ref["b"].astype(np.float64))
    assert np.array_equal(k, ref["kappa"].astype(np.float64))


@pytest.mark.parametrize("counts", [None, [8] * 13, [1, 2, 3, 4, 2, 3, 4, 2, 3, 4, 2, 3, 4]])
def test_afda_weights_match_oracle_bitwise(oracle_mod, counts):
    """The library's H-AFDA table (host code, no device) equals the oracle's, which is pinned
    against the exact Taylor solutions; the default point counts agree too."""
    from viscoelasticinversion_amd import stepper
    assert stepper.afda_default_points() == oracle_mod.afda_default_points()
    for h in (1.0, 2.5, 5.0, 7.3):
        assert np.array_equal(stepper.afda_weights(h, counts), oracle_mod.afda_weights(h, counts))


def test_afda_weights_reject_bad_counts():
    from viscoelasticinversion_amd import _lib, stepper
    for bad in ([0] + [4] * 12, [4] * 12 + [9], [4] * 4 + [1] + [4] * 8):
        with pytest.raises(_lib.ViError):
            stepper.afda_weights(1.0, bad)
    with pytest.raises(ValueError):
        stepper.free_surface_mode("lamb")
    assert [stepper.free_surface_mode(m) for m in (False, True, "none", "image", "afda")] == [0, 1, 0, 1, 2]


@pytest.mark.parametrize("points", range(2, 9))
def test_hermitian_weights_match_oracle_bitwise(oracle_mod, points):
    """The library's formula #3 weights (host code, no device) equal the oracle's, which are
    pinned against the exact solution of the reference's augmented Vandermonde system."""
    from viscoelasticinversion_amd import stepper
    for h in (1.0, 2.5, 5.0, 7.3):
        assert np.array_equal(stepper.afda_hermitian_weights(h, points),
                              oracle_mod.afda_hermitian(h, points))


def test_hermitian_weights_reject_bad_counts():
    from viscoelasticinversion_amd import _lib, stepper
    for bad in (0, 1, 9):
        with pytest.raises(_lib.ViError):
            stepper.afda_hermitian_weights(1.0, bad)
    assert stepper.free_surface_mode("hermitian") == 3


def test_cpml_auto_d0_matches_config():
    from viscoelasticinversion_amd import CPMLConfig, _lib
    lib = _lib.lib()
    d0 = lib.vi_cpml_auto_d0(10, 3000.0, 10.0, 1e-5, 2.0)
    cfg = CPMLConfig(thickness=10, vmax=3000.0, dx=10.0, reflection_coef=1e-5)
    assert d0 == pytest.approx(cfg.d0, rel=1e-15)
    assert d0 == pytest.approx(-(3.0) * 3000.0 * np.log(1e-5) / (2 * 10 * 10.0))


def test_create_without_device_or_with_bad_args_fails_loudly():
    from viscoelasticinversion_amd import Simulation, _lib
    lib = _lib.lib()
    ndev = lib.vi_device_count()
    h = ctypes.c_void_p()
    bad = _lib.GridDesc(8, 8, 1.0, 1.0, 1e-3, 0, 0, 8, 8)
    assert lib.vi_create(ctypes.byref(bad), ctypes.byref(h)) == _lib.VI_EINVAL
    assert b"nx and nz" in lib.vi_last_error()
    for ox, oz in ((3, 8), (8, 10), (0, 8), (8, -2)):  # orders must be 2, 4, 6 or 8
        bad = _lib.GridDesc(32, 32, 1.0, 1.0, 1e-3, 0, 0, ox, oz)
        assert lib.vi_create(ctypes.byref(bad), ctypes.byref(h)) == _lib.VI_EINVAL
        assert b"fd_order" in lib.vi_last_error()
    bad = _lib.GridDesc(32, 32, 1.0, 1.0, 1e-3, 0, 7, 8, 8)
    assert lib.vi_create(ctypes.byref(bad), ctypes.byref(h)) == _lib.VI_EINVAL
    if ndev == 0:
        good = _lib.GridDesc(32, 32, 1.0, 1.0, 1e-3, 0, 0, 8, 8)
        assert lib.vi_create(ctypes.byref(good), ctypes.byref(h)) == _lib.VI_EDEVICE
        with pytest.raises(_lib.ViError):
            Simulation(32, 32, 1.0, 1.0, 1e-3)
    assert lib.vi_step(None, 1) == _lib.VI_EINVAL


# ---------------------------------------------------------------- reference KATs (source)
def test_ricker_values():
    """tests/test_source.jl:8-31."""
    from viscoelasticinversion_amd import _ricker
    assert _ricker(25.0, 0.0) == pytest.approx(1.0)
    assert _ricker(25.0, 0.01) == pytest.approx(_ricker(25.0, -0.01))
    assert abs(_ricker(25.0, 0.1)) < 0.01
    assert _ricker(50.0, 0.0) == pytest.approx(1.0)


def test_ricker_source_windows():
    """tests/test_source.jl:33-64, 142-166 and test_viscoelastic_problem.jl:106-122."""
    from viscoelasticinversion_amd import RickerSource
    s = RickerSource(25.0, 0.1, 10, 15)
    assert s.grid == (10, 15) and s.coefs == 1.0
    assert s.tspan[0] < 0.1 < s.tspan[1]
    assert abs((s.tspan[1] - s.tspan[0]) - 2.0 / 25.0) < 0.1
    assert RickerSource(1.0, 0.1, 4, 6).tspan[1] - RickerSource(1.0, 0.1, 4, 6).tspan[0] > 2.0
    hi = RickerSource(100.0, 0.1, 4, 6)
    assert hi.tspan[1] - hi.tspan[0] < 0.1
    assert s.value(0.1 + 0.1) == 0.0  # closed after tc + 0.1 at 25 Hz
    assert s.value(s.tspan[0] - 0.01) == 0.0 and s.value(s.tspan[1] + 0.01) == 0.0
    assert s.value(0.1) == pytest.approx(1.0)
    assert s.value(0.11) == pytest.approx(float(np.squeeze((1 - 2 * (np.pi * 25 * 0.01) ** 2)
                                                           * np.exp(-(np.pi * 25 * 0.01) ** 2))))
    du = np.zeros((15 + 1, 10 + 1), np.float32)
    s(du, 0.1)
    s(du, 0.1)
    assert np.count_nonzero(du) == 1 and du[15, 10] == pytest.approx(2.0)


def test_source_samples_times():
    from viscoelasticinversion_amd import RickerSource
    f = RickerSource(10.0, 0.1, 5, 5, kind="force_z")
    e = RickerSource(10.0, 0.1, 5, 5, kind="explosive")
    dt = 1e-3
    assert f.samples(300, dt)[100] == pytest.approx(1.0)  # t = 100 dt = tc
    np.testing.assert_allclose(e.samples(300, dt), e.value((np.arange(300) + 0.5) * dt))


# ---------------------------------------------------------------- reference KATs (CPML)
def test_cpml_config():
    """tests/test_cpml.jl:7-26."""
    from viscoelasticinversion_amd import CPMLConfig
    c = CPMLConfig(thickness=10, d0=100.0)
    assert (c.thickness, c.d0, c.left, c.right, c.top, c.bottom) == (10, 100.0, True, True, False, True)
    c = CPMLConfig(thickness=10, vmax=3000.0, dx=10.0, reflection_coef=1e-5)
    assert c.d0 > 0 and c.thickness == 10
    with pytest.raises(ValueError, match="Must provide either d0"):
        CPMLConfig(thickness=10)


def test_free_surface_bc():
    """tests/test_cpml.jl:71-77."""
    from viscoelasticinversion_amd import FreeSurfaceBC
    assert FreeSurfaceBC().depth == 1
    assert FreeSurfaceBC(2).depth == 2


def test_parameters_defaults_and_errors():
    """Parameters.jl:47-86 keyword defaults; unsupported options raise."""
    from viscoelasticinversion_amd import CPMLConfig, Parameters
    p = Parameters()
    assert (p.Nx, p.Nz, p.dx, p.dz, p.dt, p.fd_order_x) == (64, 64, 10.0, 10.0, 0.001, 8)
    assert p.source.freq == 40.0 and p.source.tc == 0.2 and p.source.grid == (32, 32)
    assert np.all(p.model.b == 1 / 2000.0) and np.all(p.model.lam == 1e9)
    p = Parameters(Nx=32, Nz=32, dt=0.001, cpml=CPMLConfig(thickness=5, d0=100.0),
                   free_surface=True)
    assert p.cpml is not None and p.free_surface_bc.depth == 1
    with pytest.raises(NotImplementedError):
        Parameters(bc="periodic")
    with pytest.raises(ValueError):
        Parameters(bc="bogus")
    assert Parameters(fd_order_x=4, fd_order_z=6).fd_order_z == 6
    with pytest.raises(NotImplementedError):
        Parameters(fd_order_x=10)
    with pytest.raises(ValueError):
        Parameters(device="cpu")


def test_make_problem_shapes():
    """test_viscoelastic_problem.jl:44-61: zero initial state, source kept on the problem."""
    from viscoelasticinversion_amd import Parameters, RickerSource, make_problem
    src = RickerSource(25.0, 0.1, 16, 16)
    p = Parameters(Nx=32, Nz=32, source=src)
    prob = make_problem(p, tspan=(0.0, 0.5))
    assert prob.u0[0].shape == (3, 32, 32) and prob.u0[1].shape == (2, 32, 32)
    assert np.all(prob.u0[0] == 0) and np.all(prob.u0[1] == 0)
    assert prob.p.source is src and prob.tspan == (0.0, 0.5)


def test_qmodel_fixture_band():
    """Models.jl:65-107: tau = N/(Kbar Q0) and Q(w) = N/(tau K(w)); with the stored 2-40 Hz,
    L=3 relaxation times Q stays within 8 % of Q0 = 50 over the band."""
    from viscoelasticinversion_amd import QModel
    from viscoelasticinversion_amd.configs import TAUN_2_40HZ_L3
    q = QModel(2.0, 40.0, 8, 8, Q0=50, N=3, taun=TAUN_2_40HZ_L3)
    w = 2 * np.pi * np.linspace(2.0, 40.0, 100)
    Q = q.Q(w)
    assert np.all(np.abs(Q / 50 - 1) < 0.08)
    assert np.mean(1.0 / Q) == pytest.approx(1.0 / 50, rel=0.01)
    q.update(np.full((8, 8), 100.0))
    assert np.allclose(q.tau, 3 / (q.K * 100.0))
