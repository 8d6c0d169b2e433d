"""Pins of the oracle's H-AFDA formula #3 (free_surface = "hermitian").

Formula #3 of Kristek, Moczo & Archuleta (2002), as the reference builds it
(src/common/Stencil.jl:52-105 and _stencil_hermitian, Stencil.jl:119-189) and applies it
(src/free-surface-kernels.jl:154-214): d vz/dz at z = h (row 5) from the surface derivative
vz,z(0) and vz at the half points (m+1/2) h,
    vz,z(h) = alpha vz,z(0) + sum_m beta_m vz((m+1/2) h),
with vz,z(0) = -lam/(lam + 2 mu) vx,x(0) from szz(0) = 0 (the memory terms added when L > 0,
-C13/C33 for VTI). Every other row is the H-AFDA mode's.

What fixes the oracle here:
  * alpha and beta are the exact rational solution of the reference's augmented Vandermonde
    system (derivative column at 0), equal bit for bit to the correctly rounded fractions;
  * from a polynomial vz whose slope at z = 0 is the surface value, row 5 gets the exact
    derivative (any degree up to n), and with a cubic every row below does, in isotropic,
    VTI and L = 3 media with memory variables; a wrong alpha, sign, surface ratio or row
    fails it;
  * a Rayleigh wave runs along the surface at the speed of the Rayleigh equation;
  * the fused and multi-pass restatements agree bit for bit;
  * the default point count stays bounded over long runs in a layered medium with cell noise.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import helpers
from helpers import small_cfg
from test_oracle_pins import _rayleigh_speed, _ricker, _uniform_model

F64 = np.float64


def _hermitian_exact(x, xd, x0):
    """Exact rational solve of _stencil_hermitian's system (Stencil.jl:136-189) for one first
    derivative known at xd: column 0 holds (xd - x0)^(l-1)/(l-1)! (l >= 1), column 1 + k holds
    (x_k - x0)^l / l!, right-hand side e_1. Returns [alpha, beta_0, ...]."""
    n = len(x) + 1
    A = [[Fraction(0)] * n for _ in range(n)]
    for l in range(n):
        if l >= 1:
            A[l][0] = Fraction(xd - x0) ** (l - 1) / math.factorial(l - 1)
        for k, xk in enumerate(x):
            A[l][1 + k] = Fraction(xk - x0) ** l / math.factorial(l)
    M = [A[r] + [Fraction(int(r == 1))] for r in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        piv = M[c][c]
        M[c] = [v / piv for v in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [a - f * b for a, b in zip(M[r], M[c])]
    return [M[r][n] for r in range(n)]


@pytest.mark.parametrize("n", range(2, 9))
@pytest.mark.parametrize("h", [1.0, 5.0, 3.7])
def test_hermitian_weights_are_the_exact_solution(oracle_mod, n, h):
    w = _hermitian_exact([Fraction(2 * m + 1, 2) for m in range(n)], Fraction(0), Fraction(1))
    got = oracle_mod.afda_hermitian(h, n)
    want = np.zeros(9)
    want[0] = float(w[0])
    want[1:1 + n] = [float(b) / h for b in w[1:]]
    assert np.array_equal(got, want)
    # plausible mistakes: the derivative point at the target, or alpha = 0 (n > 2)
    if n > 2:
        assert got[0] != 0.0


def test_hermitian_weights_known_values_and_range(oracle_mod):
    # n = 2: vz,z(0) adds nothing to the centred difference (alpha = 0, beta = -1, 1);
    # n = 3: alpha = -1/23, beta = (-25/23, 26/23, -1/23)
    assert np.array_equal(oracle_mod.afda_hermitian(1.0, 2)[:3], [0.0, -1.0, 1.0])
    w3 = oracle_mod.afda_hermitian(1.0, 3)
    assert np.array_equal(w3[:4], [-1 / 23, -25 / 23, 26 / 23, -1 / 23])
    for bad in (0, 1, 9):
        with pytest.raises(ValueError):
            oracle_mod.afda_hermitian(1.0, bad)


def _poly(c, z):
    return sum(ck * z ** k for k, ck in enumerate(c))


def _dpoly(c, z):
    return sum(k * ck * z ** (k - 1) for k, ck in enumerate(c) if k)


def _surface_case(oracle_mod, kind, n, coefs, nrelax=0):
    """Oracle with vx = a x everywhere and vz(z) = es z + coefs (z from the surface, at the
    vz rows); returns (ezz of the rates on every row, the surface value es)."""
    nx, nz, dx, dz, dt = 24, 40, 4.0, 5.0, 1e-4
    rho, vp, vs = 2000.0, 2000.0, 1150.0
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    if kind == "vti":
        c33 = rho * vp ** 2
        c13 = 0.7 * lam
        model = dict(rho=np.full((nz, nx), rho), c11=np.full((nz, nx), 1.2 * c33),
                     c13=np.full((nz, nx), c13), c33=np.full((nz, nx), c33),
                     c55=np.full((nz, nx), mu))
        lr, pz = c13, c33
    else:
        model = _uniform_model(nx, nz, rho=rho, vp=vp, vs=vs)
        lr, pz = lam, lam + 2 * mu
    taun = None
    tau = 0.0
    if nrelax:
        tau = 0.05
        model["tau"] = np.full((nz, nx), tau)
        taun = np.array([1e-3, 1e-2, 1e-1])
    a = 2.5e-3
    x = np.arange(nx) * dx
    rows = np.arange(nz)
    zh = (rows - 4 + 0.5) * dz
    o = oracle_mod.Oracle(nx, nz, dx, dz, dt, model, dtype=F64, nrelax=nrelax, taun=taun,
                          free_surface="hermitian", hermitian_points=n)
    SMxx = SMzz = 0.0
    if nrelax:
        mxx = np.stack([np.full((nz, nx), v) for v in (1e-4, -2e-4, 3e-4)])
        mzz = np.stack([np.full((nz, nx), v) for v in (-5e-4, 1e-4, 2e-4)])
        o.set(oracle_mod.MXX, mxx)
        o.set(oracle_mod.MZZ, mzz)
        SMxx, SMzz = float(mxx[:, 0, 0].sum()), float(mzz[:, 0, 0].sum())
    # the surface value: dszz = 0 on row 4 (elastic: -lam/(lam + 2 mu) a)
    l0, pz0 = lr * (1 + tau), pz * (1 + tau)
    es = -((l0 * a + lr * tau * SMxx) + pz * tau * SMzz) / pz0
    o.set(oracle_mod.VX, np.ones((nz, 1)) * (a * x)[None, :])
    c = [0.0, es] + list(coefs)
    o.set(oracle_mod.VZ, np.where(rows < 4, 0.0, _poly(c, zh))[:, None] * np.ones((1, nx)))
    _, ds, dm = o.rates(None)
    o.close()
    # ezz from dszz = l0 exx + pz0 ezz + l1 SMxx + pz1 SMzz
    ezz = (ds[1] - (l0 * a + lr * tau * SMxx + pz * tau * SMzz)) / pz0
    zi = (rows - 4) * dz
    want = _dpoly(c, zi)[:, None] * np.ones((1, nx))
    return ezz, want, es


@pytest.mark.parametrize("n", range(2, 9))
def test_hermitian_row_exact_to_degree_n(oracle_mod, n):
    """Row 5's d vz/dz is exact for a vz of degree n whose slope at the surface is the
    surface value (the only information formula #3 adds)."""
    coefs = [(-1) ** k * 10.0 ** (-3 - k) for k in range(n - 1)]  # z^2 .. z^n
    ezz, want, es = _surface_case(oracle_mod, "iso", n, coefs)
    assert es != 0.0
    got, w = ezz[5, 4:-4], want[5, 4:-4]
    assert np.allclose(got, w, rtol=1e-9, atol=1e-12), (got[0], w[0])


@pytest.mark.parametrize("kind,nrelax", [("iso", 0), ("vti", 0), ("iso", 3), ("vti", 3)])
def test_hermitian_rows_exact_on_cubics(oracle_mod, kind, nrelax):
    """With a cubic vz (slope es at z = 0), every row from 5 down is exact at the default
    count: row 5 by formula #3, rows 6, 7 by the H-AFDA rows, the rest centred. The surface
    value carries lam/(lam + 2 mu) (C13/C33 for VTI) and, with L = 3, the memory terms."""
    ezz, want, es = _surface_case(oracle_mod, kind, None, [3e-3, -1e-4], nrelax)
    sl = (slice(5, 36), slice(4, 20))
    assert np.allclose(ezz[sl], want[sl], rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_hermitian_rayleigh_wave_speed(oracle_mod):
    nx, nz, h, dt, f = 700, 110, 2.0, 2.5e-4, 25.0
    vp, vs = 2000.0, 2000.0 / np.sqrt(3.0)
    m = _uniform_model(nx, nz, vp=vp, vs=vs)
    nt = 3000
    w = _ricker(f, np.arange(nt) * dt - 1.2 / f)[None, :]
    T = 30
    d0 = -(3) * vp * np.log(1e-5) / (2 * T * h)
    cp = dict(thickness=T, d0=d0, alpha_max=np.pi * f, kappa_max=1.0, power=2.0, left=True,
              right=True, top=False, bottom=True)
    o = oracle_mod.Oracle(nx, nz, h, h, dt, m, dtype=F64, free_surface="hermitian", cpml=cp)
    o.set_sources([60], [5], [0], w)
    r1, r2 = 160, 400
    o.set_receivers([r1, r2], [4, 4], [1, 1])
    tr = o.run(nt)
    a, b = tr[:, 0], tr[:, 1]
    cc = np.correlate(b, a, mode="full")
    k = int(np.argmax(cc))
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    lag = (k - (len(a) - 1)) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    c_meas = (r2 - r1) * h / (lag * dt)
    assert abs(c_meas / _rayleigh_speed(vp, vs) - 1) < 0.02, c_meas
    assert np.all(o.get(oracle_mod.SZZ)[4, 4:-4] == 0)
    for f_ in (oracle_mod.SZZ, oracle_mod.SXZ, oracle_mod.VX, oracle_mod.VZ):
        assert not np.any(o.get(f_)[:4])
    o.close()


@pytest.mark.parametrize("nrelax", [0, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_hermitian_fused_and_multipass_agree_bitwise(oracle_mod, nrelax, dtype):
    cfg = small_cfg(nrelax=nrelax, nt=50, fs="hermitian", cpml=True)
    a = helpers.run_oracle(cfg, dtype, mode=0)
    b = helpers.run_oracle(cfg, dtype, mode=1)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_hermitian_differs_from_afda_only_below_the_surface(oracle_mod):
    """One stress half step: formula #3 changes row 5 only (n > 2); n = 2 equals H-AFDA's
    2-point row bit for bit."""
    cfg = small_cfg(nrelax=0, nt=1, fs="afda")
    rng = np.random.default_rng(4)
    init = {k: rng.standard_normal((cfg["nz"], cfg["nx"])) for k in (0, 1)}
    base = helpers.run_oracle(cfg, F64, init=init)
    for n, same in ((2, True), (5, False)):
        c = dict(cfg, free_surface="hermitian", hermitian_points=n)
        got = helpers.run_oracle(c, F64, init=init)
        for k in ("sxx", "szz"):
            diff = np.nonzero(np.any(got[k] != base[k], axis=1))[0]
            assert list(diff) == ([] if same else [5]), (n, k, diff)


def test_hermitian_refuses_bad_counts_and_top_cpml(oracle_mod):
    m = _uniform_model(40, 40)
    for n in (1, 9):
        with pytest.raises(ValueError):
            oracle_mod.Oracle(40, 40, 5.0, 5.0, 1e-4, m, free_surface="hermitian",
                              hermitian_points=n)
    cp = dict(thickness=8, d0=100.0, top=True)
    with pytest.raises(ValueError):
        oracle_mod.Oracle(40, 40, 5.0, 5.0, 1e-4, m, free_surface="hermitian", cpml=cp)


def test_hermitian_bounded_in_layered_noisy_medium(oracle_mod):
    """Elastic, 8 layers with 2 % cell noise, Dirichlet walls, the default point count: the
    wavefield norm after 20000 steps stays within a few percent of its value after 2000."""
    nx, nz, dx, dz, dt = 100, 60, 4.0, 5.0, 4e-4
    rng = np.random.default_rng(7)
    layer = np.minimum((np.arange(nz) * 8) // nz, 7)
    vp_l = np.linspace(1800.0, 3200.0, 8)
    vp = vp_l[layer][:, None] * np.ones((1, nx)) * (1 + 0.02 * rng.standard_normal((nz, nx)))
    vs = vp / 1.8
    rho = 310.0 * vp ** 0.25 * (1 + 0.02 * rng.standard_normal((nz, nx)))
    mu = rho * vs ** 2
    model = dict(rho=rho, lam=rho * vp ** 2 - 2 * mu, mu=mu)
    o = oracle_mod.Oracle(nx, nz, dx, dz, dt, model, dtype=F64, free_surface="hermitian")
    v0 = np.zeros((nz, nx))
    v0[4:-4, 4:-4] = np.random.default_rng(3).standard_normal((nz - 8, nx - 8))
    o.set(oracle_mod.VX, v0)
    o.set(oracle_mod.VZ, np.roll(v0, 3, axis=1))
    norms = []
    for _ in range(10):
        o.run(2000)
        norms.append(sum(float(np.linalg.norm(o.get(f))) for f in (oracle_mod.VX, oracle_mod.VZ)))
    o.close()
    assert np.all(np.isfinite(norms))
    assert max(norms[1:]) < 1.05 * norms[0], norms
