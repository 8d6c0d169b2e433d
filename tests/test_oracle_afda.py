"""Pins of the oracle's H-AFDA free surface (free_surface = "afda").

The scheme is the adjusted one-sided FD approximation of Kristek, Moczo & Archuleta (2002) that
the reference drafts in src/free-surface-kernels.jl:1-214 and src/common/Stencil.jl:52-105,
carried to the rows an 8th-order centred stencil needs: the surface row 4 holds sxx, szz = 0
and vx; vz and sxz sit half a cell below; no value above the surface is used.

What fixes the oracle here:
  * the one-sided weights are the exact rational Taylor (Vandermonde) solutions, equal bit for
    bit to the correctly rounded fractions;
  * one step from polynomial fields gives the exact derivative on every row near the surface
    (pins the row-to-stencil mapping, the stencil positions and the boundary values);
  * a Rayleigh wave runs along the surface at the speed of the Rayleigh equation;
  * the fused and multi-pass restatements agree bit for bit;
  * the scheme stays bounded over long runs in a layered medium with 2 % cell noise.
"""
from fractions import Fraction

import numpy as np
import pytest

import helpers
from helpers import small_cfg
from test_oracle_pins import _rayleigh_speed, _ricker, _uniform_model, _vandermonde_weights

F64 = np.float64
ALL8 = [8] * 13


def _exact_table(counts):
    half = [Fraction(2 * m + 1, 2) for m in range(8)]
    ints = [Fraction(m) for m in range(8)]
    rows = []
    for r, n in enumerate(counts):
        if r < 4:    # A_k: d sxz/dz at z = k h from sxz(0) = 0 and the half rows
            w = _vandermonde_weights([Fraction(0)] + half[:n], Fraction(r), 1)[1:]
        elif r < 7:  # B_k: d szz/dz at z = (k+1/2) h from the integer rows (szz(0) = 0 stored)
            w = _vandermonde_weights(ints[:n], Fraction(2 * (r - 4) + 1, 2), 1)
        elif r < 10:  # C_k: d vz/dz at z = k h, k = 1..3, from the half rows
            w = _vandermonde_weights(half[:n], Fraction(r - 6), 1)
        else:        # D_k: d vx/dz at z = (k+1/2) h from the integer rows
            w = _vandermonde_weights(ints[:n], Fraction(2 * (r - 10) + 1, 2), 1)
        rows.append(list(w) + [Fraction(0)] * (8 - len(w)))
    return rows


@pytest.mark.parametrize("counts", [None, ALL8, [1, 2, 3, 4, 2, 3, 4, 2, 3, 4, 2, 3, 4],
                                    [8, 7, 6, 5, 8, 5, 3, 8, 5, 3, 8, 6, 4]])
@pytest.mark.parametrize("h", [1.0, 5.0, 3.7])
def test_afda_weights_are_the_exact_taylor_solutions(oracle_mod, counts, h):
    got = oracle_mod.afda_weights(h, counts)
    table = _exact_table(counts or oracle_mod.afda_default_points())
    want = np.array([[float(x) for x in row] for row in table]) / h
    assert np.array_equal(got, want)


def test_afda_default_points_and_range(oracle_mod):
    assert oracle_mod.afda_default_points() == [3, 4, 6, 8, 3, 4, 6, 2, 4, 6, 3, 4, 6]
    for bad in ([0] + [4] * 12, [4] * 12 + [9], [4] * 4 + [1] + [4] * 8):
        with pytest.raises(ValueError):
            oracle_mod.afda_weights(1.0, bad)


def _poly(c, z):
    return sum(ck * z ** k for k, ck in enumerate(c))


def _dpoly(c, z):
    return sum(k * ck * z ** (k - 1) for k, ck in enumerate(c) if k)


@pytest.mark.parametrize("counts", [None, ALL8])
def test_afda_rows_exact_on_polynomials(oracle_mod, counts):
    """One step from z-polynomial fields: every row from the surface down gets the exact
    vertical derivative (the degrees are within each stencil's exactness)."""
    nx, nz, dx, dz, dt = 24, 40, 4.0, 5.0, 1e-4
    rho, vp, vs = 2000.0, 2000.0, 1150.0
    m = _uniform_model(nx, nz, rho=rho, vp=vp, vs=vs)
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    pi = lam + 2 * mu
    full = counts is not None
    rows = np.arange(nz)
    zi = (rows - 4) * dz          # sxx, szz, vx rows
    zh = (rows - 4 + 0.5) * dz    # vz, sxz rows
    above = rows < 4
    # velocity half: sxz(0) = szz(0) = 0
    p = [0.0, 3e3, -40.0, 0.3 if full else 0.2]   # cubic: A_k exact to degree >= 3
    q = [0.0, -2e3, 25.0] + ([0.1] if full else [])  # B_0 of the default is exact to degree 2
    o = oracle_mod.Oracle(nx, nz, dx, dz, dt, m, dtype=F64, free_surface="afda",
                          afda_points=counts)
    o.set(oracle_mod.SXZ, np.where(above, 0.0, _poly(p, zh))[:, None] * np.ones((1, nx)))
    o.set(oracle_mod.SZZ, np.where(above, 0.0, _poly(q, zi))[:, None] * np.ones((1, nx)))
    o.run(1)
    vx, vz = o.get(oracle_mod.VX), o.get(oracle_mod.VZ)
    o.close()
    sl = (slice(4, nz - 4), slice(4, nx - 4))
    want_vx = dt / rho * _dpoly(p, zi)[:, None] * np.ones((1, nx))
    want_vz = dt / rho * _dpoly(q, zh)[:, None] * np.ones((1, nx))
    assert np.allclose(vx[sl], want_vx[sl], rtol=1e-9, atol=1e-9 * np.abs(want_vx).max())
    assert np.allclose(vz[sl], want_vz[sl], rtol=1e-9, atol=1e-9 * np.abs(want_vz).max())
    # stress half: polynomial velocities, zero stress
    P = [0.3, -2e-3, 5e-5] + ([-2e-7] if full else [])  # D_0 of the default: degree 2
    Q = [-0.1, 4e-3] + ([3e-5, 1e-7] if full else [])   # C_1 of the default: degree 1
    o = oracle_mod.Oracle(nx, nz, dx, dz, dt, m, dtype=F64, free_surface="afda",
                          afda_points=counts)
    o.set(oracle_mod.VX, np.where(above, 0.0, _poly(P, zi))[:, None] * np.ones((1, nx)))
    o.set(oracle_mod.VZ, np.where(above, 0.0, _poly(Q, zh))[:, None] * np.ones((1, nx)))
    o.run(1)
    sxx, szz, sxz = (o.get(f) for f in (oracle_mod.SXX, oracle_mod.SZZ, oracle_mod.SXZ))
    o.close()
    ezz = _dpoly(Q, zi)[:, None] * np.ones((1, nx))
    ezz[4] = 0.0  # surface row: dszz = 0 with exx = 0
    ezx = _dpoly(P, zh)[:, None] * np.ones((1, nx))
    for got, want in ((sxx, dt * lam * ezz), (szz, dt * pi * ezz), (sxz, dt * mu * ezx)):
        assert np.allclose(got[sl], want[sl], rtol=1e-9, atol=1e-9 * np.abs(want).max())
    assert np.all(szz[4, 4:-4] == 0)


def test_afda_rayleigh_wave_speed(oracle_mod):
    """A vertical force just below the H-AFDA surface launches a Rayleigh wave whose speed
    along the surface matches the Rayleigh equation."""
    nx, nz, h, dt, f = 700, 110, 2.0, 2.5e-4, 25.0
    vp, vs = 2000.0, 2000.0 / np.sqrt(3.0)
    m = _uniform_model(nx, nz, vp=vp, vs=vs)
    nt = 3000
    w = _ricker(f, np.arange(nt) * dt - 1.2 / f)[None, :]
    T = 30
    d0 = -(3) * vp * np.log(1e-5) / (2 * T * h)
    cp = dict(thickness=T, d0=d0, alpha_max=np.pi * f, kappa_max=1.0, power=2.0, left=True,
              right=True, top=False, bottom=True)
    o = oracle_mod.Oracle(nx, nz, h, h, dt, m, dtype=F64, free_surface="afda", cpml=cp)
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
    # nothing above the surface is written
    for f_ in (oracle_mod.SZZ, oracle_mod.SXZ, oracle_mod.VX, oracle_mod.VZ):
        assert not np.any(o.get(f_)[:4])
    o.close()


@pytest.mark.parametrize("nrelax", [0, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_afda_fused_and_multipass_agree_bitwise(oracle_mod, nrelax, dtype):
    cfg = small_cfg(nrelax=nrelax, nt=50, fs="afda")
    a = helpers.run_oracle(cfg, dtype, mode=0)
    b = helpers.run_oracle(cfg, dtype, mode=1)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_afda_bounded_in_layered_noisy_medium(oracle_mod):
    """Elastic (no attenuation to hide growth), 8 layers with 2 % cell noise, Dirichlet
    walls: the wavefield norm after 20000 steps stays within a few percent of its value
    after 2000 (the energy is trapped; the surface neither creates nor destroys it)."""
    nx, nz, dx, dz, dt = 100, 60, 4.0, 5.0, 4e-4
    rng = np.random.default_rng(7)
    layer = np.minimum((np.arange(nz) * 8) // nz, 7)
    vp_l = np.linspace(1800.0, 3200.0, 8)
    vp = vp_l[layer][:, None] * np.ones((1, nx)) * (1 + 0.02 * rng.standard_normal((nz, nx)))
    vs = vp / 1.8
    rho = 310.0 * vp ** 0.25 * (1 + 0.02 * rng.standard_normal((nz, nx)))
    mu = rho * vs ** 2
    model = dict(rho=rho, lam=rho * vp ** 2 - 2 * mu, mu=mu)
    o = oracle_mod.Oracle(nx, nz, dx, dz, dt, model, dtype=F64, free_surface="afda")
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
