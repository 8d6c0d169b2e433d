"""Staggered operators of order 2, 4, 6 and 8 (the reference's Stencil(xorder, zorder, dx, dz;
stagger), src/common/Stencil.jl:280-320, with _stencil, Stencil.jl:9-14), on the oracle and on
the library's host side.

What fixes them:
  * the weights are the exact rational solution of _stencil's Vandermonde system at the
    staggered nodes, equal bit for bit to the correctly rounded fractions, in the oracle
    (closed form) and in the library (Lagrange derivative): two derivations that agree;
  * one velocity half step from polynomial stresses is exact for degree 2K (order 2K, x and z
    separately, with different orders per axis) and not for degree 2K + 1, so a wrong order
    or a swapped axis fails;
  * the von Neumann bound stable_dt of each order separates bounded from growing runs;
  * invalid orders and H-AFDA with a z order other than 8 are refused.
The Dirichlet ring stays 4 cells for every order (the reference pads by order/2)."""
import math
from fractions import Fraction

import numpy as np
import pytest

from test_oracle_pins import _uniform_model, _vandermonde_weights

ORDERS = (2, 4, 6, 8)


def _exact(order):
    K = order // 2
    x = [Fraction(2 * k - 1, 2) for k in range(-K + 1, K + 1)]
    w = dict(zip(x, _vandermonde_weights(x, Fraction(0), 1)))
    return [w[Fraction(2 * l - 1, 2)] for l in range(1, K + 1)]


@pytest.mark.parametrize("order", ORDERS)
def test_fd_weights_are_the_exact_staggered_solution(oracle_mod, order):
    from viscoelasticinversion_amd import stepper
    ex = _exact(order)
    for h in (1.0, 2.5, 5.0, 7.3):
        want = np.zeros(4)
        want[:len(ex)] = [float(c) / h for c in ex]
        assert np.array_equal(oracle_mod.fd_weights(h, order), want)
        assert np.array_equal(stepper.fd_weights(h, order), want)
    assert np.array_equal(oracle_mod.fd_weights(3.0, 8), oracle_mod.fd_coeffs(3.0))


def test_fd_weights_known_values_and_refusals(oracle_mod):
    from viscoelasticinversion_amd import _lib, stepper
    assert np.array_equal(oracle_mod.fd_weights(1.0, 4)[:2], [9 / 8, -1 / 24])
    assert np.array_equal(oracle_mod.fd_weights(1.0, 2), [1.0, 0, 0, 0])
    for bad in (0, 3, 10):
        with pytest.raises(ValueError):
            oracle_mod.fd_weights(1.0, bad)
        with pytest.raises(_lib.ViError):
            stepper.fd_weights(1.0, bad)
    m = _uniform_model(40, 40)
    with pytest.raises(ValueError):
        oracle_mod.Oracle(40, 40, 5.0, 5.0, 1e-4, m, fd_order=(8, 5))
    for fs in ("afda", "hermitian"):
        with pytest.raises(ValueError):
            oracle_mod.Oracle(40, 40, 5.0, 5.0, 1e-4, m, free_surface=fs, fd_order=(8, 4))
    oracle_mod.Oracle(40, 40, 5.0, 5.0, 1e-4, m, free_surface="afda", fd_order=(4, 8)).close()


@pytest.mark.parametrize("ox,oz", [(2, 2), (4, 6), (6, 4), (8, 2), (4, 8)])
def test_velocity_rates_exact_to_the_order(oracle_mod, ox, oz):
    """dvx = (d sxx/dx + d sxz/dz) / rho from sxx(x) and sxz(z) polynomials: exact for degree
    = order (the symmetric stencils also cancel the even term), off for degree order + 1."""
    nx, nz, h = 48, 44, 0.2
    rho = 2000.0
    m = _uniform_model(nx, nz, rho=rho)
    xi = (np.arange(nx) - nx / 2) * h       # sxx nodes (integer columns)
    zh = (np.arange(nz) - nz / 2 + 0.5) * h  # sxz nodes (half rows)
    xv = xi + 0.5 * h                        # vx columns
    zv = zh - 0.5 * h                        # vx rows
    sl = (slice(4, nz - 4), slice(4, nx - 4))
    for axis, order in (("x", ox), ("z", oz)):
        for deg, exact in ((order, True), (order + 1, False)):
            c = [0.3 * (-1) ** k / (k + 1) for k in range(deg + 1)]
            p = np.polynomial.Polynomial(c)
            o = oracle_mod.Oracle(nx, nz, h, h, 1e-4, m, dtype=np.float64, fd_order=(ox, oz))
            if axis == "x":
                o.set(oracle_mod.SXX, np.ones((nz, 1)) * p(xi)[None, :])
                want = np.ones((nz, 1)) * p.deriv()(xv)[None, :] / rho
            else:
                o.set(oracle_mod.SXZ, p(zh)[:, None] * np.ones((1, nx)))
                want = p.deriv()(zv)[:, None] * np.ones((1, nx)) / rho
            dv, _, _ = o.rates(None)
            o.close()
            err = np.abs(dv[0][sl] - want[sl]).max() / np.abs(want[sl]).max()
            if exact:
                assert err < 1e-13, (axis, order, deg, err)
            else:
                assert err > 1e-11, (axis, order, deg, err)


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_stable_dt_of_each_order(oracle_mod, order):
    """Homogeneous elastic medium, random initial velocity: bounded over 1500 steps at 0.995
    of the order's von Neumann bound, growing past 1e3 (or overflowing) at 1.005 of it: the
    bound is sharp for each order."""
    from viscoelasticinversion_amd.api import stable_dt
    nx = nz = 64
    dx, dz = 5.0, 4.0
    m = _uniform_model(nx, nz, vp=2500.0, vs=1400.0)
    lim = stable_dt(m, dx, dz, (order, order))
    lim8 = stable_dt(m, dx, dz)
    assert lim > lim8 or order == 8  # lower orders have smaller symbols
    v0 = np.zeros((nz, nx))
    v0[4:-4, 4:-4] = np.random.default_rng(2).standard_normal((nz - 8, nx - 8))
    grow = {}
    for f in (0.995, 1.005):
        o = oracle_mod.Oracle(nx, nz, dx, dz, f * lim, m, dtype=np.float64,
                              fd_order=(order, order))
        o.set(oracle_mod.VX, v0)
        o.set(oracle_mod.VZ, -v0)
        o.run(1500)
        grow[f] = float(np.abs(o.get(oracle_mod.VX)).max())
        o.close()
    assert grow[0.995] < 10 and not grow[1.005] < 1e3, grow
