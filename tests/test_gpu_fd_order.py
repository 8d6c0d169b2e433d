"""Staggered operators of order 2, 4, 6 and 8 per axis (Stencil(xorder, zorder, dx, dz;
stagger), src/common/Stencil.jl:280-320) on the GPU: the kernels take the order's weights
(zero beyond order/2), so every order runs the same code. Checked bit for bit against the
CPU oracle (FP32 and FP64, imaging, CPML, L = 3, VTI, every tiling), in a batch that mixes
orders, in the right-hand-side entry, and for the adjoint by the dot-product test."""
import numpy as np
import pytest

from helpers import make_oracle, run_gpu, run_oracle, small_cfg

pytestmark = pytest.mark.gpu


def _prec(dtype):
    return "float64" if dtype == np.float64 else "float32"


def _same(gpu, ref, label):
    for k in ref:
        assert np.array_equal(gpu[k], ref[k]), (label, k)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", [(2, 4), (4, 8), (6, 6), (8, 2), (4, 4)])
def test_orders_equal_oracle(order, dtype):
    cfg = small_cfg(nx=110, nz=72, nrelax=3, nt=60, fs=True)
    cfg["fd_order"] = order
    print(f"[case] fd order {order} {_prec(dtype)}", flush=True)
    _same(run_gpu(cfg, _prec(dtype)), run_oracle(cfg, dtype), f"order {order}")


@pytest.mark.parametrize("rows", [1, 2, 4, 8])
def test_order_4_every_tiling_and_vti(rows):
    for vti in (False, True):
        cfg = small_cfg(nx=130, nz=80, nrelax=2, nt=40, fs=True, vti=vti)
        cfg["fd_order"] = (4, 6)
        print(f"[case] fd order (4, 6) rows={rows} vti={vti}", flush=True)
        _same(run_gpu(cfg, "float32", tiling=rows), run_oracle(cfg, np.float32),
              f"rows {rows} vti {vti}")


def test_mixed_orders_in_one_batch():
    from viscoelasticinversion_amd import Batch
    from viscoelasticinversion_amd.configs import build_simulation
    orders = [(4, 4), (8, 8), (2, 6), (6, 2)]
    cfgs = []
    for k, o in enumerate(orders):
        c = small_cfg(nx=70 + 19 * k, nz=50 + 6 * k, nrelax=2, nt=40, seed=k + 5, fs=True)
        c["fd_order"] = o
        cfgs.append(c)
    print("[case] batch mixing orders", flush=True)
    sims = [build_simulation(c, "float32") for c in cfgs]
    b = Batch(sims)
    b.step(25)
    b.step(15)
    b.sync()
    for k, (c, sim) in enumerate(zip(cfgs, sims)):
        got = {f: sim.get(f) for f in ("vx", "vz", "sxx", "szz", "sxz", "mxx", "mzz", "mxz")}
        got["traces"] = sim.traces(40)
        _same(got, {f: v for f, v in run_oracle(c, np.float32).items() if f in got},
              f"batch shot {k}")
    b.close()
    for s in sims:
        s.close()


def test_rates_order_4_equal_oracle():
    from viscoelasticinversion_amd.configs import build_simulation
    cfg = small_cfg(nx=83, nz=57, nrelax=3, nt=12, fs=True)
    cfg["fd_order"] = (4, 4)
    rng = np.random.default_rng(9)
    o = make_oracle(cfg, np.float64)
    sim = build_simulation(cfg, "float64")
    for k in range(5):
        v = rng.standard_normal((cfg["nz"], cfg["nx"]))
        o.set(k, v)
        sim.set(k, v)
    sv = np.arange(1.0, 1.0 + len(cfg["src_ix"]))
    for a, b in zip(sim.rates(sv), o.rates(sv)):
        assert np.array_equal(a, b)
    sim.close()
    o.close()


def test_adjoint_dot_product_order_4_6():
    """<J dm, dd> = <dm, J^T dd> with operators of orders (4, 6), L = 3, imaging, CPML."""
    from test_gpu_gradient import _setup, _sim
    cfg, _ = _setup(3, True, True, seed=41)
    cfg["fd_order"] = (4, 6)
    nt = cfg["nt"]
    rng = np.random.default_rng(78)
    sim = _sim(cfg, cfg["model"])
    sim.step(nt)
    d0 = sim.traces(nt)
    sim.close()
    dd = rng.standard_normal(d0.shape) * np.abs(d0).max()
    print("[case] adjoint dot product orders (4, 6)", flush=True)
    sim = _sim(cfg, cfg["model"])
    _, g = sim.gradient(d0 - dd, nt, checkpoint_every=0)
    sim.close()
    nz, nx = cfg["nz"], cfg["nx"]
    for key in ("rho", "lam", "mu", "tau"):
        dm = np.zeros((nz, nx))
        dm[4:-4, 4:-4] = rng.standard_normal((nz - 8, nx - 8))
        base = cfg["model"][key]
        eps = 1e-6
        tr = []
        for sgn in (1, -1):
            m = dict(cfg["model"]); m[key] = base * (1 + sgn * eps * dm)
            s_ = _sim(cfg, m)
            s_.step(nt)
            tr.append(s_.traces(nt))
            s_.close()
        lhs = float(np.sum((tr[0] - tr[1]) / (2 * eps) * dd))
        rhs = float(np.sum(g[key] * base * dm))
        assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs)), (key, lhs, rhs)
