"""Long windows on the GPU: configurations the one-step parity tests do not reach.

* 64x64x97 (Lx = 2 pi, Lz = pi, Re = 3000, dt = 0.004): a 10-step window
  of primal, homogeneous tangent and forced adjoint (strength 1) against the
  oracle's sketches in tests/golden/window_64x64x97_10steps.npz (written by
  tests/golden/make_window_fixtures.py; the oracle needs minutes for it).
  Bounds: 3x the FMA/plain floor of the oracle on the same window (the
  fixture script's --floor mode: 9.7e-11 primal, 1.7e-9 tangent, 1.9e-9
  adjoint), on the sketch scale (x 1.5, tests/golden/sketch.py).
* The 500-step tangent of the minimal channel (32x32x65, Lx = pi,
  Lz = 0.3 pi, Re = 2000, dt = 0.005): tangent-adjoint duality in the
  Gauss-weighted physical inner product of drivers/testAdj.py:55-56, for the
  homogeneous tangent, and with the mean-velocity source (inhomo 1):
  <AC_N, IC_N> = <AC_0, IC_0> + <AC_N, S_N>, S_N the tangent from zero
  with the source.  1e-10 relative.
"""
import os
import sys

import numpy as np
import pytest

from channel.synthetic import synthetic_state

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from sketch import rel, sketch  # noqa: E402

pytestmark = pytest.mark.gpu

BENCH = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0, dt=0.004)
G0 = dict(Nx=32, Ny=65, Nz=32, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)
FIXTURE = os.path.join(HERE, "golden", "window_64x64x97_10steps.npz")

# FMA/plain floors of the oracle on the fixture's window (relative, full
# arrays), measured with tests/golden/make_window_fixtures.py --floor
FLOOR = {"primal_1": 9.7e-11, "primal_mid": 1.06e-10, "primal_end": 9.3e-11,
         "tangent_end": 1.68e-9, "adjoint_end": 1.90e-9}


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.c_destroy()


def _state(cfg, seed, **kw):
    c = cfg
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=seed, **kw)


def _init(cc, cfg, n):
    c = cfg
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     n) == 0


def test_ten_step_window_at_benchmark_grid(cc):
    fx = np.load(FIXTURE)
    n = int(fx["nstep"])
    print("case: 64x64x97 window of %d steps, primal/tangent/adjoint vs fixture" % n, flush=True)
    _init(cc, BENCH, n)
    C = _state(BENCH, 0)
    cc.c_primal(0, C)
    got = {"primal_1": cc.c_getsoln(1), "primal_mid": cc.c_getsoln(n // 2),
           "primal_end": cc.c_getsoln(n)}
    IC = _state(BENCH, 3, unit=True)
    cc.c_tangent(0, n, IC, 0)
    got["tangent_end"] = IC
    AC = _state(BENCH, 4, unit=True)
    cc.c_adjoint(n, 0, AC, 0, 1.0)
    got["adjoint_end"] = AC
    errs = {}
    for k, a in got.items():
        assert abs(np.linalg.norm(a) / fx[k + "_norm"] - 1) <= 1.5 * 3 * FLOOR[k] + 1e-12, k
        errs[k] = rel(sketch(a), fx[k])
    print(" ".join("%s %.2e" % kv for kv in errs.items()), flush=True)
    for k, e in errs.items():
        assert e <= 1.5 * 3 * FLOOR[k], (k, e)


def _phys_ip(cc, cfg, w):
    shape = (3, 3 * cfg["Nx"] // 2, len(w), 3 * cfg["Nz"] // 2)

    def phys(C):
        f = np.zeros(shape)
        cc.c_spec2phys(np.ascontiguousarray(C), f)
        return f
    return lambda a, b: float(np.sum(phys(a) * phys(b) * w[None, None, :, None]))


def test_five_hundred_step_tangent_duality(cc):
    n = 500
    print("case: 32x32x65, 500-step tangent and adjoint, duality", flush=True)
    _init(cc, G0, n)
    cc.c_primal(0, _state(G0, 0))
    y, w = cc.c_quad()
    ip = _phys_ip(cc, G0, w)
    IC0, AC0 = _state(G0, 1, unit=True), _state(G0, 2, unit=True)
    AC = AC0.copy()
    cc.c_adjoint(n, 0, AC, 0, 0.0)
    # homogeneous tangent
    IC = IC0.copy()
    cc.c_tangent(0, n, IC, 0)
    left, right = ip(AC0, IC), ip(AC, IC0)
    gap = abs(left - right) / abs(left)
    # with the mean-velocity source: affine in IC0
    ICs = IC0.copy()
    cc.c_tangent(0, n, ICs, 1)
    S = np.zeros_like(IC0)
    cc.c_tangent(0, n, S, 1)
    left_s = ip(AC0, ICs)
    right_s = ip(AC, IC0) + ip(AC0, S)
    gap_s = abs(left_s - right_s) / abs(left_s)
    print("duality gap: homogeneous %.2e, with source %.2e (<AC0, IC_N> = %.6e)"
          % (gap, gap_s, left), flush=True)
    assert gap <= 1e-10
    assert gap_s <= 1e-10
