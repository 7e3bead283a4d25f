"""Mixed precision (sc_set_precision(1)): the x/z transforms, the nonlinear
products, the arrays between the transforms and the y-transform to the
Gauss points (on the FP32 matrix cores; the FP64 chunked rebuild where
dimR > 96, as at 128x128x129) in FP32; Galerkin projections, solves, flux
correction, the trapezoid projection and states in FP64.  Errors are measured
against the FP64 library path (itself at the oracle's rounding floor, see
tests/test_gpu_parity.py) from the same seeded state.

* 128x128x129 (Lx 2 pi, Lz pi, Re 5000, dt 0.002), 100 steps: state at steps
  1 and 100, the time-averaged mean-velocity profile and wall shear, a
  100-step tangent and a forced adjoint.  Measured 5.7e-7, 6.2e-7, 2e-12,
  9e-11, 2.2e-6, 7e-5; the bounds are 3x those.
* 64x64x97 (Lx 2 pi, Lz pi, Re 3000, dt 0.004): the same (measured 2.2e-7,
  3.4e-7, 4.9e-11, 4.8e-10, 1.5e-6, 4.8e-5 with the FP32 rebuild; bounds 3x
  those), and the time-averaged lower-wall shear over a 1000-step window
  with its gradient (channel.gradient.WallShearWindow, checkpoints every
  50): J, dJ/dmeanU, dJ/dRe and dJ/dC0 against FP64 (measured 1.8e-9,
  5.3e-10, 6.1e-6, 2.7e-5; bound 1e-4).
* Mixed-precision members and the field cache: bit for bit equal to the
  single-member and the uncached mixed runs (16x33x12 and 64x17x64)."""
import numpy as np
import pytest

from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.sc_set_precision(0)
    c_channel.sc_set_members(1)
    c_channel.sc_set_field_cache(1)
    c_channel.c_destroy()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _runs(cc, cfg, N, mixed):
    Nx, Ny, Nz, Lx, Lz, Re, dt = cfg
    C0 = synthetic_state(Nx, Ny, Nz, Lx, Lz, 4.0 / 3.0, seed=0)
    cc.sc_set_precision(mixed)
    assert cc.c_init(Nx, Ny, Nz, Lx, Lz, Re, 4.0 / 3.0, dt, N) == 0
    assert cc.sc_precision() == mixed
    C = C0.copy()
    cc.c_primal(0, C)
    prof, shear = cc.sc_mean_profile(0, 0, N + 1, 1.0 / (N + 1))
    out = dict(s1=cc.c_getsoln(1).copy(), sN=C, prof=prof.copy(), shear=shear.copy())
    I = synthetic_state(Nx, Ny, Nz, Lx, Lz, 4.0 / 3.0, seed=5, unit=True)
    cc.c_tangent(0, N, I, 0)
    out["tangent"] = I
    A = synthetic_state(Nx, Ny, Nz, Lx, Lz, 4.0 / 3.0, seed=6, unit=True)
    cc.c_adjoint(N, 0, A, 0, 1.0)
    out["adjoint"] = A
    cc.c_destroy()
    cc.sc_set_precision(0)
    return out


@pytest.mark.parametrize("grid,bounds", [
    ("128x128x129", dict(s1=2e-6, sN=2e-6, prof=1e-11, shear=3e-10, tangent=7e-6, adjoint=2.1e-4)),
    ("64x64x97", dict(s1=7e-7, sN=1.1e-6, prof=1.5e-10, shear=1.5e-9, tangent=4.5e-6,
                      adjoint=1.5e-4)),
])
def test_mixed_state_errors(cc, grid, bounds):
    cfg = {"128x128x129": (128, 129, 128, 2 * np.pi, np.pi, 5000.0, 0.002),
           "64x64x97": (64, 97, 64, 2 * np.pi, np.pi, 3000.0, 0.004)}[grid]
    N = 100
    print("case: mixed precision against FP64, %s, %d steps" % (grid, N), flush=True)
    ref, mix = _runs(cc, cfg, N, 0), _runs(cc, cfg, N, 1)
    errs = {k: rel(mix[k], ref[k]) for k in bounds}
    print(" ".join("%s %.2e" % kv for kv in errs.items()), flush=True)
    for k, b in bounds.items():
        assert errs[k] <= b, (k, errs[k])


def test_mixed_wall_shear_gradient(cc):
    from channel.gradient import WallShearWindow
    c = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, meanU=2.0 / 3.0, dt=0.004)
    N = 1000
    print("case: mixed-precision wall-shear gradient, 64x64x97, %d steps" % N, flush=True)
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], 2 * c["meanU"], seed=0)
    out = {}
    for mixed in (0, 1):
        cc.sc_set_precision(mixed)
        win = WallShearWindow(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["dt"], N, every=50)
        try:
            J, cps, dmu, dre = win.forward(C0, c["meanU"], c["Re"], params=True)
            out[mixed] = (J, dmu, dre, win.state_gradient(cps, c["meanU"], c["Re"]))
        finally:
            win.release()
            cc.sc_set_precision(0)
    (J0, mu0, re0, g0), (J1, mu1, re1, g1) = out[0], out[1]
    errs = {"J": abs(J1 - J0) / abs(J0), "dJ/dmeanU": abs(mu1 - mu0) / abs(mu0),
            "dJ/dRe": abs(re1 - re0) / abs(re0), "dJ/dC0": rel(g1, g0)}
    print(" ".join("%s %.2e" % kv for kv in errs.items()), flush=True)
    for k, e in errs.items():
        assert e <= 1e-4, (k, e)


@pytest.mark.parametrize("grid", ["16x33x12", "64x17x64"])
def test_mixed_members_and_cache_are_bitwise(cc, grid):
    g = {"16x33x12": (16, 33, 12, np.pi, 0.3 * np.pi, 2000.0, 0.005),
         "64x17x64": (64, 17, 64, 2 * np.pi, np.pi, 3000.0, 0.004)}[grid]
    Nx, Ny, Nz, Lx, Lz, Re, dt = g
    N, nm = 3, 3
    print("case: mixed-precision members and field cache, %s" % grid, flush=True)
    st = lambda s, **kw: synthetic_state(Nx, Ny, Nz, Lx, Lz, 4.0 / 3.0, seed=s, **kw)
    C0 = [st(s, amp=0.1) for s in range(nm)]
    D0 = [st(10 + s, unit=True) for s in range(nm)]
    A0 = [st(20 + s, unit=True) for s in range(nm)]
    cc.sc_set_precision(1)
    try:
        single = []
        for cache in (1, 0):
            cc.sc_set_field_cache(cache)
            assert cc.c_init(Nx, Ny, Nz, Lx, Lz, Re, 4.0 / 3.0, dt, N) == 0
            assert cc.sc_field_cache_active() == cache
            res = []
            for m in range(nm):
                C, I, A = C0[m].copy(), D0[m].copy(), A0[m].copy()
                cc.c_primal(0, C)
                cc.c_tangent(0, N, I, 0)
                cc.c_adjoint(N, 0, A, 0, 1.0)
                res.append((C, I, A))
            cc.c_destroy()
            single.append(res)
        cc.sc_set_field_cache(1)
        for a, b in zip(single[0], single[1]):
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)
        cc.sc_set_members(nm)
        assert cc.c_init(Nx, Ny, Nz, Lx, Lz, Re, 4.0 / 3.0, dt, N) == 0
        outs = []
        for fn, x, args in ((cc.sc_primal_members, C0, (0,)), (cc.sc_tangent_members, D0, (0, N, 0, False)),
                            (cc.sc_adjoint_members, A0, (N, 0, 1.0, False))):
            d = cc.DeviceArray(np.ascontiguousarray(np.stack(x)))
            try:
                if fn is cc.sc_primal_members:
                    fn(0, d)
                elif fn is cc.sc_tangent_members:
                    fn(0, N, d, 0, False)
                else:
                    fn(N, 0, d, 1.0, False)
                outs.append(d.download())
            finally:
                d.free()
        cc.c_destroy()
        for m in range(nm):
            for k in range(3):
                np.testing.assert_array_equal(outs[k][m], single[0][m][k])
    finally:
        cc.sc_set_members(1)
        cc.sc_set_field_cache(1)
        cc.sc_set_precision(0)
