"""The wall-shear gradient on the device.

* sc_tangent_sources (mean-velocity and Reynolds-number sources) and
  sc_adjoint_objective against the oracle's orc_tangent_src and
  orc_adjoint_objective (pinned in tests/test_oracle_gradient.py), 16x33x12,
  at the one-step bar of tests/test_gpu_parity.py (1e-10), or 3x the
  rounding floor where that is higher (the Reynolds-number tangent from
  zero).
* sc_gradient_euclid: sum Re(conj(g) d) = <lam, d> in the physical inner
  product formed from spec2phys fields, for directions with mean, kx = 0
  and Nyquist-row content.
* The 64x64x97 workload with the time-averaged lower-wall shear over a
  1000-step window checkpointed every 50 steps (channel.gradient.WallShear,
  a torch.autograd.Function): J equal to the unwindowed run's; dJ/dmeanU,
  dJ/dRe and dJ/dC0 (along a random direction) against central differences
  (fourth-order central quotients at 1e-3 of each quantity's size: the
  chaotic growth of rounding differences over 1000 steps makes a 1e-6 step
  too noisy, see below); the adjoint gradient against the forward-mode
  (tangent) directional derivative.  Bar: 1e-4.
"""
import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

SMALL = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)
BENCH = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, meanU=2.0 / 3.0, dt=0.004)


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.c_destroy()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _st(cfg, seed, flux=4.0 / 3.0, **kw):
    c = cfg
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], flux, seed=seed, **kw)


def test_sources_and_objective_adjoint_against_oracle(cc):
    print("case: tangent sources and objective adjoint vs oracle, 16x33x12", flush=True)
    c, n = SMALL, 3
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], n)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     n) == 0
    C0 = _st(c, 0, amp=0.1)
    o.primal(0, C0.copy())
    cc.c_primal(0, C0.copy())
    # bounds: 3x the rounding floor (tests/test_oracle_rounding_floor.py),
    # 1e-10 where that is lower; the Reynolds-number tangent from zero is a
    # small difference of O(1) stage terms (floor 8.1e-10)
    for inhomo, re_src, seed, bound in ((0, 1, None, 2.5e-9), (1, 1, 2, 1e-10),
                                        (1, 0, None, 1e-10)):
        d = np.zeros_like(C0) if seed is None else _st(c, seed, unit=True)
        Io, Ig = d.copy(), d.copy()
        o.tangent_src(0, n, Io, inhomo, re_src)
        cc.sc_tangent_sources(0, n, Ig, inhomo, re_src)
        err = rel(Ig, Io)
        print("inhomo %d re_source %d: rel err %.2e" % (inhomo, re_src, err), flush=True)
        assert err <= bound, (inhomo, re_src, err)
    A0 = _st(c, 4, unit=True)
    A0[0, 0, :, 0] += np.linspace(0.1, -0.2, A0.shape[2])   # a mean column as well
    Ao, Ag = A0.copy(), A0.copy()
    o.adjoint_objective(n, 0, Ao, 0.7)
    cc.sc_adjoint_objective(n, 0, Ag, 0.7)
    assert rel(Ag, Ao) <= 1e-10
    o.destroy()


def test_euclidean_gradient_map(cc):
    print("case: Euclidean gradient map, 16x33x12", flush=True)
    c = SMALL
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     1) == 0
    y, w = cc.c_quad()
    rng = np.random.default_rng(7)
    shape = (c["Nz"], 2, c["Ny"] - 2, c["Nx"] // 2)

    def rand():
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    def phys(C):
        f = np.zeros((3, 3 * c["Nx"] // 2, len(y), 3 * c["Nz"] // 2))
        cc.c_spec2phys(np.ascontiguousarray(C), f)
        return f
    for _ in range(2):
        lam, d = rand(), rand()
        g = cc.sc_gradient_euclid(lam)
        ip = float(np.sum(phys(lam) * phys(d) * w[None, None, :, None]))
        assert abs(np.sum(g.real * d.real + g.imag * d.imag) - ip) <= 1e-12 * abs(ip) + 1e-12


def test_checkpointed_wall_shear_gradient(cc):
    """1000 steps at 64x64x97, checkpoints every 50."""
    import torch
    from channel.gradient import WallShearWindow, wall_shear
    print("case: wall-shear gradient, 64x64x97, 1000 steps, checkpoints every 50", flush=True)
    c = BENCH
    N = 1000
    win = WallShearWindow(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["dt"], N, every=50)
    try:
        C0 = _st(c, 0, flux=2 * c["meanU"])
        Ct = torch.tensor(C0, requires_grad=True)
        mu = torch.tensor(c["meanU"], dtype=torch.float64, requires_grad=True)
        Re = torch.tensor(c["Re"], dtype=torch.float64, requires_grad=True)
        J = wall_shear(win, Ct, mu, Re)
        J.backward()
        gC, gmu, gRe = Ct.grad.numpy(), float(mu.grad), float(Re.grad)

        def Jof(C, m, r):
            return win.forward(C, m, r, params=False)[0]
        # checkpointing changes nothing: one window of 1000 steps
        whole = WallShearWindow(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["dt"], N, every=N)
        J_whole = whole.forward(C0, c["meanU"], c["Re"], params=False)[0]
        whole.release()
        assert abs(float(J) - J_whole) <= 1e-13 * abs(J_whole)
        # Finite differences.  Over 1000 steps the flow amplifies rounding
        # differences between two runs to about 1e-11 of J, which divided by
        # a step of 1e-6 is a 1e-2 error in the difference quotient (measured:
        # the central quotient at 1e-6 is 0.2-2.4% off and drifts as the step
        # shrinks).  The check uses a fourth-order central quotient at a step
        # of 1e-3 of each quantity's size (truncation O(h^4), noise ~1e-8);
        # the 1e-6 central quotient is printed beside it.
        def fd4(f, x, h):
            return (8 * (f(x + h) - f(x - h)) - (f(x + 2 * h) - f(x - 2 * h))) / (12 * h)

        def fd2(f, x, h):
            return (f(x + h) - f(x - h)) / (2 * h)
        d = _st(c, 9, unit=True)
        f_mu = lambda m: Jof(C0, m, c["Re"])
        f_re = lambda r: Jof(C0, c["meanU"], r)
        f_C = lambda t: Jof(C0 + t * d, c["meanU"], c["Re"])
        scale_C = np.linalg.norm(C0)
        fd_mu = fd4(f_mu, c["meanU"], 1e-3 * c["meanU"])
        fd_re = fd4(f_re, c["Re"], 1e-3 * c["Re"])
        fd_C = fd4(f_C, 0.0, 1e-3 * scale_C)
        print("central quotients at 1e-6: dJ/dmeanU %.6e dJ/dRe %.6e dJ/dC0.d %.6e" % (
            fd2(f_mu, c["meanU"], 1e-6 * c["meanU"]), fd2(f_re, c["Re"], 1e-6 * c["Re"]),
            fd2(f_C, 0.0, 1e-6 * scale_C)), flush=True)
        gd = float(np.sum(gC.real * d.real + gC.imag * d.imag))
        # forward mode along d: the tangent from d, then the objective's tangent
        win._init(c["meanU"], c["Re"])
        I = d.copy()
        tan = 0.0
        for Cw in win.forward(C0, c["meanU"], c["Re"], params=False)[1]:
            cc.c_primal(0, Cw.copy())
            cc.sc_tangent_sources(0, win.every, I, 0, 0)
            tan += float(cc.sc_mean_profile(1, 0, win.every, 1.0 / N)[1][0])
        errs = {"meanU": abs(gmu - fd_mu) / abs(fd_mu), "Re": abs(gRe - fd_re) / abs(fd_re),
                "C0": abs(gd - fd_C) / abs(fd_C), "C0_vs_tangent": abs(gd - tan) / abs(tan)}
        print("J %.10e dJ/dmeanU %.6e (fd %.6e) dJ/dRe %.6e (fd %.6e) dJ/dC0.d %.6e (fd %.6e, "
              "tangent %.6e)" % (float(J), gmu, fd_mu, gRe, fd_re, gd, fd_C, tan), flush=True)
        print(" ".join("%s %.2e" % kv for kv in errs.items()), flush=True)
        # adjoint against tangent: the exact transpose, equal up to rounding
        # grown over 1000 steps (measured 2.9e-9)
        for k, e in errs.items():
            assert e <= (1e-8 if k == "C0_vs_tangent" else 1e-4), (k, e)
    finally:
        win.release()
