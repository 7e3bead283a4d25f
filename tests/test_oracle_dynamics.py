"""Pin the oracle's time stepping with closed forms and identities.

* Stokes limit (no base flow): the one-step tangent map of a Fourier mode has
  the eigenvalues prod_k (1 + a_k h)/(1 - b_k h), h = sigma dt, with sigma the
  analytic Stokes eigenvalues (v: mu tan mu = -k tanh k, mu cot mu = k coth k;
  omega_y: sigma = -(k^2 + (n pi/2)^2)/Re).  Pins the Galerkin operators, the
  RK3/CN coefficients (channel/project.c:44-47) and the band solves.
* Orr-Sommerfeld: laminar Poiseuille at Re = 10000, alpha = 1 has
  c = 0.23752649 + 0.00373967 i (Orszag 1971).  Pins the linearised
  convective term of the tangent (channel/pass1.c, channel/increproject.c).
* Central finite differences of the primal against the tangent, for the
  homogeneous tangent and for the mean-velocity tangent
  (channel/increproject0.c:209-214: inhomo = d/d(meanU) with flux = 2 meanU).
* Duality of tangent and adjoint in the Gauss-weighted physical inner
  product used by drivers/testAdj.py:55-56.
* Gauss-point reconstruction (channel/initAlphaBeta2.c): continuity and
  wall-normal vorticity identities.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

import pyorc
from channel.synthetic import synthetic_state

A = [1 / 3, -1 / 2, 1 / 3]
B = [1 / 6, 2 / 3, 0.0]


def stage_product(h):
    r = 1.0
    for a, b in zip(A, B):
        r *= (1 + a * h) / (1 - b * h)
    return r


def tangent_map(o, z, x, nsteps=1):
    nQ, nR = o.dimQ, o.dimR
    cols = []
    for v, j in [(0, j) for j in range(nQ)] + [(1, j) for j in range(nR)]:
        IC = np.zeros(o.cshape, complex)
        IC[z, v, j, x] = 1.0
        assert o.tangent(0, nsteps, IC, 0) == 0
        cols.append(np.concatenate([IC[z, 0, :nQ, x], IC[z, 1, :, x]]))
    return np.array(cols).T


def test_stokes_eigenvalues():
    Re, dt, k = 20.0, 0.01, 1.5
    o = pyorc.Oracle(4, 33, 2, 2 * np.pi / k, 1.0, Re, 0.0, dt, 1)
    o.primal(0, np.zeros(o.cshape, complex))
    lam = np.linalg.eigvals(tangent_map(o, 0, 1))
    mus_e = [brentq(lambda m: m * np.tan(m) + k * np.tanh(k), (n - .5) * np.pi + 1e-9,
                    (n + .5) * np.pi - 1e-9) for n in range(1, 6)]
    mus_o = [brentq(lambda m: m / np.tan(m) - k / np.tanh(k), n * np.pi + 1e-9,
                    (n + 1) * np.pi - 1e-9) for n in range(1, 6)]
    sig = [-(m * m + k * k) / Re for m in mus_e + mus_o]
    sig += [-(k * k + (n * np.pi / 2) ** 2) / Re for n in range(1, 9)]
    for s in sig:
        assert np.abs(lam - stage_product(s * dt)).min() < 1e-10


def test_orr_sommerfeld_growth_rate():
    Re, dt = 10000.0, 0.01
    o = pyorc.Oracle(4, 65, 2, 2 * np.pi, 1.0, Re, 4.0 / 3.0, dt, 1)
    C = np.zeros(o.cshape, complex)
    C[0, 0, 0, 0] = 1.0            # u = 1 - y^2: R_0 coefficient
    o.primal(0, C)
    lam = np.linalg.eigvals(tangent_map(o, 0, 1))
    s = np.log(lam[np.argmax(np.abs(lam))]) / dt
    assert abs(s.real - 0.00373967) < 2e-8
    assert abs(-s.imag - 0.23752649) < 5e-8


CFG = dict(Nx=8, Ny=17, Nz=8, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, dt=0.005)
FLUX = 4.0 / 3.0


def make(n, flux=FLUX):
    c = CFG
    return pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], flux, c["dt"], n)


def state(seed, **kw):
    c = CFG
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], FLUX, seed=seed, **kw)


def test_tangent_matches_central_differences_and_is_linear():
    n = 4
    o = make(n)
    C0 = state(0, amp=0.2)
    d = state(1, unit=True)

    def run(C):
        C = C.copy()
        o.primal(0, C)
        return C
    eps = 1e-4
    fd = (run(C0 + eps * d) - run(C0 - eps * d)) / (2 * eps)
    o.primal(0, C0.copy())
    IC = d.copy()
    o.tangent(0, n, IC, 0)
    assert np.linalg.norm(fd - IC) / np.linalg.norm(IC) < 1e-8
    # MIC holds the tangent at step boundaries
    np.testing.assert_array_equal(o.getincresoln(n), IC)
    np.testing.assert_array_equal(o.getincresoln(0), d)
    IC2 = 2.5 * d
    o.tangent(0, n, IC2, 0)
    assert np.linalg.norm(IC2 - 2.5 * IC) <= 1e-11 * np.linalg.norm(IC)


def test_mean_velocity_tangent_matches_differences_in_flux():
    n = 4
    C0 = state(0, amp=0.2)
    e = 1e-6

    def run_flux(fl):
        o = make(n, fl)
        C = C0.copy()
        o.primal(0, C)
        return C
    fd = (run_flux(FLUX + 2 * e) - run_flux(FLUX - 2 * e)) / (2 * e)
    o = make(n)
    o.primal(0, C0.copy())
    IC = np.zeros_like(C0)
    o.tangent(0, n, IC, 1)
    assert np.linalg.norm(fd - IC) / np.linalg.norm(IC) < 1e-7


def test_adjoint_is_dual_to_tangent():
    n = 4
    o = make(n)
    o.primal(0, state(0, amp=0.2))
    y, w = o.quad()
    IC0 = state(2, unit=True)
    AC0 = state(3, unit=True)
    IC = IC0.copy()
    o.tangent(0, n, IC, 0)
    AC = AC0.copy()
    o.adjoint(n, 0, AC, 0, 0.0)

    def ip(a, b):
        return np.sum(o.spec2phys(a) * o.spec2phys(b) * w[None, None, :, None])
    left, right = ip(AC0, IC), ip(AC, IC0)
    assert abs(left - right) <= 1e-10 * abs(left)


def test_gauss_reconstruction_identities():
    o = make(1)
    C = state(4, amp=0.3)
    U = o.coef_to_gauss(C)
    q, nQ, nR = o.qpts, o.dimQ, o.dimR
    Q = o.array("Q").reshape(q, nQ)
    Qp = o.array("Qp").reshape(q, nQ)
    Qpp = o.array("Qpp").reshape(q, nQ)
    R = o.array("R").reshape(q, nR)
    Rp = o.array("Rp").reshape(q, nR)
    kx = o.array("Kx")
    kz = o.array("Kz")
    for z in range(o.Nz):
        for x in range(o.Nx // 2):
            if (z, x) == (0, 0):
                np.testing.assert_allclose(U[0, 0, :, 0], R @ C[0, 0, :, 0], atol=1e-14)
                np.testing.assert_allclose(U[0, 3, :, 0], Rp @ C[0, 0, :, 0], atol=1e-13)
                np.testing.assert_allclose(U[0, 2, :, 0], R @ C[0, 1, :, 0], atol=1e-14)
                np.testing.assert_allclose(U[0, 4, :, 0], Rp @ C[0, 1, :, 0], atol=1e-13)
                continue
            a, b = C[z, 0, :nQ, x], C[z, 1, :, x]
            u, v, w_, du, dw = U[z, :, :, x]
            np.testing.assert_allclose(v, Q @ a, atol=1e-14)
            # continuity: i kx u + dv/dy + i kz w = 0 (and its y derivative)
            np.testing.assert_allclose(1j * kx[x] * u + Qp @ a + 1j * kz[z] * w_, 0, atol=1e-12)
            np.testing.assert_allclose(1j * kx[x] * du + Qpp @ a + 1j * kz[z] * dw, 0, atol=1e-11)
            # wall-normal vorticity omega_y = i kz u - i kx w = R beta
            np.testing.assert_allclose(1j * kz[z] * u - 1j * kx[x] * w_, R @ b, atol=1e-12)
            np.testing.assert_allclose(1j * kz[z] * du - 1j * kx[x] * dw, Rp @ b, atol=1e-11)


def test_ddt_project_makes_orthogonal():
    n = 3
    o = make(n)
    o.primal(0, state(0, amp=0.2))
    y, w = o.quad()
    for i in [0, 2, n]:
        v = state(5 + i, unit=True)
        v0 = v.copy()
        r = o.ddt_project(i, v)
        ip_, im_ = (3, 0) if i == 0 else (3 * i, 3 * (i - 1))
        dC = (o.getsoln(1) - o.getsoln(0)) / CFG["dt"] if i == 0 else \
            (o.getsoln(i) - o.getsoln(i - 1)) / CFG["dt"]
        np.testing.assert_allclose(v, v0 - r * dC, atol=1e-15)
        ip = np.sum(o.spec2phys(dC) * o.spec2phys(v) * w[None, None, :, None])
        nn = np.sum(o.spec2phys(dC) ** 2 * w[None, None, :, None])
        assert abs(ip) < 1e-12 * nn


def test_cfl_violation_truncates_run():
    c = CFG
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], FLUX, 5.0, 3)
    C0 = state(0, amp=0.2)
    C = C0.copy()
    o.primal(0, C)
    assert o.last_status() == -1
    np.testing.assert_array_equal(C, C0)
    assert np.all(o.getsoln(1) == 0)


def test_tangent_forcing_against_dense_quadrature_operators():
    """Re-sensitivity forcing (channel/tangent_forcing.c:10-75,
    channel/tangent_forcing0.c:10-44) against dense operators assembled
    independently of the band storage: Gram matrices by Gauss quadrature of
    the basis values (exact at these degrees), f_a = -re^2 (K2^2 G_QQ +
    2 K2 G_Q'Q' + G_Q''Q'') alpha, f_b = re^2 (G_R'R' + K2 G_RR) beta, mean
    column re^2 G_R'R' alpha(0,0) and beta(0,0); column (0,0) of f_a, f_b zero."""
    Nx, Ny, Nz = 16, 33, 6
    o = pyorc.Oracle(Nx, Ny, Nz, 2.0, 1.3, 1000.0, 4.0 / 3.0, 0.01, 1)
    C = synthetic_state(Nx, Ny, Nz, 2.0, 1.3, 4.0 / 3.0, seed=3, amp=0.2)
    o.primal(0, C.copy())
    q, nQ, nR, Mx = o.qpts, o.dimQ, o.dimR, Nx // 2
    y, w = o.quad()
    Q, Qp, Qpp = (o.array(n).reshape(q, nQ) for n in ("Q", "Qp", "Qpp"))
    R, Rp = (o.array(n).reshape(q, nR) for n in ("R", "Rp"))
    gram = lambda A, B: (A * w[:, None]).T @ B
    GQ, GQp, GQpp, GR, GRp = gram(Q, Q), gram(Qp, Qp), gram(Qpp, Qpp), gram(R, R), gram(Rp, Rp)
    Kx, Kz = o.array("Kx"), o.array("Kz")
    re = o.scalar("re")
    for count in (0, 3):
        fa, fb, f0a, f0b = o.tangent_forcing(count)
        S = o.getsoln(count // 3)            # stored state MC[count]
        al, be = S[:, 0], S[:, 1]
        for z in range(Nz):
            for x in range(Mx):
                if z == 0 and x == 0:
                    assert np.all(fa[z, :, x] == 0) and np.all(fb[z, :, x] == 0)
                    continue
                k2 = Kz[z] ** 2 + Kx[x] ** 2
                ea = -re * re * ((k2 * k2) * GQ + 2 * k2 * GQp + GQpp) @ al[z, :nQ, x]
                eb = re * re * (GRp + k2 * GR) @ be[z, :, x]
                assert np.linalg.norm(fa[z, :, x] - ea) <= 1e-12 * np.linalg.norm(ea) + 1e-300
                assert np.linalg.norm(fb[z, :, x] - eb) <= 1e-12 * np.linalg.norm(eb) + 1e-300
        for f0, v in ((f0a, al[0, :, 0]), (f0b, be[0, :, 0])):
            e0 = re * re * GRp @ v
            assert np.linalg.norm(f0 - e0) <= 1e-12 * np.linalg.norm(e0) + 1e-300
    o.destroy()
