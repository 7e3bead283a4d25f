"""Pin the oracle's incremental adjoint stage by independent routes.

Reference: channel/increadjproject.c:10-645 (modes (kx, kz) != (0, 0)) and
channel/increadjproject0.c:14-357 (the mean mode), called for every mode in
the order of shan/increadjoint_test7.c:673-685.

- The wall-lift terms and the tangent coupling (mean mode included) are
  checked against dense
  Galerkin matrices formed here by Gauss quadrature from the basis values at
  the Gauss points (not from the banded Gram arrays), dense solves (not the
  band LU), and the lift polynomials' own derivatives (numpy.polyder).  The
  reference's lift tables (channel/increadjproject.c:40-46) are thereby
  pinned as the Galerkin projections (Q_i, lv), (Q_i, lv''), (R_i, lg) of
  lv = (1-y)^2 (1+y)/4 and lg = (1-y)/2: a wrong entry or sign fails.
- The lift polynomials are checked at the walls: lv(+-1) = 0, lv'(+1) = 0,
  lv'(-1) = 1, lg(-1) = 1, lg(+1) = 0.
- The wall values IAUxb, IAUzb are checked against du/dy, dw/dy of the
  rebuilt field at y = -1 (Q_j''(-1) and R_j'(-1) from Legendre fits to the
  Gauss-point values); the reference doubles the alpha sum's share of their
  imaginary parts (channel/increadjproject.c:469-476), which the check
  carries as a stated factor.
- With zero wall planes and zero tangent states the stage reduces to the
  adjoint stage (channel/adjproject.c), already pinned by the adjoint tests:
  every mode except the mean (whose flux correction the incremental stage
  drops) agrees.
"""
import numpy as np
import pytest

import pyorc

A = [1. / 3., -1. / 2, 1. / 3.]
B = [2. / 3, 1. / 6, 0.]
C = [1., 1. / 3., 1. / 2.]
D = [0., -2. / 3, -1. / 6]
RE_NUM, DT, NSTEPS = 800.0, 0.01, 2


@pytest.fixture(scope="module", params=[(8, 17, 8), (12, 21, 10)])
def case(request):
    Nx, Ny, Nz = request.param
    o = pyorc.Oracle(Nx, Ny, Nz, 2.0, 1.3, RE_NUM, 4.0 / 3.0, DT, NSTEPS)
    rng = np.random.default_rng(7)
    o.primal(0)
    IC = 1e-2 * (rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape))
    IC[Nz // 2] = 0
    IC[0, :, :, 0] = IC[0, :, :, 0].real
    o.tangent(0, NSTEPS, IC, 0)
    return o, rng


def basis(o):
    y, w = o.quad()
    q = y.size
    nR, nQ = o.dimR, o.dimQ
    g = lambda n, m: o.array(n).reshape(q, m)
    Q, Qp, Qpp, R, Rp = g("Q", nQ), g("Qp", nQ), g("Qpp", nQ), g("R", nR), g("Rp", nR)
    ip = lambda X, Y: X.T @ (w[:, None] * Y)
    return dict(y=y, w=w, Q=Q, Qp=Qp, Qpp=Qpp, R=R, Rp=Rp, Mq=ip(Q, Q), Mq1=ip(Qp, Qp),
                Mq2=ip(Qpp, Qpp), Mr=ip(R, R), Mr1=ip(Rp, Rp))


LV = np.polymul(np.polymul([-1.0, 1.0], [-1.0, 1.0]), [1.0, 1.0]) / 4.0   # (1-y)^2 (1+y)/4
LG = np.array([-0.5, 0.5])                                                  # (1-y)/2


def test_lift_polynomials_at_the_walls():
    d1 = np.polyder(LV)
    assert np.polyval(LV, 1.0) == 0 and np.polyval(LV, -1.0) == 0
    assert np.polyval(d1, 1.0) == 0 and np.polyval(d1, -1.0) == 1.0
    assert np.polyval(LG, -1.0) == 1.0 and np.polyval(LG, 1.0) == 0


def k2_of(o):
    Kx, Kz = o.array("Kx"), o.array("Kz")
    return Kx, Kz, Kz[:, None] ** 2 + Kx[None, :] ** 2


def rand_c(rng, shape, s=1.0):
    return s * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def expected_modes(o, bs, k, count, IAC, W, Fk=None, Fk1=None):
    """Dense Galerkin restatement of one stage for the modes (z, x) != (0, 0),
    z != Nz/2, with zero products: returns IAC after the solve and after the
    next stage's explicit part."""
    re, dt = o.scalar("re"), o.scalar("dt")
    nQ, nR = o.dimQ, o.dimR
    Kx, Kz, K2 = k2_of(o)
    y, w = bs["y"], bs["w"]
    lv, lv2 = np.polyval(LV, y), np.polyval(np.polyder(LV, 2), y)
    lg, lg1 = np.polyval(LG, y), np.polyval(np.polyder(LG), y)
    Qlv, Qlv2 = bs["Q"].T @ (w * lv), bs["Q"].T @ (w * lv2)
    Rlg, Rplg1 = bs["R"].T @ (w * lg), bs["Rp"].T @ (w * lg1)
    M0 = o.getincrestage(count)
    M1 = o.getincrestage(count - 1) if k < 2 else None
    sol, nxt = IAC.copy(), IAC.copy()
    Nz, Mx = IAC.shape[0], IAC.shape[3]
    for z in range(Nz):
        if z == Nz // 2:
            continue
        for x in range(Mx):
            if z == 0 and x == 0:
                # mean mode: Rs, Rps operators, lifts (R_i, lg), no viscous lift term
                ub, wb, ubt, wbt = W[:4, 0, 0]
                Mr, Mr1 = bs["Mr"], bs["Mr1"]
                for v, (wt, w0) in enumerate(((ubt, ub), (wbt, wb))):
                    c0 = IAC[0, v, :, 0].copy()
                    if k == 0:
                        c0 = (Mr - re * A[0] * dt * Mr1) @ c0 - re * dt * (A[0] + B[0]) * Mr1 @ M0[0, v, :, 0]
                    c0 = c0 + Rlg * (wt - w0)
                    if Fk is not None:
                        c0 = c0 + dt * C[k] * Fk[v, 0, 0, :].real
                    c0 = np.linalg.solve(Mr + re * B[k] * dt * Mr1, c0)
                    sol[0, v, :, 0] = c0
                    if k < 2:
                        cn = (Mr - re * A[k + 1] * dt * Mr1) @ c0 - \
                            re * dt * (A[k + 1] + B[k + 1]) * Mr1 @ M1[0, v, :, 0]
                        if Fk1 is not None:
                            cn = cn + dt * D[k + 1] * Fk1[v, 0, 0, :]
                        nxt[0, v, :, 0] = cn
                continue
            k2 = K2[z, x]
            Dq = k2 * k2 * bs["Mq"] + 2 * k2 * bs["Mq1"] + bs["Mq2"]
            Hq = -(k2 * bs["Mq"] + bs["Mq1"])
            Dr = -(bs["Mr1"] + k2 * bs["Mr"])
            ub, wb, ubt, wbt, aub, awb = W[:, z, x]
            al, be = IAC[z, 0, :nQ, x].copy(), IAC[z, 1, :, x].copy()
            if k == 0:
                al = (Hq + re * A[0] * dt * Dq) @ al + re * dt * (A[0] + B[0]) * Dq @ M0[z, 0, :nQ, x]
                be = (bs["Mr"] + re * A[0] * dt * Dr) @ be + re * dt * (A[0] + B[0]) * Dr @ M0[z, 1, :, x]
            vis = A[k] * wbt + B[k] * wb + (A[k] + B[k]) * awb
            al = al + (Qlv2 - k2 * Qlv) * (wbt - wb) + dt * re * vis * (k2 * k2 * Qlv - 2 * k2 * Qlv2)
            visb = A[k] * ubt + B[k] * ub + (A[k] + B[k]) * aub
            be = be + Rlg * (ubt - ub) - dt * re * visb * (k2 * Rlg + Rplg1)
            if Fk is not None:
                al = al + dt * C[k] * Fk[0, z, x, :nQ]
                be = be + dt * C[k] * Fk[1, z, x, :]
            al = np.linalg.solve(Hq - re * B[k] * dt * Dq, al)
            be = np.linalg.solve(bs["Mr"] - re * B[k] * dt * Dr, be)
            sol[z, 0, :nQ, x], sol[z, 1, :, x] = al, be
            if k < 2:
                an = (Hq + re * A[k + 1] * dt * Dq) @ al + re * dt * (A[k + 1] + B[k + 1]) * Dq @ M1[z, 0, :nQ, x]
                bn = (bs["Mr"] + re * A[k + 1] * dt * Dr) @ be + \
                    re * dt * (A[k + 1] + B[k + 1]) * Dr @ M1[z, 1, :, x]
                if Fk1 is not None:
                    an = an + dt * D[k + 1] * Fk1[0, z, x, :nQ]
                    bn = bn + dt * D[k + 1] * Fk1[1, z, x, :]
                nxt[z, 0, :nQ, x], nxt[z, 1, :, x] = an, bn
    return sol, (nxt if k < 2 else sol)


def inputs(o, rng, walls=True):
    Nz, Mx = o.cshape[0], o.cshape[3]
    IAC = rand_c(rng, o.cshape)
    W = rand_c(rng, (6, Nz, Mx)) if walls else np.zeros((6, Nz, Mx), complex)
    return IAC, W


@pytest.mark.parametrize("k", [0, 1, 2])
def test_walls_and_tangent_coupling_against_dense_galerkin(case, k):
    o, rng = case
    bs = basis(o)
    count = 4 - k
    IAC, W = inputs(o, rng)
    Z = np.zeros(o.ushape, complex)
    out, IAU, IAUb = o.increadjproject(k, count, IAC, Z, Z, W)
    sol, nxt = expected_modes(o, bs, k, count, IAC, W)
    Nz = IAC.shape[0]
    m = np.ones(IAC.shape, bool)
    m[1:, 0, o.dimQ:, :] = False         # alpha rows beyond dimQ are not used
    m[0, 0, o.dimQ:, 1:] = False
    scale = np.abs(nxt[m]).max()
    np.testing.assert_allclose(out[m], nxt[m], rtol=0, atol=1e-10 * scale)
    # row Nz/2 is left as it was; its Gauss-point values are zero
    np.testing.assert_array_equal(out[Nz // 2], IAC[Nz // 2])
    assert not np.any(IAU[Nz // 2])


@pytest.mark.parametrize("k", [0, 2])
def test_forcing_enters_before_the_solve_and_in_the_next_rhs(case, k):
    o, rng = case
    bs = basis(o)
    Nz, Mx = o.cshape[0], o.cshape[3]
    count = 5 - k
    IAC, W = inputs(o, rng)
    F = rand_c(rng, (2, 2, Nz, Mx, o.dimR))
    Z = np.zeros(o.ushape, complex)
    out, _, _ = o.increadjproject(k, count, IAC, Z, Z, W, F)
    _, nxt = expected_modes(o, bs, k, count, IAC, W, F[0], F[1])
    m = np.ones(IAC.shape, bool)
    m[1:, 0, o.dimQ:, :] = False
    m[0, 0, o.dimQ:, 1:] = False
    np.testing.assert_allclose(out[m], nxt[m], rtol=0, atol=1e-10 * np.abs(nxt[m]).max())


def test_gauss_point_values_and_wall_values(case):
    o, rng = case
    bs = basis(o)
    Nz, Mx = o.cshape[0], o.cshape[3]
    nQ = o.dimQ
    IAC, W = inputs(o, rng)
    Z = np.zeros(o.ushape, complex)
    k = 2
    sol, IAU, IAUb = o.increadjproject(k, 3, IAC, Z, Z, W)
    # coefficients -> Gauss points through the pinned rebuild, with the
    # omega_y sums over the first dimQ coefficients (the reference's loops)
    tr = sol.copy()
    tr[:, 1, nQ:, :] = 0
    tr[0, 1, :, 0] = sol[0, 1, :, 0]
    base = o.coef_to_gauss(tr)
    Kx, Kz, K2 = k2_of(o)
    y = bs["y"]
    lv, lv1, lv2 = (np.polyval(np.polyder(LV, n), y) for n in range(3))
    lg, lg1 = np.polyval(LG, y), np.polyval(np.polyder(LG), y)
    ub, wb = W[0], W[1]
    exp = base.copy()
    exp[Nz // 2] = 0
    for z in range(Nz):
        if z == Nz // 2:
            continue
        for x in range(Mx):
            if z == 0 and x == 0:
                exp[0, 0, :, 0] += ub[0, 0] * lg
                exp[0, 3, :, 0] += ub[0, 0] * lg1
                exp[0, 2, :, 0] += wb[0, 0] * lg
                exp[0, 4, :, 0] += wb[0, 0] * lg1
                continue
            kx, kz, k2 = Kx[x], Kz[z], K2[z, x]
            f, g = -lv1 * wb[z, x], lg * ub[z, x]          # -dv/dy and omega_y of the lifts
            qq, p = lv2 * wb[z, x], lg1 * ub[z, x]
            exp[z, 1, :, x] += lv * wb[z, x]
            # u = (i kx dv/dy - i kz omega_y)/k2, w = (i kz dv/dy + i kx omega_y)/k2
            exp[z, 0, :, x] += -1j * (kx * f + kz * g) / k2
            exp[z, 2, :, x] += 1j * (kx * g - kz * f) / k2
            exp[z, 3, :, x] += 1j * (kx * qq - kz * p) / k2
            exp[z, 4, :, x] += 1j * (kz * qq + kx * p) / k2
    np.testing.assert_allclose(IAU, exp, rtol=0, atol=1e-11 * np.abs(exp).max())
    # wall values: du/dy and dw/dy at y = -1 of the rebuilt field
    from numpy.polynomial import legendre as L
    q2 = L.legval(-1.0, L.legfit(y, bs["Qpp"], y.size - 1))
    assert np.allclose(q2, 8.0 * (-1.0) ** np.arange(nQ))
    rp0 = o.array("Rp0")
    ex = np.zeros((2, Nz, Mx), complex)
    for z in range(Nz):
        if z == Nz // 2:
            continue
        for x in range(Mx):
            if z == 0 and x == 0:
                ex[0, 0, 0] = rp0 @ sol[0, 0, :, 0] - ub[0, 0] / 2
                ex[1, 0, 0] = rp0 @ sol[0, 1, :, 0] - wb[0, 0] / 2
                continue
            kx, kz, k2 = Kx[x], Kz[z], K2[z, x]
            v2 = q2 @ sol[z, 0, :nQ, x]                    # d2v/dy2 of the coefficients
            w1 = rp0 @ sol[z, 1, :, x] + ub[z, x] * np.polyval(np.polyder(LG), -1.0)
            v2l = wb[z, x] * np.polyval(np.polyder(LV, 2), -1.0)
            # du/dy = (i kx v'' - i kz omega')/k2, dw/dy = (i kz v'' + i kx omega')/k2;
            # the reference doubles the alpha sum's share of their imaginary
            # parts, i.e. the real part of v''
            dv = 2 * v2.real + 1j * v2.imag
            ex[0, z, x] = (1j * kx * (dv + v2l) - 1j * kz * w1) / k2
            ex[1, z, x] = (1j * kz * (dv + v2l) + 1j * kx * w1) / k2
    np.testing.assert_allclose(IAUb, ex, rtol=0, atol=1e-10 * np.abs(ex).max())


@pytest.mark.parametrize("k", [0, 1, 2])
def test_reduces_to_the_adjoint_stage(k):
    """Zero walls, zero tangent states, no forcing: the adjoint stage."""
    o = pyorc.Oracle(8, 17, 8, 2.0, 1.3, RE_NUM, 4.0 / 3.0, DT, NSTEPS)
    rng = np.random.default_rng(11 + k)
    o.primal(0)
    o.tangent(0, NSTEPS, np.zeros(o.cshape, complex), 0)
    assert not np.any(o.getincrestage(3))
    IAC = rand_c(rng, o.cshape)
    IU, LIU = rand_c(rng, o.ushape), rand_c(rng, o.ushape)
    W = np.zeros((6,) + o.cshape[:1] + o.cshape[3:], complex)
    out, _, _ = o.increadjproject(k, 3, IAC, IU, LIU, W)
    ref, _ = o.adjproject_stage(k, IAC, IU, LIU)
    Nz = IAC.shape[0]
    m = np.ones(IAC.shape, bool)
    m[0, :, :, 0] = False
    m[Nz // 2] = False
    m[:, 0, o.dimQ:, :] = False
    np.testing.assert_allclose(out[m], ref[m], rtol=0, atol=1e-12 * np.abs(ref[m]).max())


def test_mean_mode_forcing_parts(case):
    """The stage's forcing reaches the mean mode through its real part only;
    the next stage's through both parts (channel/increadjproject0.c:130-141,
    290-300)."""
    o, rng = case
    Nz, Mx = o.cshape[0], o.cshape[3]
    IAC, W = inputs(o, rng)
    Z = np.zeros(o.ushape, complex)
    F = np.zeros((2, 2, Nz, Mx, o.dimR), complex)
    base, _, _ = o.increadjproject(1, 3, IAC, Z, Z, W, F)
    F[0, :, 0, 0, :] = 1j * rand_c(rng, (2, o.dimR)).real
    same, _, _ = o.increadjproject(1, 3, IAC, Z, Z, W, F)
    np.testing.assert_array_equal(same[0, :, :, 0], base[0, :, :, 0])
    F[1, :, 0, 0, :] = 1j * rand_c(rng, (2, o.dimR)).real
    moved, _, _ = o.increadjproject(1, 3, IAC, Z, Z, W, F)
    dt = o.scalar("dt")
    np.testing.assert_allclose(moved[0, :, :, 0] - base[0, :, :, 0], dt * D[2] * F[1, :, 0, 0, :],
                               rtol=0, atol=1e-14)


@pytest.mark.parametrize("grid", [(8, 17, 8), (12, 21, 10)])
def test_incremental_products_are_two_adjoint_product_sets(grid):
    """pass2's incremental side (channel/pass2.c:227-330: IS = u x I lambda +
    Iu x lambda, IT = lambda x I omega + I lambda x omega) is bilinear: it
    equals the pinned adjoint operator with the state u and I lambda plus the
    same with the tangent Iu and lambda; zero when I lambda and Iu vanish."""
    Nx, Ny, Nz = grid
    o = pyorc.Oracle(Nx, Ny, Nz, 2.0, 1.3, RE_NUM, 4.0 / 3.0, DT, 1)

    def field(seed):
        r = np.random.default_rng(seed)
        C = r.standard_normal(o.cshape) + 1j * r.standard_normal(o.cshape)
        C[Nz // 2] = 0
        C[0, :, :, 0] = C[0, :, :, 0].real
        return o.coef_to_gauss(C)

    U, IU, AU, IAU = field(1), field(2), field(3), field(4)
    IH = o.pass2_incremental(U, IU, AU, IAU)
    ref = o.pass2(U, IAU) + o.pass2(IU, AU)
    m = np.ones(o.ushape, bool)
    m[Nz // 2] = False
    np.testing.assert_allclose(IH[m], ref[m], rtol=0, atol=1e-12 * np.abs(ref[m]).max())
    assert not np.any(IH[Nz // 2])
    Z = np.zeros(o.ushape, complex)
    assert np.abs(o.pass2_incremental(U, Z, AU, Z)).max() == 0.0
