"""Pin the oracle's wall-plane functions by independent routes.

Reference: channel/initAlphaBeta2.c:70-81, 210-223 (the wall plane Uxb, Uzb of
a coefficient array) and channel/increBoundary.c:20-162 (the padded product
of both wall planes with the fixed single-mode plane w.n, then the rewrite of
channel/increBoundary.c:137-158).

- The wall plane is checked against the y-derivatives of the rebuilt
  velocity at the wall: with the alpha (v) coefficients of the nonzero modes
  set to zero, Uxb and Uzb are du/dy and dw/dy at y = -1, obtained here by
  projecting the Gauss-point values of the U slots onto Legendre polynomials
  (exact for their degree) and evaluating the series at -1.
- increBoundary is checked against its closed form: w.n is the real field
  2 cos(kx1 x + kz1 z) of the mode (kx, kz) = (1, 1), so the product's
  spectrum is the input spectrum shifted by (+1, +1) plus the one shifted by
  (-1, -1) (Hermitian extension to kx < 0, the kx = 0 column taken through
  its real part in z as the c2r transform does); the 3/2 padding makes the
  product alias-free, so the forward transform returns exactly that shifted
  sum.  No FFT is used on this side.
"""
import numpy as np
import pytest

import pyorc

LX, LZ = 2.0, 1.3


@pytest.fixture(scope="module", params=[(8, 17, 8), (16, 17, 12), (12, 25, 10)])
def orc(request):
    Nx, Ny, Nz = request.param
    return pyorc.Oracle(Nx, Ny, Nz, LX, LZ, 1000.0, 4.0 / 3.0, 0.01, 1)


def kz_signed(Nz):
    r = np.arange(Nz)
    return np.where(r <= Nz // 2, r, r - Nz)


def closed_form_incre_boundary(Uxb, Uzb, Nx, Nz):
    Mx = Nx // 2
    ks = kz_signed(Nz)

    def full(P):
        # A[kz + Nz/2, kx + Mx], kz in [-Nz/2, Nz/2], kx in [-Mx, Mx]
        A = np.zeros((Nz + 1, 2 * Mx + 1), complex)
        for r in range(Nz):
            if r == Nz // 2:
                continue
            kz = ks[r]
            for x in range(1, Mx):
                A[kz + Nz // 2, x + Mx] = P[r, x]
                A[-kz + Nz // 2, -x + Mx] = np.conj(P[r, x])
            # kx = 0: the c2r transform keeps the real part of the column in
            # physical z, i.e. the Hermitian part of its spectrum
            rm = (Nz - r) % Nz
            A[kz + Nz // 2, Mx] = 0.5 * (P[r, 0] + np.conj(P[rm, 0]))
        return A

    out = []
    for P in (Uxb, Uzb):
        A = full(P)
        G = np.zeros((Nz, Mx), complex)
        for r in range(Nz):
            if r == Nz // 2:
                continue
            kz = ks[r]
            for x in range(Mx):
                G[r, x] = (A[kz - 1 + Nz // 2, x - 1 + Mx] + A[kz + 1 + Nz // 2, x + 1 + Mx])
        out.append(G)
    gx, gz = out
    Kx = 2 * np.pi * np.arange(Mx) / LX
    Kz = 2 * np.pi * ks / LZ
    kx = Kx[None, :]
    kz = Kz[:, None]
    nx = 1j * kz * gx - 1j * kx * gz
    nz = -1j * kx * gx - 1j * kz * gz
    nx[0, 0], nz[0, 0] = gx[0, 0], gz[0, 0]
    return nx, nz


def random_plane(rng, Nz, Mx):
    return rng.standard_normal((Nz, Mx)) + 1j * rng.standard_normal((Nz, Mx))


def test_incre_boundary_closed_form(orc):
    rng = np.random.default_rng(11)
    Nz, Mx = orc.Nz, orc.Nx // 2
    # row Nz/2 and a non-Hermitian kx = 0 column included on purpose
    ub, wb = random_plane(rng, Nz, Mx), random_plane(rng, Nz, Mx)
    got_x, got_z = orc.incre_boundary(ub, wb)
    ref_x, ref_z = closed_form_incre_boundary(ub, wb, orc.Nx, Nz)
    scale = max(np.abs(ref_x).max(), np.abs(ref_z).max())
    assert np.abs(got_x - ref_x).max() <= 1e-13 * scale
    assert np.abs(got_z - ref_z).max() <= 1e-13 * scale
    assert np.all(got_x[Nz // 2] == 0) and np.all(got_z[Nz // 2] == 0)


def test_incre_boundary_is_linear(orc):
    rng = np.random.default_rng(12)
    Nz, Mx = orc.Nz, orc.Nx // 2
    a = [random_plane(rng, Nz, Mx) for _ in range(4)]
    x1, z1 = orc.incre_boundary(a[0], a[1])
    x2, z2 = orc.incre_boundary(a[2], a[3])
    x3, z3 = orc.incre_boundary(a[0] - 2.5 * a[2], a[1] - 2.5 * a[3])
    assert np.abs(x3 - (x1 - 2.5 * x2)).max() <= 1e-12 * np.abs(x3).max()
    assert np.abs(z3 - (z1 - 2.5 * z2)).max() <= 1e-12 * np.abs(z3).max()


def wall_value(f, y, w):
    """Value at y = -1 of the polynomial (degree < len(y)) through the
    Gauss-point values f[..., i] (Gauss-Legendre projection, exact)."""
    n = len(y)
    P = np.polynomial.legendre.legvander(y, n - 1)          # [i][n]
    c = (f * w) @ P * ((2 * np.arange(n) + 1) / 2.0)
    return c @ ((-1.0) ** np.arange(n))


def test_wall_plane_is_wall_derivative(orc):
    rng = np.random.default_rng(13)
    Nz, Mx, nR = orc.Nz, orc.Nx // 2, orc.dimR
    C = np.zeros(orc.cshape, complex)
    C[:, 1] = rng.standard_normal((Nz, nR, Mx)) + 1j * rng.standard_normal((Nz, nR, Mx))
    C[0, 0, :, 0] = rng.standard_normal(nR) + 1j * rng.standard_normal(nR)   # mean u
    ub, wb = orc.wall_plane(C)
    U = orc.coef_to_gauss(C)
    y, w = orc.quad()
    du = np.moveaxis(U[:, 3], 1, 2)          # [Nz][Mx][qpts]
    dw = np.moveaxis(U[:, 4], 1, 2)
    ref_u = wall_value(du, y, w)
    ref_w = wall_value(dw, y, w)
    scale = max(np.abs(ref_u).max(), np.abs(ref_w).max())
    assert np.abs(ub - ref_u).max() <= 1e-11 * scale
    assert np.abs(wb - ref_w).max() <= 1e-11 * scale
    # the v coefficients of the nonzero modes do not enter the wall plane
    C2 = C.copy()
    C2[:, 0] += rng.standard_normal(C2[:, 0].shape)
    C2[0, 0, :, 0] = C[0, 0, :, 0]
    ub2, wb2 = orc.wall_plane(C2)
    assert np.array_equal(ub2, ub) and np.array_equal(wb2, wb)
