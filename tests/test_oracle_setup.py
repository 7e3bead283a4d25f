"""Pin the oracle's setup (quadrature, bases, Gram bands, wavenumbers) against
numpy's independent Gauss-Legendre and Legendre-series arithmetic.

Reference: channel/LegendreSetup.c:25-527, channel/RsQs.c:55-261,
channel/waveNumsAndCFL.c:22-96.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as L

import pyorc

ONE_MINUS_Y2 = np.array([2.0 / 3.0, 0.0, -2.0 / 3.0])   # 1 - y^2 as a Legendre series


def basis_series(j, power):
    c = np.zeros(j + 1)
    c[j] = 1.0
    for _ in range(power):
        c = L.legmul(c, ONE_MINUS_Y2)
    return c


@pytest.fixture(scope="module", params=[(8, 17, 8), (16, 33, 6), (32, 65, 32)])
def orc(request):
    Nx, Ny, Nz = request.param
    o = pyorc.Oracle(Nx, Ny, Nz, 2.0, 1.3, 1000.0, 4.0 / 3.0, 0.01, 1)
    return o


def test_gauss_nodes_and_weights(orc):
    y, w = orc.quad()
    yr, wr = L.leggauss(orc.qpts)
    assert np.all(np.diff(y) > 0)
    np.testing.assert_allclose(y, yr, rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, wr, rtol=1e-11, atol=0)
    np.testing.assert_array_equal(y, -y[::-1])
    # Gauss rule with q nodes integrates y^k exactly for k <= 2q - 1
    for k in range(0, 2 * orc.qpts, 2):
        assert abs(np.sum(w * y ** k) - 2.0 / (k + 1)) < 1e-15 * (1 + k)


def test_bases_and_derivatives(orc):
    y, _ = orc.quad()
    q, nQ, nR = orc.qpts, orc.dimQ, orc.dimR
    R = orc.array("R").reshape(q, nR)
    Rp = orc.array("Rp").reshape(q, nR)
    Q = orc.array("Q").reshape(q, nQ)
    Qp = orc.array("Qp").reshape(q, nQ)
    Qpp = orc.array("Qpp").reshape(q, nQ)
    for j in range(nR):
        c = basis_series(j, 1)
        np.testing.assert_allclose(R[:, j], L.legval(y, c), atol=1e-12 * (1 + j))
        np.testing.assert_allclose(Rp[:, j], L.legval(y, L.legder(c)), atol=1e-11 * (1 + j) ** 2)
    for j in range(nQ):
        c = basis_series(j, 2)
        np.testing.assert_allclose(Q[:, j], L.legval(y, c), atol=1e-12 * (1 + j))
        np.testing.assert_allclose(Qp[:, j], L.legval(y, L.legder(c)), atol=1e-11 * (1 + j) ** 2)
        np.testing.assert_allclose(Qpp[:, j], L.legval(y, L.legder(c, 2)), atol=1e-10 * (1 + j) ** 3)
    Rp0 = orc.array("Rp0")
    Rpp0 = orc.array("Rpp0")
    for j in range(nR):
        c = basis_series(j, 1)
        assert Rp0[j] == pytest.approx(L.legval(-1.0, L.legder(c)), abs=1e-9 * (1 + j) ** 2)
        assert Rpp0[j] == pytest.approx(L.legval(-1.0, L.legder(c, 2)), abs=1e-8 * (1 + j) ** 4)


def _dense_from_band(B, N, dl):
    A = np.zeros((N, N))
    for i in range(N):
        for m in range(-dl, dl + 1):
            j = i + 2 * m
            if 0 <= j < N:
                A[i, j] = B[i, dl + m]
    return A


def test_gram_bands_match_exact_integrals(orc):
    """Band storage (entry [i][dl+m] = matrix element (i, i+2m)) equals the
    exact L2 inner products, computed by Legendre-series integration."""
    nQ, nR = orc.dimQ, orc.dimR
    names = [("Qs", nQ, 4, 2, 0), ("Qps", nQ, 4, 2, 1), ("Qpps", nQ, 4, 2, 2),
             ("Rs", nR, 2, 1, 0), ("Rps", nR, 2, 1, 1)]
    for name, N, dl, power, der in names:
        B = orc.array(name).reshape(N, 2 * dl + 1)
        A = _dense_from_band(B, N, dl)
        series = [L.legder(basis_series(j, power), der) if der else basis_series(j, power)
                  for j in range(N)]
        # spot-check a set of rows against exact integration
        for i in sorted(set([0, 1, 2, N // 2, N - 2, N - 1])):
            for j in range(max(0, i - 2 * dl - 2), min(N, i + 2 * dl + 3)):
                prod = L.legmul(series[i], series[j])
                integ = L.legint(prod)
                exact = L.legval(1.0, integ) - L.legval(-1.0, integ)
                scale = max(1.0, abs(exact))
                assert abs(A[i, j] - exact) <= 1e-10 * scale * (1 + i) ** 2, (name, i, j)
        assert np.allclose(A, A.T, rtol=0, atol=0)


def test_weighted_transposes(orc):
    y, w = orc.quad()
    q, nQ, nR = orc.qpts, orc.dimQ, orc.dimR
    for name, base, n in [("Qw", "Q", nQ), ("Qpw", "Qp", nQ), ("Qppw", "Qpp", nQ),
                          ("Rw", "R", nR), ("Rpw", "Rp", nR)]:
        Mw = orc.array(name).reshape(n, q)
        M = orc.array(base).reshape(q, n)
        np.testing.assert_array_equal(Mw, (w[:, None] * M).T)


def test_wavenumbers_and_cfl(orc):
    Nx, Nz = orc.Nx, orc.Nz
    Lx, Lz = 2.0, 1.3
    kx = orc.array("Kx")
    kz = orc.array("Kz")
    np.testing.assert_allclose(kx, 2 * np.pi * np.fft.rfftfreq(Nx, d=Lx / Nx)[:Nx // 2], rtol=1e-15)
    ref = 2 * np.pi * np.fft.fftfreq(Nz, d=Lz / Nz)
    mask = np.arange(Nz) != Nz // 2
    np.testing.assert_allclose(kz[mask], ref[mask], rtol=1e-14)
    y, _ = orc.quad()
    cfl2 = orc.array("cfl2")
    np.testing.assert_allclose(cfl2[1:-1], 2 * np.pi / (y[2:] - y[:-2]), rtol=1e-15)
    assert cfl2[0] == cfl2[-1] == pytest.approx(np.pi / (y[0] + 1.0))
    assert orc.scalar("cfl1") == pytest.approx(np.pi * Nx / Lx)
    assert orc.scalar("cfl3") == pytest.approx(np.pi * Nz / Lz)


def test_glibc_rand_stream():
    # the first outputs of glibc's unseeded rand() (srand(1))
    assert pyorc.glibc_rand_sequence(5, 1) == [1804289383, 846930886, 1681692777,
                                               1714636915, 1957747793]
