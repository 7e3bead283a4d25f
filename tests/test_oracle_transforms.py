"""Pin the oracle's dealiased transforms and nonlinear terms against numpy.

Reference: channel/pass1.c:13-376 (padding layout, FFTW_BACKWARD unnormalised
inverse, c2r along x, scaled forward and truncation), channel/pass2.c:23-570
(adjoint convective operator), channel/mvOps.c (stride-2 banded LU).
The numpy side builds fields in physical space and differentiates them
spectrally with numpy.fft, so the curl and the cross products are checked by
an independent route.
"""
import numpy as np
import pytest

import pyorc

LX, LZ = 2.0, 1.3


def np_inv2d(S, Nx, Nz):
    Mx, Pz, Px = Nx // 2, 3 * Nz // 2, 3 * Nx // 2
    T = np.zeros((Pz, Px // 2 + 1), complex)
    T[:Nz // 2, :Mx] = S[:Nz // 2]
    T[Nz + 1:, :Mx] = S[Nz // 2 + 1:]
    T = np.fft.ifft(T, axis=0) * Pz
    T[:, 0] = T[:, 0].real
    return np.fft.irfft(T, n=Px, axis=1) * Px


def np_fwd2d(phys, Nx, Nz):
    Mx, Pz, Px = Nx // 2, 3 * Nz // 2, 3 * Nx // 2
    G = np.fft.rfft(phys, axis=1)[:, :Mx]
    G = np.fft.fft(G, axis=0) / (Px * Pz)
    S = np.zeros((Nz, Mx), complex)
    S[:Nz // 2] = G[:Nz // 2]
    S[Nz // 2 + 1:] = G[Nz + 1:]
    return S


def ddx(f):
    Px = f.shape[1]
    k = 2 * np.pi * np.fft.rfftfreq(Px, d=LX / Px)
    return np.fft.irfft(1j * k[None, :] * np.fft.rfft(f, axis=1), n=Px, axis=1)


def ddz(f):
    Pz = f.shape[0]
    k = 2 * np.pi * np.fft.fftfreq(Pz, d=LZ / Pz)
    return np.real(np.fft.ifft(1j * k[:, None] * np.fft.fft(f, axis=0), axis=0))


def rand_plane(rng, Nz, Mx):
    return rng.standard_normal((Nz, Mx)) + 1j * rng.standard_normal((Nz, Mx))


@pytest.fixture(scope="module", params=[(8, 9, 8), (16, 9, 12), (12, 9, 10)])
def orc(request):
    Nx, Ny, Nz = request.param
    return pyorc.Oracle(Nx, Ny, Nz, LX, LZ, 1000.0, 4.0 / 3.0, 0.01, 1)


def test_inverse_and_forward_transforms(orc):
    rng = np.random.default_rng(3)
    Nx, Nz, Mx = orc.Nx, orc.Nz, orc.Nx // 2
    S = rand_plane(rng, Nz, Mx)
    phys = orc.inv2d(S)
    np.testing.assert_allclose(phys, np_inv2d(S, Nx, Nz), rtol=0, atol=1e-12)
    back = orc.fwd2d(phys)
    np.testing.assert_allclose(back, np_fwd2d(phys, Nx, Nz), rtol=0, atol=1e-14)
    # forward(inverse(S)) returns S except where c2r drops Im at kx = 0
    # (Hermitian part) and the Nyquist row
    Sh = S.copy()
    Sh[Nz // 2] = 0
    herm = Sh.copy()
    for z in range(Nz):
        zc = (-z) % Nz
        herm[z, 0] = 0.5 * (Sh[z, 0] + np.conj(Sh[zc, 0]))
    herm[Nz // 2] = 0
    np.testing.assert_allclose(back, herm, atol=1e-13)


def test_pass1_nonlinear_term(orc):
    rng = np.random.default_rng(5)
    Nx, Nz, Mx, q = orc.Nx, orc.Nz, orc.Nx // 2, orc.qpts
    U = rng.standard_normal(orc.ushape) + 1j * rng.standard_normal(orc.ushape)
    rc, H = orc.pass1(U, 0)
    assert rc == 0
    for y in [0, q // 2, q - 1]:
        u, v, w, du, dw = [np_inv2d(U[:, s, y, :], Nx, Nz) for s in range(5)]
        ox = dw - ddz(v)
        oy = ddz(u) - ddx(w)
        oz = ddx(v) - du
        hx = v * oz - w * oy
        hy = w * ox - u * oz
        hz = u * oy - v * ox
        rows = np.arange(Nz) != Nz // 2
        for slot, h in [(0, hz), (1, hx), (2, hy)]:
            ref = np_fwd2d(h, Nx, Nz)
            np.testing.assert_allclose(H[rows, slot, y, :], ref[rows], atol=1e-11)
        # dudy, dwdy slots are untouched
        np.testing.assert_array_equal(H[:, 3:, y, :], U[:, 3:, y, :])


def test_pass2_adjoint_term(orc):
    rng = np.random.default_rng(7)
    Nx, Nz, Mx, q = orc.Nx, orc.Nz, orc.Nx // 2, orc.qpts
    U = rng.standard_normal(orc.ushape) + 1j * rng.standard_normal(orc.ushape)
    AU = rng.standard_normal(orc.ushape) + 1j * rng.standard_normal(orc.ushape)
    H = orc.pass2(U, AU)
    rows = np.arange(Nz) != Nz // 2
    for y in [0, q // 3, q - 1]:
        u, v, w, du, dw = [np_inv2d(U[:, s, y, :], Nx, Nz) for s in range(5)]
        lx, ly, lz = [np_inv2d(AU[:, s, y, :], Nx, Nz) for s in range(3)]
        ox = dw - ddz(v)
        oy = ddz(u) - ddx(w)
        oz = ddx(v) - du
        sx, sy, sz = v * lz - w * ly, w * lx - u * lz, u * ly - v * lx
        tx, ty, tz = ly * oz - lz * oy, lz * ox - lx * oz, lx * oy - ly * ox
        hx = -tx + ddz(sy)
        hy = -ty - ddz(sx) + ddx(sz)
        hz = -tz - ddx(sy)
        for slot, f in [(0, hz), (1, hx), (2, hy), (3, sx), (4, sz)]:
            ref = np_fwd2d(f, Nx, Nz)
            np.testing.assert_allclose(H[rows, slot, y, :], ref[rows], atol=1e-11)


def test_band_solver_against_dense():
    rng = np.random.default_rng(11)
    for N, dl in [(13, 4), (14, 4), (29, 2), (30, 2), (61, 4)]:
        B = rng.standard_normal((N, 2 * dl + 1))
        B[:, dl] += 8.0 * np.sign(B[:, dl]) + 8.0
        A = np.zeros((N, N))
        for i in range(N):
            for m in range(-dl, dl + 1):
                if 0 <= i + 2 * m < N:
                    A[i, i + 2 * m] = B[i, dl + m]
                else:
                    B[i, dl + m] = 0.0
        b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        x = pyorc.band_solve(B, b, dl)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)
