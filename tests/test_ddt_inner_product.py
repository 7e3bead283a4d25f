"""The identity behind the device's ddt_project inner products.

channel/channel.c:704-780 forms <f, g> = sum_y W_y sum_{x,z} (u_f u_g + v_f v_g
+ w_f w_g) over the velocities rebuilt on the dealiased grid (spec2phys).  The
device forms the same number from the coefficients with the Gauss-rule Gram
bands of the setup:

    per column (kz, kx) != (0, 0):  c Re[a_f^H (Qs + Qps/k2) a_g + b_f^H (Rs/k2) b_g]
    mean column:                    Re a_f^T Rs Re a_g + Re b_f^T Rs Re b_g

(c = 2 for kx > 0; at kx = 0 the Hermitian part (c(kz) + conj c(-kz))/2 of
the coefficients; row Nz/2 left out), times Px Pz.  This test evaluates that
closed form in numpy from the oracle's Gram bands and wavenumbers and checks
it against the oracle's own grid sum of its spec2phys fields, on structured
and on unstructured (non-Hermitian kx = 0 column, nonzero Nyquist row)
coefficient arrays.  A dropped 1/k2, a wrong Hermitian weight or a missing
mean term fails it.
"""
import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

LX, LZ = 2.0, 1.3


@pytest.fixture(scope="module", params=[(16, 17, 12), (12, 25, 8)])
def orc(request):
    Nx, Ny, Nz = request.param
    return pyorc.Oracle(Nx, Ny, Nz, LX, LZ, 1000.0, 4.0 / 3.0, 0.01, 1)


def band_mult(B, x, dl):
    n = len(x)
    y = np.zeros(n, complex)
    for m in range(-dl, dl + 1):
        for i in range(n):
            j = i + 2 * m
            if 0 <= j < n:
                y[i] += B[i, dl + m] * x[j]
    return y


def gram_inner(o, F, G):
    Nz, Mx, nQ, nR = o.Nz, o.Nx // 2, o.dimQ, o.dimR
    Qs = o.array("Qs").reshape(nQ, 9)
    Qps = o.array("Qps").reshape(nQ, 9)
    Rs = o.array("Rs").reshape(nR, 5)
    Kx, Kz = o.array("Kx"), o.array("Kz")
    zm = (-np.arange(Nz)) % Nz

    def herm(C):
        H = C.copy()
        H[:, :, :, 0] = 0.5 * (C[:, :, :, 0] + np.conj(C[zm, :, :, 0]))
        return H

    F, G = herm(F), herm(G)
    tot = 0.0
    for z in range(Nz):
        if z == Nz // 2:
            continue
        for x in range(Mx):
            if z == 0 and x == 0:
                tot += np.vdot(F[0, 0, :, 0], band_mult(Rs, G[0, 0, :, 0], 2)).real
                tot += np.vdot(F[0, 1, :, 0], band_mult(Rs, G[0, 1, :, 0], 2)).real
                continue
            k2 = Kx[x] ** 2 + Kz[z] ** 2
            c = 2.0 if x > 0 else 1.0
            ta = np.vdot(F[z, 0, :nQ, x], band_mult(Qs + Qps / k2, G[z, 0, :nQ, x], 4))
            tb = np.vdot(F[z, 1, :, x], band_mult(Rs / k2, G[z, 1, :, x], 2))
            tot += c * (ta + tb).real
    return tot * (3 * o.Nx // 2) * (3 * Nz // 2)


def grid_inner(o, F, G):
    y, w = o.quad()
    f, g = o.spec2phys(F), o.spec2phys(G)
    return float(np.einsum("cxyz,cxyz,y->", f, g, w))


def test_gram_form_equals_grid_sum(orc):
    o = orc
    F = synthetic_state(o.Nx, o.Ny, o.Nz, LX, LZ, 4.0 / 3.0, seed=1, amp=0.1)
    G = synthetic_state(o.Nx, o.Ny, o.Nz, LX, LZ, 4.0 / 3.0, seed=2, amp=0.1)
    for a, b in ((F, G), (F, F), (G - F, F)):
        ref = grid_inner(o, a, b)
        got = gram_inner(o, a, b)
        assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)


def test_gram_form_on_unstructured_coefficients(orc):
    o = orc
    rng = np.random.default_rng(7)
    shape = o.cshape
    F = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    G = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = grid_inner(o, F, G)
    got = gram_inner(o, F, G)
    scale = np.sqrt(grid_inner(o, F, F) * grid_inner(o, G, G))
    assert abs(got - ref) <= 1e-12 * scale, (got, ref)
