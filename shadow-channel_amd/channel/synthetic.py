"""Seeded synthetic channel states written directly as Galerkin coefficients.

The layout is the reference's C array (channel/channel.i:34-53):
``C[Nz, 2, dimR, Nx/2]`` complex128, alpha (wall-normal velocity v, Q basis
(1-y^2)^2 P_j) in ``C[:, 0]`` and beta (wall-normal vorticity omega_y, R basis
(1-y^2) P_j) in ``C[:, 1]``.  Every such state is divergence free and has
no-slip walls by construction of the bases.

Shape of the perturbation: alpha_j, beta_j ~ N(0, 1) |k|^-1 0.8^j for each
Fourier mode (kx, kz) != (0, 0), giving a k^-2 energy spectrum and Legendre
coefficients decaying as 0.8^j, then scaled so the Gauss-weighted RMS of v and
of omega_y over the channel are each ``amp``.  The kx = 0 column is made
Hermitian (C[Nz-z] = conj C[z]), the Nyquist row z = Nz/2 and alpha rows
j >= dimQ are zero, and the mean mode carries the laminar profile
(3/4) flux (1 - y^2) = (3/4) flux R_0, i.e. alpha_0(0,0) = (3/4) flux.
"""
import numpy as np
from numpy.polynomial import legendre as npleg


def _gauss(qpts):
    y, w = npleg.leggauss(qpts)
    return y, w


def _legendre_table(y, n):
    P = np.zeros((y.size, n))
    for j in range(n):
        c = np.zeros(j + 1)
        c[j] = 1.0
        P[:, j] = npleg.legval(y, c)
    return P


def wavenumbers(Nx, Nz, Lx, Lz):
    kx = np.arange(Nx // 2) * (2 * np.pi / Lx)
    kz = np.array([i if i <= Nz // 2 else i - Nz for i in range(Nz)]) * (2 * np.pi / Lz)
    return kx, kz


def rms_v_omega(C, Nx, Ny, Nz, Lx, Lz):
    """Gauss-weighted RMS of v and omega_y of the fluctuating modes.

    Uses Parseval over (x, z): the mean of |f|^2 over a plane is the sum of
    |f_hat|^2 over the full (kx, kz) spectrum, where a stored kx > 0 column
    stands for two (its conjugate partner is implied).
    """
    qpts = 3 * Ny // 2
    dimR, dimQ = Ny - 2, Ny - 4
    y, w = _gauss(qpts)
    P = _legendre_table(y, dimR)
    om = (1 - y * y)[:, None]
    R = om * P
    Q = om * R[:, :dimQ]
    Mx = Nx // 2
    mult = np.where(np.arange(Mx) == 0, 1.0, 2.0)
    a = C[:, 0, :dimQ, :]              # [Nz, dimQ, Mx]
    b = C[:, 1, :, :]                  # [Nz, dimR, Mx]
    v = np.einsum("yj,zjx->zyx", Q, a)
    g = np.einsum("yj,zjx->zyx", R, b)
    mask = np.ones((Nz, Mx))
    mask[0, 0] = 0.0
    mask[Nz // 2, :] = 0.0
    ev = np.einsum("zyx,y,x,zx->", np.abs(v) ** 2, w, mult, mask)
    eg = np.einsum("zyx,y,x,zx->", np.abs(g) ** 2, w, mult, mask)
    return np.sqrt(ev / 2.0), np.sqrt(eg / 2.0)


def synthetic_state(Nx, Ny, Nz, Lx, Lz, flux, seed=0, amp=0.02, unit=False):
    """Return a seeded state C[Nz, 2, dimR, Nx/2] (complex128).

    unit=True returns a perturbation only (no mean flow) scaled to unit
    Euclidean coefficient norm, the shape used for tangent directions.
    """
    rng = np.random.default_rng(seed)
    dimR, dimQ = Ny - 2, Ny - 4
    Mx = Nx // 2
    kx, kz = wavenumbers(Nx, Nz, Lx, Lz)
    kmag = np.sqrt(kz[:, None] ** 2 + kx[None, :] ** 2)
    kmag[0, 0] = 1.0
    decay = 0.8 ** np.arange(dimR)
    C = np.zeros((Nz, 2, dimR, Mx), complex)
    for v in range(2):
        noise = rng.standard_normal((Nz, dimR, Mx)) + 1j * rng.standard_normal((Nz, dimR, Mx))
        C[:, v] = noise * decay[None, :, None] / kmag[:, None, :]
    C[:, 0, dimQ:, :] = 0.0
    C[Nz // 2] = 0.0
    C[0, :, :, 0] = 0.0
    # Hermitian kx = 0 column: C[Nz - z] = conj(C[z])
    for z in range(1, Nz // 2):
        C[Nz - z, :, :, 0] = np.conj(C[z, :, :, 0])
    if unit:
        C /= np.linalg.norm(C)
        return C
    rv, rg = rms_v_omega(C, Nx, Ny, Nz, Lx, Lz)
    C[:, 0] *= amp / rv
    C[:, 1] *= amp / rg
    C[0, 0, 0, 0] = 0.75 * flux
    return C
