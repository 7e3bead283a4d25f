"""Pin the oracle's sensitivity machinery.

* The Reynolds-number tangent (orc_tangent_src with re_src): the stage
  operators M -+ re b_k dt D / M + re a_k dt D depend on re = 1/Re, so the
  tangent of the primal in Re gains at stage k the right-hand-side source
  dt (a_k TF(Y_k) + b_k TF(X_k)) with TF = -re^2 D the forcing of
  channel/tangent_forcing.c (the reference computes it, but never wires it
  in: channel/channel.c:519-522, increproject0.c:46-90).  Checked against
  central differences of the primal in Re, and its linearity in the
  source's combination with the mean-velocity source.
* The mean column of the adjoint.  The reference corrects the adjoint's
  mean flux after each stage solve through the solve's own matrix
  (channel/adjproject0.c:82-95, 153-167), which is the transpose of the
  tangent's mass-matrix correction (channel/increproject0.c:131-144) only
  to O(re dt): duality with directions that have a mean column is off by
  ~1e-4.  The exact transpose (P = I - add w1^T / 2 is self-adjoint in the
  mean column's pairing a^T Rs b, and falls on the right-hand side before
  the solves of adjoint stages 0 and 1 and on the state at each step's
  start) restores duality to rounding; the objective adjoint uses it.
* The objective adjoint (orc_adjoint_objective) of the time-averaged
  lower-wall shear J = w sum_m tau(C_m), tau = sum_j R_j'(-1) Re alpha_j(0,0)
  (channel/project0.c:227-234): for every tangent d from the window's
  start, <AC_0, d>_phys = w sum_m tau(T_m d) (the physical inner product of
  drivers/testAdj.py:55-56), tau read from the stored tangent states.
"""
import numpy as np

import pyorc
from channel.synthetic import synthetic_state

CFG = dict(Nx=8, Ny=17, Nz=8, Lx=np.pi, Lz=0.3 * np.pi, Re=400.0, dt=0.01)
FLUX = 4.0 / 3.0


def _oracle(n, Re=None, flux=FLUX):
    c = CFG
    return pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], Re or c["Re"], flux, c["dt"],
                        n)


def _state(seed, **kw):
    c = CFG
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], FLUX, seed=seed, **kw)


def tau(o, C):
    """Lower-wall shear of the mean profile: sum_j R_j'(-1) Re alpha_j(0,0)."""
    return float(o.array("Rp0") @ C[0, 0, :, 0].real)


def test_reynolds_tangent_matches_central_differences():
    """Fourth-order central differences with a step of 2 in Re = 400: the
    states move by only ~1e-4 of their size per unit Re, so smaller steps
    lose digits to cancellation (measured agreement 3e-9 to 1e-8).  The
    source through the corrected instead of the uncorrected mean solution
    misses by 8e-4."""
    n = 4
    C0 = _state(0, amp=0.3)
    e = 2.0

    def run(Re):
        o = _oracle(n, Re=Re)
        C = C0.copy()
        o.primal(0, C)
        traj = [o.getsoln(i) for i in range(n + 1)]
        o.destroy()
        return traj
    f = [run(CFG["Re"] + k * e) for k in (-2, -1, 1, 2)]
    o = _oracle(n)
    o.primal(0, C0.copy())
    IC = np.zeros_like(C0)
    o.tangent_src(0, n, IC, 0, 1)
    for i in range(1, n + 1):
        fd = (f[0][i] - 8 * f[1][i] + 8 * f[2][i] - f[3][i]) / (12 * e)
        rel = np.linalg.norm(o.getincresoln(i) - fd) / np.linalg.norm(fd)
        assert rel < 2e-8, (i, rel)
    # no source, no tangent
    IZ = np.zeros_like(C0)
    o.tangent_src(0, n, IZ, 0, 0)
    assert np.all(IZ == 0)
    # the two sources superpose with a homogeneous part
    d = _state(2, unit=True)
    parts = []
    for inh, re in ((0, 0), (1, 0), (0, 1)):
        v = (d if (inh, re) == (0, 0) else np.zeros_like(d)).copy()
        o.tangent_src(0, n, v, inh, re)
        parts.append(v)
    both = d.copy()
    o.tangent_src(0, n, both, 1, 1)
    assert np.linalg.norm(both - sum(parts)) <= 1e-12 * np.linalg.norm(both)


def test_objective_adjoint_is_the_gradient_of_wall_shear():
    n = 5
    w = 0.37
    o = _oracle(n)
    o.primal(0, _state(0, amp=0.3))
    y, wq = o.quad()

    def ip(a, b):
        return float(np.sum(o.spec2phys(a) * o.spec2phys(b) * wq[None, None, :, None]))
    AC = np.zeros(o.cshape, complex)
    o.adjoint_objective(n, 0, AC, w)
    for seed in (3, 4):
        d = _state(seed, unit=True)
        o.tangent(0, n, d.copy(), 0)
        dJ = w * sum(tau(o, o.getincresoln(m)) for m in range(n))
        g = ip(AC, d)
        assert abs(g - dJ) <= 1e-10 * abs(dJ), (g, dJ)
    # a later start carries the adjoint state: the same gradient in two halves
    A2 = np.zeros_like(AC)
    o.adjoint_objective(n, 2, A2, w)
    o.adjoint_objective(2, 0, A2, w)
    assert np.linalg.norm(A2 - AC) <= 1e-13 * np.linalg.norm(AC)
    o.destroy()


def _duality_with_mean(o, n, seed=5):
    """Relative duality gap for directions whose mean columns are set."""
    y, wq = o.quad()

    def ip(a, b):
        return float(np.sum(o.spec2phys(a) * o.spec2phys(b) * wq[None, None, :, None]))
    rng = np.random.default_rng(seed)
    d, x = _state(3, unit=True), _state(4, unit=True)
    d[0, 0, :, 0] = rng.standard_normal(o.dimR)
    d[0, 1, :, 0] = rng.standard_normal(o.dimR)
    x[0, 0, :, 0] = rng.standard_normal(o.dimR)
    x[0, 1, :, 0] = rng.standard_normal(o.dimR)
    IC = d.copy()
    o.tangent(0, n, IC, 0)
    AC = x.copy()
    o.adjoint(n, 0, AC, 0, 0.0)
    return abs(ip(x, IC) - ip(AC, d)) / abs(ip(x, IC))


def test_exact_mean_transpose_restores_duality():
    n = 3
    o = _oracle(n)
    o.primal(0, _state(0, amp=0.3))
    ref = _duality_with_mean(o, n)
    o.L.orc_set_exact_mean_adjoint(1)
    try:
        exact = _duality_with_mean(o, n)
    finally:
        o.L.orc_set_exact_mean_adjoint(0)
    o.destroy()
    assert 1e-6 < ref < 1e-2, ref          # the reference's mean-column gap
    assert exact < 1e-11, exact
