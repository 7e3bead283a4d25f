"""The Python wrapper `channel` (the counterpart of the reference's
channel/__init__.py:1-300, the layer the LSS drivers call) on the device
library, against the oracle on the same seeded states.

Covered: para_init and the stored trajectory (get_solution), tangent and
adjoint through the wrapper, the objectives velAvg / IvelAvg / uavgGrad
(channel/__init__.py:217-299) against the same reductions of the oracle's
physical fields, spec2phys from an array, a step index and a file name, the
save/read round trip ('.hd5' appended, channel/__init__.py:32-55), and the
wrapper's asserts (double init, step ranges).

Tolerances: 1e-10 on the stored states and the tangent/adjoint (1-4 steps,
the same bar as tests/test_gpu_parity.py), 1e-9 on the averaged objectives
(sums of physical fields of those states).
"""
import os

import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

CFG = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, meanU=2.0 / 3.0, dt=0.005)
N = 4


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def run():
    import channel
    for k, v in CFG.items():
        setattr(channel, k, v)
    c = CFG
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], 2 * c["meanU"], seed=3, amp=0.05)
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], 2 * c["meanU"], c["dt"], N)
    o.primal(0, C0.copy())
    channel.para_init(N, C0.copy())
    yield channel, o
    channel.destroy()


def oracle_mean_u(o, C):
    return o.spec2phys(C)[0].mean(2).mean(0)


def test_trajectory_and_spec2phys(run):
    print("case: wrapper para_init + trajectory, 16x33x12", flush=True)
    ch, o = run
    for i in range(N + 1):
        assert rel(ch.get_solution(i), o.getsoln(i)) <= 1e-10
    C2 = ch.get_solution(2)
    f_arr = ch.spec2phys(C2)
    f_idx = ch.spec2phys(2)
    assert np.array_equal(f_arr, f_idx)
    assert rel(f_arr, o.spec2phys(o.getsoln(2).copy())) <= 1e-10
    view = ch.get_solution(2, copy=False)
    assert not view.flags.owndata


def test_velAvg_profile_and_centreline(run):
    print("case: wrapper velAvg", flush=True)
    ch, o = run
    T = N * CFG["dt"]
    ref = sum((CFG["dt"] / T) * oracle_mean_u(o, o.getsoln(i).copy()) for i in range(N))
    got = ch.velAvg(0, N, T)
    assert got.shape == (o.qpts,)
    assert rel(got, ref) <= 1e-9
    mid = o.qpts // 2 + 1
    assert abs(ch.velAvg(0, N, T, profile=False) - ref[mid]) <= 1e-9 * abs(ref[mid])
    with pytest.raises(AssertionError):
        ch.velAvg(0, N + 1, T)


def test_tangent_IvelAvg_uavgGrad(run):
    print("case: wrapper tangent + IvelAvg + uavgGrad", flush=True)
    ch, o = run
    c = CFG
    IC0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], 2 * c["meanU"], seed=4,
                          unit=True)
    Ig, Io = IC0.copy(), IC0.copy()
    ch.tangent(0, N, Ig, 1)
    o.tangent(0, N, Io, 1)
    assert rel(Ig, Io) <= 1e-10
    T = N * c["dt"]
    vref = sum((c["dt"] / T) * oracle_mean_u(o, o.getincresoln(i).copy()) for i in range(N))
    assert rel(ch.IvelAvg(0, N, T), vref) <= 1e-9
    # uavgGrad over two chunks with time-dilation terms
    bounds = [0, 2, N]
    zeta = [0.3, -0.7]
    uavg = ch.velAvg(0, N, T)
    ref = np.zeros(o.qpts)
    for k in range(2):
        s, e = bounds[k], bounds[k + 1]
        va = sum((c["dt"] / T) * oracle_mean_u(o, o.getincresoln(i).copy()) for i in range(s, e))
        uend = oracle_mean_u(o, o.getsoln(e).copy())
        ref = ref + va + (1.0 / T) * zeta[k] * (uend - uavg)
    assert rel(ch.uavgGrad(2, bounds, T, uavg, zeta), ref) <= 1e-9
    with pytest.raises(AssertionError):
        ch.tangent(2, 1, IC0.copy(), 0)


def test_adjoint_through_wrapper(run):
    print("case: wrapper adjoint", flush=True)
    ch, o = run
    c = CFG
    a = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], 2 * c["meanU"], seed=5, unit=True)
    Ag, Ao = a.copy(), a.copy()
    ch.adjoint(N, 0, Ag, 0, 1.0)
    o.adjoint(N, 0, Ao, 0, 1.0)
    assert rel(Ag, Ao) <= 1e-10
    with pytest.raises(AssertionError):
        ch.adjoint(0, 1, a.copy(), 0, 1.0)


def test_save_read_round_trip(run, tmp_path):
    print("case: wrapper save/read_solution", flush=True)
    ch, o = run
    C = ch.get_solution(N)
    base = os.path.join(str(tmp_path), "state")
    ch.save_solution(base, C)
    assert os.path.exists(base + ".hd5")
    # the reference's HDF5 datasets, read back on the host
    import c_channel
    alpha = c_channel.h5_read(base + ".hd5", "/data_alpha")
    assert np.array_equal(alpha, np.transpose(C[:, 0], (1, 0, 2)))
    assert c_channel.h5_read(base + ".hd5", "data_Nx").tolist() == [C.shape[3] * 2]
    assert c_channel.h5_read(base + ".hd5", "data_Nz").tolist() == [C.shape[0]]
    back = ch.read_solution(base)
    assert np.array_equal(back, C)
    assert np.array_equal(ch.spec2phys(base), ch.spec2phys(C))


def test_double_init_is_refused(run):
    ch, o = run
    with pytest.raises(AssertionError):
        ch.para_init(N, ch.get_solution(0))


def _mean_profile_poly(o, C):
    """The plane-mean u profile of C as a Legendre series in y (degree
    qpts - 1, exact for the profile's polynomial of degree dimR + 1), fitted
    through the oracle's physical field at the Gauss points."""
    y, w = o.quad()
    u = oracle_mean_u(o, C)
    return np.polynomial.legendre.Legendre.fit(y, u, o.qpts - 1, domain=[-1, 1])


def test_wall_shear_against_the_profile_derivative(run):
    """wallShear (the Uxb(0,0) of channel/project0.c:227-234 at y = -1, and
    the same at y = +1) time-averaged over the window, against the
    derivative at the walls of the Legendre series through the oracle's
    physical plane means; also the tangent's (incre 1)."""
    print("case: wrapper wallShear", flush=True)
    ch, o = run
    T = N * CFG["dt"]
    ref = np.zeros(2)
    iref = np.zeros(2)
    for i in range(N):
        d = _mean_profile_poly(o, o.getsoln(i).copy()).deriv()
        ref += (CFG["dt"] / T) * np.array([d(-1.0), d(1.0)])
        di = _mean_profile_poly(o, o.getincresoln(i).copy()).deriv()
        iref += (CFG["dt"] / T) * np.array([di(-1.0), di(1.0)])
    got = ch.wallShear(0, N, T)
    assert ref[0] > 0 and ref[1] < 0          # drag on both walls
    assert rel(got, ref) <= 1e-9
    assert rel(ch.wallShear(0, N, T, incre=1), iref) <= 1e-9
    # the wall plane of one state (sc_wall_plane) gives the same lower-wall value
    import c_channel
    ub, _ = c_channel.sc_wall_plane(ch.get_solution(2))
    one = ch.wallShear(2, 3, CFG["dt"])
    assert abs(one[0] - ub[0, 0].real) <= 1e-12 * abs(ub[0, 0].real)


def test_plane_statistics_against_physical_moments(run):
    """statistics (channel/comp_stat.c:9-230: plane means of u, v, w, their
    second moments and du/dy) from the spectra, against the same moments of
    the oracle's physical field (mean over the dealiased x-z grid)."""
    print("case: wrapper statistics", flush=True)
    ch, o = run
    for i, incre in ((1, 0), (N, 0), (2, 1)):
        C = (o.getincresoln(i) if incre else o.getsoln(i)).copy()
        f = o.spec2phys(C)
        pm = lambda a: a.mean(axis=2).mean(axis=0)
        ref = [pm(f[0]), pm(f[1]), pm(f[2]), pm(f[0] * f[1]), pm(f[0] * f[2]), pm(f[1] * f[2]),
               pm(f[0] * f[0]), pm(f[1] * f[1]), pm(f[2] * f[2]),
               o.coef_to_gauss(C)[0, 3, :, 0].real]
        got = ch.statistics(i, incre)
        assert got.shape == (10, o.qpts)
        for q in range(10):
            scale = max(np.abs(ref[q]).max(), 1e-3 * np.abs(ref[6]).max())
            assert np.abs(got[q] - ref[q]).max() <= 1e-11 * scale, q
