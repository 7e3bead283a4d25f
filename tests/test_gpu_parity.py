"""GPU parity: the HIP library (through the C ABI, via c_channel) against the
CPU oracle on the same seeded states.

Tolerances (relative L2 over the coefficient array):
  * SHORT = 1e-10 for one to four steps.  The Legendre-Galerkin solves are
    ill-conditioned, so any change of rounding order moves the result by a
    few 1e-11: the oracle built with FMA contraction differs from the plain
    oracle by 4.9e-11 after one step of the 32x32x65 case, 1.3e-11 (run-up),
    3.1e-11 (tangent), 4.1e-11 (adjoint), 9.2e-11 (projection); see
    tests/test_oracle_rounding_floor.py.  The device path lands at 1.2-1.5x
    those floors.
  * PROJ = 3e-10 for the time-derivative projection (ddt_project), whose own
    FMA/plain floor is 9.2e-11 (a difference quotient of stored states).
  * 1e-8 after 100 primal steps on the 32x32x65 minimal channel (Lx = pi,
    Lz = 0.3 pi, Re = 2000, dt = 0.005).
"""
import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

G0 = dict(Nx=32, Ny=65, Nz=32, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)
SMALL = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.c_destroy()


SHORT = 1e-10
PROJ = 3e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def start(cc, cfg, nsteps, seed=0, amp=0.02):
    c = cfg
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], nsteps)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     nsteps) == 0
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=seed, amp=amp)
    return o, C0


def direction(cfg, seed):
    c = cfg
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=seed,
                           unit=True)


def test_primal_one_step_minimal_channel(cc):
    print("case: primal 1 step, 32x32x65", flush=True)
    o, C0 = start(cc, G0, 1)
    Co, Cg = C0.copy(), C0.copy()
    o.primal(0, Co)
    cc.c_primal(0, Cg)
    assert cc.last_status() == 0
    assert rel(Cg, Co) <= SHORT
    assert rel(cc.c_getsoln(1), o.getsoln(1)) <= SHORT
    np.testing.assert_array_equal(cc.c_getsoln(0), C0)


def test_primal_hundred_steps_minimal_channel(cc):
    print("case: primal 100 steps, 32x32x65", flush=True)
    o, C0 = start(cc, G0, 100)
    Co, Cg = C0.copy(), C0.copy()
    o.primal(0, Co)
    cc.c_primal(0, Cg)
    assert rel(Cg, Co) <= 1e-8
    for i in (1, 10, 50):
        assert rel(cc.c_getsoln(i), o.getsoln(i)) <= 1e-8


def test_primal_with_run_up(cc):
    print("case: primal run-up 2 + 2 steps", flush=True)
    o, C0 = start(cc, SMALL, 2, amp=0.1)
    Co, Cg = C0.copy(), C0.copy()
    o.primal(2, Co)
    cc.c_primal(2, Cg)
    assert rel(Cg, Co) <= SHORT
    for i in range(3):
        assert rel(cc.c_getsoln(i), o.getsoln(i)) <= SHORT


def test_laminar_start_with_rand(cc):
    print("case: laminar + rand() start", flush=True)
    c = SMALL
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], 3)
    cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], 3)
    pyorc.load().orc_srand(1)
    cc.sc_srand(1)
    o.primal(0, None)
    cc.c_primal(0, np.zeros([0, 0, 0, 0], complex))
    for i in range(4):
        assert rel(cc.c_getsoln(i), o.getsoln(i)) <= SHORT


def test_tangent_homogeneous_and_inhomogeneous(cc):
    print("case: tangent 4 steps", flush=True)
    o, C0 = start(cc, SMALL, 4, amp=0.1)
    o.primal(0, C0.copy())
    cc.c_primal(0, C0.copy())
    for inhomo in (0, 1):
        d = direction(SMALL, 1 + inhomo)
        Io, Ig = d.copy(), d.copy()
        o.tangent(0, 4, Io, inhomo)
        cc.c_tangent(0, 4, Ig, inhomo)
        assert rel(Ig, Io) <= SHORT
        for i in range(5):
            assert rel(cc.c_getincresoln(i), o.getincresoln(i)) <= SHORT
    d = direction(SMALL, 5)
    Io, Ig = d.copy(), d.copy()
    o.tangent(1, 3, Io, 0)
    cc.c_tangent(1, 3, Ig, 0)
    assert rel(Ig, Io) <= SHORT


def test_adjoint_with_and_without_forcing(cc):
    print("case: adjoint 4 steps", flush=True)
    o, C0 = start(cc, SMALL, 4, amp=0.1)
    o.primal(0, C0.copy())
    cc.c_primal(0, C0.copy())
    d = direction(SMALL, 3)
    o.tangent(0, 4, d.copy(), 0)
    cc.c_tangent(0, 4, d.copy(), 0)
    for strength in (0.0, 1.0):
        a = direction(SMALL, 4)
        Ao, Ag = a.copy(), a.copy()
        o.adjoint(4, 0, Ao, 0, strength)
        cc.c_adjoint(4, 0, Ag, 0, strength)
        assert rel(Ag, Ao) <= SHORT


def test_ddt_project_and_spec2phys(cc):
    print("case: ddt_project and spec2phys", flush=True)
    o, C0 = start(cc, SMALL, 3, amp=0.1)
    o.primal(0, C0.copy())
    cc.c_primal(0, C0.copy())
    for i in (0, 2, 3):
        v = direction(SMALL, 10 + i)
        vo, vg = v.copy(), v.copy()
        ro = o.ddt_project(i, vo)
        rg = cc.c_ddt_project(i, vg)
        assert abs(rg - ro) <= PROJ * max(1.0, abs(ro))
        assert rel(vg, vo) <= PROJ
    flow = np.zeros((3, 3 * SMALL["Nx"] // 2, 3 * SMALL["Ny"] // 2, 3 * SMALL["Nz"] // 2))
    cc.c_spec2phys(C0, flow)
    fo = o.spec2phys(C0)
    assert rel(flow, fo) <= SHORT


RAGGED = dict(Nx=20, Ny=17, Nz=10, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)


def test_ragged_grid_ddt_project_and_adjoint(cc):
    """Nx/2 = 10 columns (not a multiple of the inner-product kernel's 4
    columns per block, so its last block is partly empty), Nz = 10, and
    transform lengths 30 and 15 (generic-length FFT kernels): ddt_project and
    a forced adjoint step against the oracle."""
    print("case: ragged grid 20x17x10, ddt_project + adjoint", flush=True)
    o, C0 = start(cc, RAGGED, 2, amp=0.1)
    o.primal(0, C0.copy())
    cc.c_primal(0, C0.copy())
    for i in (0, 1, 2):
        v = direction(RAGGED, 20 + i)
        vo, vg = v.copy(), v.copy()
        ro = o.ddt_project(i, vo)
        rg = cc.c_ddt_project(i, vg)
        assert abs(rg - ro) <= PROJ * max(1.0, abs(ro))
        assert rel(vg, vo) <= PROJ
    # a tangent first: MIC nonzero, so the adjoint's forcing is active
    d = direction(RAGGED, 29)
    o.tangent(0, 2, d.copy(), 0)
    cc.c_tangent(0, 2, d.copy(), 0)
    a = direction(RAGGED, 30)
    Ao, Ag = a.copy(), a.copy()
    o.adjoint(2, 0, Ao, 0, 1.0)
    cc.c_adjoint(2, 0, Ag, 0, 1.0)
    assert rel(Ag, Ao) <= SHORT


@pytest.mark.parametrize("grid", ["16x33x12", "64x17x64"])
def test_field_cache_is_bitwise(cc, grid):
    """The tangent and adjoint read the primal's transformed velocity and
    vorticity from the per-state cache instead of rebuilding and transforming
    MC[j]: both paths must give identical bits, including state 0 after a
    run-up (where the cache entry is not used).  On the 96-point grid the
    adjoint's first two stages run their two passes as one merged x-kernel
    launch when both states are cached; the same bits again."""
    print("case: transformed-field cache on/off, " + grid, flush=True)
    c, n = (SMALL if grid == "16x33x12" else X96), 3
    out = {}
    try:
        for enable in (True, False):
            cc.sc_set_field_cache(enable)
            for ru in (0, 2):
                assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                                 c["dt"], n) == 0
                assert cc.sc_field_cache_active() == enable
                C = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"],
                                    seed=0, amp=0.1)
                cc.c_primal(ru, C)
                I = direction(c, 5)
                cc.c_tangent(0, n, I, 0)
                A = direction(c, 6)
                cc.c_adjoint(n, 0, A, 0, 1.0)
                out[(enable, ru)] = (C, I, A)
    finally:
        cc.sc_set_field_cache(True)
    for ru in (0, 2):
        for a, b in zip(out[(True, ru)], out[(False, ru)]):
            np.testing.assert_array_equal(a, b)


# Nx = Nz = 64 selects the fixed-length (96-point) transform kernels; Ny = 17
# keeps the oracle fast.  The benchmark grid itself (64x64x97) is checked
# against the oracle for one step with tolerances set by its own rounding floor
# (tests/test_oracle_rounding_floor.py: 9.7e-11 step, 1.8e-9 adjoint).
X96 = dict(Nx=64, Ny=17, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0, dt=0.004)
BENCH = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0, dt=0.004)


def test_fixed_length_transforms(cc):
    print("case: 96-point transforms, 64x16x64, primal/tangent/adjoint 2 steps", flush=True)
    n = 2
    o, C0 = start(cc, X96, n)
    Co, Cg = C0.copy(), C0.copy()
    o.primal(0, Co)
    cc.c_primal(0, Cg)
    assert rel(Cg, Co) <= SHORT
    d = direction(X96, 3)
    Io, Ig = d.copy(), d.copy()
    o.tangent(0, n, Io, 0)
    cc.c_tangent(0, n, Ig, 0)
    assert rel(Ig, Io) <= SHORT
    a = direction(X96, 4)
    Ao, Ag = a.copy(), a.copy()
    o.adjoint(n, 0, Ao, 0, 1.0)
    cc.c_adjoint(n, 0, Ag, 0, 1.0)
    assert rel(Ag, Ao) <= SHORT


def _forced_adjoint_step(cc, o, cfg):
    """One tangent step (so that MIC[0], MIC[3] and the adjoint's forcing
    are nonzero), ddt_project of a direction at steps 0 and 1, and one
    forced adjoint step, device against oracle: the relative errors."""
    d = direction(cfg, 3)
    Io, Ig = d.copy(), d.copy()
    o.tangent(0, 1, Io, 0)
    cc.c_tangent(0, 1, Ig, 0)
    errs = {"tangent": rel(Ig, Io)}
    for i in (0, 1):
        v = direction(cfg, 10 + i)
        vo, vg = v.copy(), v.copy()
        ro, rg = o.ddt_project(i, vo), cc.c_ddt_project(i, vg)
        errs["ddt%d" % i] = max(rel(vg, vo), abs(rg - ro) / max(1.0, abs(ro)))
    a = direction(cfg, 4)
    Ao, Ag = a.copy(), a.copy()
    o.adjoint(1, 0, Ao, 0, 1.0)
    cc.c_adjoint(1, 0, Ag, 0, 1.0)
    errs["adjoint"] = rel(Ag, Ao)
    print(" ".join("%s %.2e" % kv for kv in errs.items()), flush=True)
    return errs


def test_benchmark_grid_one_step(cc):
    """64x64x97: one primal, tangent and forced adjoint step and ddt_project.
    The FMA/plain floors of the oracle here (tests/test_oracle_rounding_floor.py)
    are 9.7e-11 (step), 1.7e-9 (tangent), 7.3e-10 and 4.1e-9 (ddt_project at
    steps 0 and 1), 1.8e-9 (forced adjoint); the device lands at 1.0-1.05x
    those, and the bounds allow 3x."""
    print("case: benchmark grid 64x64x97, primal + tangent + adjoint 1 step", flush=True)
    o, C0 = start(cc, BENCH, 1)
    Co, Cg = C0.copy(), C0.copy()
    o.primal(0, Co)
    cc.c_primal(0, Cg)
    assert rel(Cg, Co) <= 3e-10
    e = _forced_adjoint_step(cc, o, BENCH)
    assert e["tangent"] <= 5e-9
    assert e["ddt0"] <= 2.2e-9 and e["ddt1"] <= 1.2e-8
    assert e["adjoint"] <= 5.4e-9


def test_duality_on_device(cc):
    print("case: tangent/adjoint duality", flush=True)
    n = 5
    o, C0 = start(cc, SMALL, n, amp=0.1)
    cc.c_primal(0, C0.copy())
    y, w = cc.c_quad()
    IC0, AC0 = direction(SMALL, 20), direction(SMALL, 21)
    IC, AC = IC0.copy(), AC0.copy()
    cc.c_tangent(0, n, IC, 0)
    cc.c_adjoint(n, 0, AC, 0, 0.0)

    def phys(C):
        f = np.zeros((3, 3 * SMALL["Nx"] // 2, len(y), 3 * SMALL["Nz"] // 2))
        cc.c_spec2phys(np.ascontiguousarray(C), f)
        return f

    left = np.sum(phys(AC0) * phys(IC) * w[None, None, :, None])
    right = np.sum(phys(AC) * phys(IC0) * w[None, None, :, None])
    assert abs(left - right) <= 1e-10 * abs(left)


def test_cfl_violation_stops_the_run(cc):
    print("case: CFL violation", flush=True)
    c = dict(SMALL)
    c["dt"] = 5.0
    o, C0 = start(cc, c, 3, amp=0.1)
    Cg = C0.copy()
    cc.c_primal(0, Cg)
    assert cc.last_status() == -1
    np.testing.assert_array_equal(Cg, C0)
    assert np.all(cc.c_getsoln(1) == 0)


def test_invalid_calls_raise(cc):
    print("case: invalid calls", flush=True)
    o, C0 = start(cc, SMALL, 2)
    cc.c_primal(0, C0.copy())
    with pytest.raises(AssertionError):
        cc.c_tangent(1, 1, C0.copy(), 0)
    with pytest.raises(AssertionError):
        cc.c_adjoint(0, 1, C0.copy(), 0, 0.0)
    with pytest.raises(AssertionError):
        cc.c_primal(0, np.zeros((SMALL["Nz"], 2, SMALL["Ny"] - 2, 3), complex))


def test_tangent_forcing(cc):
    """Re-sensitivity forcing of stored states (channel/tangent_forcing.c,
    channel/tangent_forcing0.c) on the device against the oracle: a direct
    band product, so agreement is at rounding level."""
    print("case: tangent forcing of stored states", flush=True)
    o, C0 = start(cc, SMALL, 2, amp=0.1)
    o.primal(0, C0.copy())
    cc.c_primal(0, C0.copy())
    for count in (0, 1, 4, 6):
        got = cc.sc_tangent_forcing(count)
        ref = o.tangent_forcing(count)
        for g, r in zip(got, ref):
            assert np.linalg.norm(g - r) <= 1e-12 * np.linalg.norm(r)


@pytest.mark.parametrize("cfg", [SMALL, G0, BENCH], ids=["16x33x12", "32x65x32", "64x97x64"])
def test_wall_plane_and_incre_boundary(cc, cfg):
    """The wall plane of a state (channel/initAlphaBeta2.c:70-81, 210-223) and
    increBoundary on it and on random planes (channel/increBoundary.c:20-162)
    against the oracle.  The random planes carry a nonzero row Nz/2 (not
    read) and a non-Hermitian kx = 0 column (read through its real part in
    z).  Both are short sums and products, so agreement is at rounding
    level: 1e-13 (wall plane), 1e-12 (increBoundary)."""
    print("case: wall plane + increBoundary, %dx%dx%d" % (cfg["Nx"], cfg["Ny"], cfg["Nz"]),
          flush=True)
    o, C0 = start(cc, cfg, 1, amp=0.1)
    ub, wb = cc.sc_wall_plane(C0)
    rub, rwb = o.wall_plane(C0)
    assert rel(ub, rub) <= 1e-13 and rel(wb, rwb) <= 1e-13
    rng = np.random.default_rng(21)
    Nz, Mx = cfg["Nz"], cfg["Nx"] // 2
    planes = [(ub, wb)]
    planes.append(tuple(rng.standard_normal((Nz, Mx)) + 1j * rng.standard_normal((Nz, Mx))
                        for _ in range(2)))
    for x0, z0 in planes:
        gx, gz = x0.copy(), z0.copy()
        cc.sc_incre_boundary(gx, gz)
        rx, rz = o.incre_boundary(x0, z0)
        assert rel(gx, rx) <= 1e-12 and rel(gz, rz) <= 1e-12
        assert np.all(gx[Nz // 2] == 0) and np.all(gz[Nz // 2] == 0)
    with pytest.raises(TypeError):
        cc.sc_incre_boundary(np.zeros((Nz, Mx + 1), complex), np.zeros((Nz, Mx), complex))


def test_device_resident_entry_points(cc):
    """sc_primal_dev / sc_adjoint_dev (states in device memory) run the same
    launches as c_primal / c_adjoint on host arrays: bitwise equal results,
    and the primal trajectory they leave agrees with the oracle's."""
    print("case: device-resident primal/adjoint, 32x32x65", flush=True)
    o, C0 = start(cc, G0, 3)
    Ch = C0.copy()
    cc.c_primal(0, Ch)
    ACh = direction(G0, 5)
    AC0 = ACh.copy()
    cc.c_adjoint(3, 0, ACh, 0, 1.0)
    Cd, ACd = cc.DeviceArray(C0), cc.DeviceArray(AC0)
    try:
        cc.sc_primal_dev(0, Cd)
        Cs = [cc.c_getsoln(i).copy() for i in range(4)]
        cc.sc_adjoint_dev(3, 0, ACd, 1.0)
        assert np.array_equal(Cd.download(), Ch)
        assert np.array_equal(ACd.download(), ACh)
    finally:
        Cd.free()
        ACd.free()
    o.primal(0, C0.copy())
    for i in range(1, 4):
        assert rel(Cs[i], o.getsoln(i)) <= 1e-9
    with pytest.raises(AssertionError):
        cc.sc_adjoint_dev(3, 0, 0, 1.0)


G4 = dict(Nx=128, Ny=129, Nz=128, Lx=2 * np.pi, Lz=np.pi, Re=5000.0, flux=4.0 / 3.0, dt=0.002)


def test_largest_grid_one_step(cc):
    """The largest configuration, 128x128x129 (dimR = 127): the chunked
    rebuild, the 28-k-step projections and generic 192-point transforms.
    One primal, tangent and forced adjoint step and ddt_project against the
    oracle.  The rounding floors at this grid (the oracle with and without
    FMA contraction, tests/test_oracle_rounding_floor.py) are 9.8e-10
    (step), 1.9e-8 (tangent), 8.1e-9 and 1.2e-9 (ddt_project at steps 0 and
    1), 1.8e-8 (forced adjoint); the device lands at about 1.05x those, and
    the bounds allow 3x."""
    print("case: largest grid 128x128x129, primal + tangent + adjoint 1 step", flush=True)
    o, C0 = start(cc, G4, 1)
    Co, Cg = C0.copy(), C0.copy()
    o.primal(0, Co)
    cc.c_primal(0, Cg)
    ep = rel(Cg, Co)
    print("largest grid: step %.2e" % ep, flush=True)
    e = _forced_adjoint_step(cc, o, G4)
    assert ep <= 3e-9
    assert e["tangent"] <= 6e-8
    assert e["ddt0"] <= 2.5e-8 and e["ddt1"] <= 4e-9
    assert e["adjoint"] <= 6e-8
