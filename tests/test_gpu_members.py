"""Members: several solutions advanced together in every kernel launch
(sc_*_members), each against the single-member run (c_primal / c_tangent /
c_adjoint) of the same input.  Every member's kernels do the same
arithmetic in the same order as a single run, so the results are equal bit
for bit.

Cases: independent trajectories (primal, then tangent and forced adjoint
along each member's own trajectory) and the shared base of the
least-squares-shadowing segments (tangent directions along member 0's
trajectory, one of them with the mean-velocity source, then their forced
adjoints), on three grids that select different kernel families:
64x17x64 (fixed-length transforms, the merged adjoint x-kernel), 16x33x12
(generic-length transforms) and 8x129x8 (dimR = 127: the chunked rebuild and
the staged-basis projection)."""
import numpy as np
import pytest

from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

GRIDS = {
    "64x17x64": dict(Nx=64, Ny=17, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0,
                     dt=0.004),
    "16x33x12": dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0,
                     dt=0.005),
    "8x129x8": dict(Nx=8, Ny=129, Nz=8, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0,
                    dt=0.002),
}
NM, N = 3, 3


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.c_destroy()
    c_channel.sc_set_members(1)


def _st(cfg, seed, **kw):
    c = cfg
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=seed, **kw)


def _init(cc, cfg, members):
    c = cfg
    cc.sc_set_members(members)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     N) == 0
    assert cc.sc_members() == members


def _run_members(cc, fn, arrays, *args):
    d = cc.DeviceArray(np.ascontiguousarray(np.stack(arrays)))
    try:
        fn(*args[:-1], d, *args[-1])
        return d.download()
    finally:
        d.free()


@pytest.mark.parametrize("grid", list(GRIDS))
def test_members_equal_single_runs(cc, grid):
    cfg = GRIDS[grid]
    print("case: %d members, %s: independent and shared-base runs" % (NM, grid), flush=True)
    _init(cc, cfg, NM)
    C0 = [_st(cfg, s, amp=0.1) for s in range(NM)]
    D0 = [_st(cfg, 10 + s, unit=True) for s in range(NM)]
    A0 = [_st(cfg, 20 + s, unit=True) for s in range(NM)]
    # independent trajectories: primal, tangent, forced adjoint per member
    Cb = _run_members(cc, cc.sc_primal_members, C0, 0, ())
    trajb = [[cc.sc_member_state(0, m, i) for i in range(N + 1)] for m in range(NM)]
    Ib = _run_members(cc, cc.sc_tangent_members, D0, 0, N, (0b010, False))
    Ab = _run_members(cc, cc.sc_adjoint_members, A0, N, 0, (1.0, False))
    for m in range(NM):
        C = C0[m].copy()
        cc.c_primal(0, C)
        np.testing.assert_array_equal(Cb[m], C)
        for i in range(N + 1):
            np.testing.assert_array_equal(trajb[m][i], cc.c_getsoln(i))
        I = D0[m].copy()
        cc.c_tangent(0, N, I, 1 if m == 1 else 0)
        np.testing.assert_array_equal(Ib[m], I)
        A = A0[m].copy()
        cc.c_adjoint(N, 0, A, 0, 1.0)
        np.testing.assert_array_equal(Ab[m], A)
    # shared base: every member along member 0's trajectory
    C = C0[0].copy()
    Cs = _run_members(cc, cc.sc_primal_members, C0, 0, ())
    np.testing.assert_array_equal(Cs[0], Cb[0])
    Is = _run_members(cc, cc.sc_tangent_members, D0, 0, N, (0b100, True))
    mic = [[cc.sc_member_state(1, m, i) for i in range(N + 1)] for m in range(NM)]
    As = _run_members(cc, cc.sc_adjoint_members, A0, N, 0, (1.0, True))
    cc.c_primal(0, C)                     # member 0's trajectory again
    for m in range(NM):
        I = D0[m].copy()
        cc.c_tangent(0, N, I, 1 if m == 2 else 0)
        np.testing.assert_array_equal(Is[m], I)
        for i in range(N + 1):
            np.testing.assert_array_equal(mic[m][i], cc.c_getincresoln(i))
        A = A0[m].copy()
        cc.c_adjoint(N, 0, A, 0, 1.0)
        np.testing.assert_array_equal(As[m], A)


WIN = [6, 2, 5, 3]


@pytest.mark.parametrize("grid", ["16x33x12", "64x17x64"])
def test_member_windows_equal_single_runs(cc, grid):
    """Uneven windows along one base trajectory (sc_set_member_windows):
    member m's tangent over steps 0 .. WIN[m] (members 1 and 2 with the
    mean-velocity source) and its forced adjoint from WIN[m] back to 0,
    against c_tangent(0, WIN[m]) and c_adjoint(WIN[m], 0) on member 0's
    trajectory; eager, recorded and replayed graph runs alike."""
    cfg = GRIDS[grid]
    c = cfg
    nm, n = len(WIN), max(WIN)
    print("case: %d members with windows %s, %s" % (nm, WIN, grid), flush=True)
    cc.sc_set_members(nm)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     n) == 0
    C0 = [_st(cfg, s, amp=0.1) for s in range(nm)]
    D0 = [_st(cfg, 10 + s, unit=True) for s in range(nm)]
    A0 = [_st(cfg, 20 + s, unit=True) for s in range(nm)]
    _run_members(cc, cc.sc_primal_members, C0, 0, ())
    with pytest.raises(AssertionError):
        cc.sc_set_member_windows([-1, 2, 5, 3])
    with pytest.raises(AssertionError):
        cc.sc_set_member_windows([n + 1, 2, 5, 3])
    cc.sc_set_member_windows(WIN)
    runs = []
    try:
        for graphs in (0, 1, 1):
            cc.sc_set_graphs(graphs)
            Iw = _run_members(cc, cc.sc_tangent_members, D0, 0, n, (0b0110, True))
            ends = [cc.sc_member_state(1, m, WIN[m]) for m in range(nm)]
            after = [[cc.sc_member_state(1, m, i) for i in range(WIN[m] + 1, n + 1)]
                     for m in range(nm)]
            Aw = _run_members(cc, cc.sc_adjoint_members, A0, n, 0, (1.0, True))
            runs.append((Iw, ends, after, Aw))
    finally:
        cc.sc_set_graphs(0)
    for Iw, ends, after, Aw in runs[1:]:
        for m in range(nm):
            np.testing.assert_array_equal(ends[m], runs[0][1][m])
            np.testing.assert_array_equal(Aw[m], runs[0][3][m])
    Iw, ends, after, Aw = runs[0]
    for m in range(nm):
        # a member past its window is zero, and so are its stored states
        if WIN[m] < n:
            assert not np.any(Iw[m])
        for a in after[m]:
            assert not np.any(a)
    cc.sc_set_member_windows(None)
    C = C0[0].copy()
    cc.c_primal(0, C)                     # member 0's trajectory again
    for m in range(nm):
        I = D0[m].copy()
        cc.c_tangent(0, WIN[m], I, 1 if m in (1, 2) else 0)
        np.testing.assert_array_equal(ends[m], I)
        A = A0[m].copy()
        cc.c_adjoint(WIN[m], 0, A, 0, 1.0)   # forced by member m's tangent, stored just now
        np.testing.assert_array_equal(Aw[m], A)
    cc.c_destroy()
    cc.sc_set_members(1)


def test_member_count_is_checked(cc):
    with pytest.raises(AssertionError):
        cc.sc_set_members(0)
    with pytest.raises(AssertionError):
        cc.sc_set_members(33)
