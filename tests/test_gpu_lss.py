"""Least-squares-shadowing KKT operator (channel/lss.py; drivers/wrapper.py:
6-54) on the device, 16x33x12, segments of 4 steps.

* Segments (the reference's loop on one trajectory) against the same loop
  run by the oracle (orc_tangent, orc_adjoint, orc_ddt_project): the KKT
  residual and the projection coefficients zeta, with and without the
  mean-velocity source, at the one-window bar of tests/test_gpu_windows.py
  (1e-9).
* BatchedSegments (every segment a member; one launch per stage for all
  segments) equal to Segments bit for bit, in FP64 and in mixed precision,
  for segments of equal length and of uneven length (3, 5 and 4 steps:
  the shorter segments end inside the members' common window).
* The drivers' solve: the right-hand side -K(0, inhomo 1) - K(0, inhomo 0)
  and a few MINRES iterations on the batched operator reduce the KKT
  residual (drivers/wrapper.py:78-134)."""
import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

G = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)
NSEG, L = 3, 4


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.sc_set_members(1)
    c_channel.sc_set_precision(0)
    c_channel.c_destroy()


def _C0():
    c = G
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0, amp=0.1)


def _x(seg, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(2 * seg.m * (2 * seg.n - 1))


class _OracleBackend:
    """The calls channel/lss.py makes, answered by the oracle."""

    def __init__(self, o):
        self.o = o

    def c_ddt_project(self, i, v):
        return self.o.ddt_project(i, v)

    def c_tangent(self, start, end, IC, inhomo):
        self.o.tangent(start, end, IC, inhomo)

    def c_adjoint(self, start, end, AC, inhomo, strength):
        self.o.adjoint(start, end, AC, inhomo, strength)


def _sequential(cc, x, inhomo, bounds=None):
    from channel.lss import Segments
    c = G
    bounds = bounds or [L * s for s in range(NSEG + 1)]
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     bounds[-1]) == 0
    cc.c_primal(0, _C0())
    seg = Segments((c["Nx"], c["Ny"], c["Nz"]), bounds)
    r = seg.matvec(x, inhomo)
    cc.c_destroy()
    return r, seg.zeta.copy()


def test_kkt_operator_against_oracle(cc):
    from channel import lss
    c = G
    print("case: KKT operator, %d segments of %d steps, vs oracle" % (NSEG, L), flush=True)
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     NSEG * L)
    o.primal(0, _C0())
    seg_o = lss.Segments((c["Nx"], c["Ny"], c["Nz"]), [L * s for s in range(NSEG + 1)])
    x = _x(seg_o, 1)
    for inhomo in (0, 1):
        saved = lss.c_channel
        lss.c_channel = _OracleBackend(o)
        try:
            ro = seg_o.matvec(x, inhomo)
            zo = seg_o.zeta.copy()
        finally:
            lss.c_channel = saved
        rg, zg = _sequential(cc, x, inhomo)
        e = np.linalg.norm(rg - ro) / np.linalg.norm(ro)
        ez = np.linalg.norm(zg - zo) / np.linalg.norm(zo)
        print("inhomo %d: residual rel err %.2e, zeta rel err %.2e" % (inhomo, e, ez), flush=True)
        assert e <= 1e-9 and ez <= 1e-9
    o.destroy()


@pytest.mark.parametrize("mixed,bounds", [(0, None), (1, None), (0, [0, 3, 8, 12]),
                                          (1, [0, 3, 8, 12])])
def test_batched_segments_equal_the_loop(cc, mixed, bounds):
    from channel.lss import BatchedSegments, Segments
    c = G
    print("case: batched segments vs the segment loop, mixed %d, bounds %s" % (mixed, bounds),
          flush=True)
    x = _x(Segments((c["Nx"], c["Ny"], c["Nz"]), [L * s for s in range(NSEG + 1)]), 2)
    ref = {}
    cc.sc_set_precision(mixed)
    try:
        for inhomo in (0, 1):
            ref[inhomo] = _sequential(cc, x, inhomo, bounds)
    finally:
        cc.sc_set_precision(0)
    if bounds is None:
        bs = BatchedSegments((c["Nx"], c["Ny"], c["Nz"]), (c["Lx"], c["Lz"]), c["Re"], c["flux"],
                             c["dt"], NSEG, L, mixed=bool(mixed))
    else:
        bs = BatchedSegments((c["Nx"], c["Ny"], c["Nz"]), (c["Lx"], c["Lz"]), c["Re"], c["flux"],
                             c["dt"], mixed=bool(mixed), bounds=bounds)
        assert bs.uneven and bs.L == 5
    bs.setup(_C0())
    try:
        for inhomo in (0, 1):
            r = bs.matvec(x, inhomo)
            np.testing.assert_array_equal(r, ref[inhomo][0])
            np.testing.assert_array_equal(bs.zeta, ref[inhomo][1])
    finally:
        bs.release()


def test_minres_on_the_batched_operator(cc):
    from channel.lss import BatchedSegments, minres
    c = G
    print("case: MINRES on the batched KKT operator", flush=True)
    bs = BatchedSegments((c["Nx"], c["Ny"], c["Nz"]), (c["Lx"], c["Lz"]), c["Re"], c["flux"],
                         c["dt"], NSEG, L)
    bs.setup(_C0())
    try:
        z = np.zeros(2 * bs.m * (2 * bs.n - 1))
        rhs = -bs.matvec(z, 1) - bs.matvec(z, 0)
        assert np.linalg.norm(rhs) > 0
        res = []
        x, (flag, it) = minres(lambda v: bs.matvec(v, 0), rhs, tol=1e-12, maxiter=6,
                               callback=lambda v: res.append(np.linalg.norm(bs.matvec(v, 0) - rhs)))
        print("residuals", " ".join("%.3e" % r for r in res), flush=True)
        assert len(res) == 6
        assert np.all(np.isfinite(x)) and res[-1] < 0.9 * np.linalg.norm(rhs)
    finally:
        bs.release()


def test_uneven_segments_at_benchmark_grid(cc):
    """Configuration 3 at the benchmark grid (64x64x97, Re 3000, dt 0.004),
    scaled to 8 segments: lengths drawn from 100 .. 400 steps, every segment
    a member (one launch per stage for all, HIP graphs on, launch shapes from
    the autotuner's cache), 8 tangent directions and one inhomogeneous
    tangent from zero per segment.  Each direction's end states and
    segment-boundary projection coefficients equal the segment loop's bit
    for bit; the projection residuals (the coefficient left after a second
    projection at each boundary) are reported and held to 1e-12."""
    from channel import tuning
    from channel.lss import BatchedSegments, Segments
    grid, box = (64, 97, 64), (2 * np.pi, np.pi)
    Re, flux, dt = 3000.0, 4.0 / 3.0, 0.004
    lens = [int(x) for x in np.random.default_rng(3).integers(100, 401, size=8)]
    bounds = [0] + list(np.cumsum(lens))
    print("case: 8 uneven segments %s at 64x64x97, 9 tangents each" % lens, flush=True)
    C0 = synthetic_state(*grid, *box, flux, seed=0)
    dirs = []
    for d in range(9):
        if d < 8:
            V = np.stack([synthetic_state(*grid, *box, flux, seed=100 * d + s, unit=True)
                          for s in range(len(lens))])
        else:
            V = np.zeros((len(lens), grid[2], 2, grid[1] - 2, grid[0] // 2), complex)
        dirs.append((V, int(d == 8)))
    bs = BatchedSegments(grid, box, Re, flux, dt, bounds=bounds)
    batched, resid = [], []
    cc.sc_set_graphs(1)
    try:
        bs.setup(C0)
        tuning.autotune(grid, box, Re, flux, dt, bs.starts, cache=tuning.SHAPES_FILE)
        for V, inhomo in dirs:
            V1, eta = bs.forward_all(V, inhomo)
            batched.append((V1, eta))
            resid.append(max(abs(cc.sc_member_ddt_project(s, bs.lens[s], V1[s].copy())) /
                             max(abs(eta[s]), 1e-300) for s in range(bs.n)))
    finally:
        cc.sc_set_graphs(0)
        tuning.reset()
        bs.release()
    print("projection residuals per direction:", " ".join("%.1e" % r for r in resid), flush=True)
    assert max(resid) <= 1e-12
    assert cc.c_init(*grid, *box, Re, flux, dt, bounds[-1]) == 0
    try:
        cc.c_primal(0, C0.copy())
        seg = Segments(grid, bounds)
        for (V, inhomo), (V1, eta) in zip(dirs, batched):
            for s in range(seg.n):
                v, e = seg.forward(s, V[s], inhomo)
                np.testing.assert_array_equal(v, V1[s])
                assert e == eta[s], (s, e, eta[s])
    finally:
        cc.c_destroy()
