"""HIP graphs (sc_set_graphs): a stepping loop recorded once and replayed
gives the same results, bit for bit, as the loop launched kernel by kernel.

Each run signature is called three times with graphs on (the recording
call, then two replays) with other eager work and host synchronisations in
between (a run of another signature, copies of the stored trajectory), and
every call is checked against the eager run of the same input:

* c_primal, c_tangent (with and without the mean-velocity source),
  sc_tangent_sources with the Reynolds-number source, c_adjoint (forced)
  and the members forms, on 16x33x12 (generic-length transforms) and
  64x17x64 (fixed-length transforms, the merged adjoint x-kernel);
* a primal run stopped by the CFL test: the replayed graph's kernels skip
  their work once the flag is set, so the output and the cleared
  trajectory equal the eager run's."""
import numpy as np
import pytest

from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

GRIDS = {
    "16x33x12": dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0,
                     dt=0.005),
    "64x17x64": dict(Nx=64, Ny=17, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0,
                     dt=0.004),
}
N = 4


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.sc_set_graphs(0)
    c_channel.c_destroy()
    c_channel.sc_set_members(1)


def _st(cfg, seed, **kw):
    c = cfg
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=seed, **kw)


def _init(cc, cfg, n=N, members=1):
    c = cfg
    cc.sc_set_members(members)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     n) == 0


def _runs(cc, C0, D0, A0):
    """Primal, three tangents and a forced adjoint: outputs and the stored
    trajectories."""
    out = {}
    C = C0.copy()
    cc.c_primal(0, C)
    out["primal"] = C
    out["MC"] = np.stack([cc.c_getsoln(i) for i in range(N + 1)])
    for name, inhomo in (("tangent", 0), ("tangent_inhomo", 1)):
        I = D0.copy()
        cc.c_tangent(0, N, I, inhomo)
        out[name] = I
    I = D0.copy()
    cc.sc_tangent_sources(1, N, I, 0, 1)
    out["tangent_re"] = I
    out["MIC"] = np.stack([cc.c_getincresoln(i) for i in range(1, N + 1)])
    A = A0.copy()
    cc.c_adjoint(N, 0, A, 0, 1.0)
    out["adjoint"] = A
    return out


@pytest.mark.parametrize("grid", list(GRIDS))
def test_graph_replays_equal_eager_runs(cc, grid):
    cfg = GRIDS[grid]
    print("case: graph replays vs eager runs, %s" % grid, flush=True)
    _init(cc, cfg)
    C0, D0, A0 = _st(cfg, 1, amp=0.1), _st(cfg, 2, unit=True), _st(cfg, 3, unit=True)
    cc.sc_set_graphs(0)
    ref = _runs(cc, C0, D0, A0)
    other = _st(cfg, 4, amp=0.1)
    ref_other = other.copy()
    cc.c_primal(0, ref_other)
    cc.sc_set_graphs(1)
    try:
        for rep in range(3):
            got = _runs(cc, C0, D0, A0)
            for k, v in ref.items():
                np.testing.assert_array_equal(got[k], v, err_msg="%s, call %d" % (k, rep))
            # eager work of another signature between the replays
            cc.sc_set_graphs(0)
            x = other.copy()
            cc.c_primal(0, x)
            np.testing.assert_array_equal(x, ref_other)
            cc.sc_set_graphs(1)
            cc.sc_synchronize()
    finally:
        cc.sc_set_graphs(0)


def test_graph_replays_of_members(cc):
    cfg = GRIDS["16x33x12"]
    nm = 3
    print("case: graph replays of %d members, 16x33x12" % nm, flush=True)
    _init(cc, cfg, members=nm)
    C0 = np.stack([_st(cfg, s, amp=0.1) for s in range(nm)])
    D0 = np.stack([_st(cfg, 10 + s, unit=True) for s in range(nm)])
    A0 = np.stack([_st(cfg, 20 + s, unit=True) for s in range(nm)])

    def runs():
        out = []
        for fn, x, args in ((cc.sc_primal_members, C0, (0,)),
                            (cc.sc_tangent_members, D0, (0, N)),
                            (cc.sc_adjoint_members, A0, (N, 0))):
            d = cc.DeviceArray(np.ascontiguousarray(x))
            try:
                if fn is cc.sc_primal_members:
                    fn(0, d)
                elif fn is cc.sc_tangent_members:
                    fn(0, N, d, 0b001, True)
                else:
                    fn(N, 0, d, 1.0, True)
                out.append(d.download())
            finally:
                d.free()
        return out
    cc.sc_set_graphs(0)
    ref = runs()
    cc.sc_set_graphs(1)
    try:
        for rep in range(3):
            for a, b in zip(runs(), ref):
                np.testing.assert_array_equal(a, b, err_msg="call %d" % rep)
    finally:
        cc.sc_set_graphs(0)


def test_graph_replay_of_a_cfl_stop(cc):
    print("case: graph replay of a CFL stop, 16x33x12 at dt 5", flush=True)
    c = dict(GRIDS["16x33x12"])
    c["dt"] = 5.0
    _init(cc, c, n=3)
    C0 = _st(c, 0, amp=0.1)
    cc.sc_set_graphs(1)
    try:
        for rep in range(2):
            C = C0.copy()
            cc.c_primal(0, C)
            assert cc.last_status() == -1
            np.testing.assert_array_equal(C, C0)
            assert np.all(cc.c_getsoln(1) == 0)
    finally:
        cc.sc_set_graphs(0)
