"""Launch shapes (sc_set_tuning) and the autotuner (channel/tuning.py).

Every shape gives the same results bit for bit: primal, tangent and forced
adjoint runs of 5 members (so that with four members per solve block the
last block holds one) under each members-per-solve-block choice and a few
y-transform block counts, on 16x33x12 and 64x17x64 (dimR 31 and 15; both
take the two-row-slot solve), against the built-in shapes.  The autotuner
returns a choice from its candidates, leaves it set, and applies a cached
choice without timing."""
import json

import numpy as np
import pytest

from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

GRIDS = {
    "16x33x12": (16, 33, 12, np.pi, 0.3 * np.pi, 2000.0, 0.005),
    "64x17x64": (64, 17, 64, 2 * np.pi, np.pi, 3000.0, 0.004),
}


@pytest.fixture(scope="module")
def cc():
    import c_channel
    from channel import tuning
    yield c_channel
    tuning.reset()
    c_channel.sc_set_members(1)
    c_channel.c_destroy()


def _runs(cc, g, nm, N):
    Nx, Ny, Nz, Lx, Lz, Re, dt = g
    st = lambda s, **kw: synthetic_state(Nx, Ny, Nz, Lx, Lz, 4.0 / 3.0, seed=s, **kw)
    out = []
    cc.sc_set_members(nm)
    try:
        assert cc.c_init(Nx, Ny, Nz, Lx, Lz, Re, 4.0 / 3.0, dt, N) == 0
        for fn, seeds, kw in ((cc.sc_primal_members, range(nm), dict(amp=0.1)),
                              (cc.sc_tangent_members, range(10, 10 + nm), dict(unit=True)),
                              (cc.sc_adjoint_members, range(20, 20 + nm), dict(unit=True))):
            d = cc.DeviceArray(np.stack([st(s, **kw) for s in seeds]))
            try:
                if fn is cc.sc_primal_members:
                    fn(0, d)
                elif fn is cc.sc_tangent_members:
                    fn(0, N, d, 0b00101, False)
                else:
                    fn(N, 0, d, 1.0, False)
                out.append(d.download())
            finally:
                d.free()
        cc.c_destroy()
    finally:
        cc.sc_set_members(1)
    return out


@pytest.mark.parametrize("grid", list(GRIDS))
def test_shapes_do_not_change_results(cc, grid):
    from channel import tuning
    g = GRIDS[grid]
    print("case: launch shapes, 5 members, %s" % grid, flush=True)
    tuning.reset()
    ref = _runs(cc, g, 5, 3)
    for name, value in (("solve_members_per_block", 1), ("solve_members_per_block", 2),
                        ("solve_members_per_block", 4), ("rebuild_blocks", 32),
                        ("rebuild_blocks", 128), ("force_blocks", 64)):
        cc.sc_set_tuning(name, value)
        assert cc.sc_get_tuning(name) == value
        got = _runs(cc, g, 5, 3)
        tuning.reset()
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b, err_msg="%s=%d" % (name, value))


def test_autotuner_choice_and_cache(cc, tmp_path):
    from channel import tuning
    Nx, Ny, Nz, Lx, Lz, Re, dt = GRIDS["64x17x64"]
    print("case: autotuner, 4 members, 64x17x64", flush=True)
    states = np.stack([synthetic_state(Nx, Ny, Nz, Lx, Lz, 4.0 / 3.0, seed=s) for s in range(4)])
    cache = str(tmp_path / "shapes.json")
    best = tuning.autotune((Nx, Ny, Nz), (Lx, Lz), Re, 4.0 / 3.0, dt, states, cache=cache)
    for name, value in best.items():
        assert value in tuning.SHAPES[name]
        assert cc.sc_get_tuning(name) == value
    assert len(tuning.autotune.last_timings) >= 8
    known = json.load(open(cache))
    assert list(known.values()) == [best]
    tuning.reset()
    tuning.autotune.last_timings = {}
    again = tuning.autotune((Nx, Ny, Nz), (Lx, Lz), Re, 4.0 / 3.0, dt, states, cache=cache)
    assert again == best and tuning.autotune.last_timings == {}
    assert all(cc.sc_get_tuning(n) == v for n, v in best.items())
    tuning.reset()
    with pytest.raises(AssertionError):
        cc.sc_set_tuning("solve_members_per_block", 3)
