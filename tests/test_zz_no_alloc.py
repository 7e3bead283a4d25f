"""sc_ddt_project_noalloc (no device allocation; partial sums in the
library's fixed device scratch array) against c_ddt_project, through the C
ABI: two calls in a row on a 16x33x12 run, then a smaller 8x17x8 run, each
result compared with the allocating entry on the same vector (the same
kernels in the same order: equal bits), the coefficient as well; invalid
calls are refused."""
import numpy as np
import pytest

from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

LARGE = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)
SMALLER = dict(Nx=8, Ny=17, Nz=8, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)


def _run(cc, cfg, n, i_steps):
    c = cfg
    print("case: no-alloc ddt_project, %dx%dx%d" % (c["Nx"], c["Ny"], c["Nz"]), flush=True)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     n) == 0
    cc.c_primal(0, synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"],
                                   seed=0, amp=0.1))
    for k, i in enumerate(i_steps):
        v = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"],
                            seed=40 + k, unit=True)
        vh = v.copy()
        r_ref = cc.c_ddt_project(i, vh)
        vd = cc.DeviceArray(v)
        rd = cc.DeviceArray(np.zeros(2, complex))       # 4 doubles: num, den, num/den
        try:
            cc.sc_ddt_project_noalloc(i, vd, rd)
            cc.sc_synchronize()
            got = vd.download()
            r = rd.download().view(np.float64)
        finally:
            vd.free()
            rd.free()
        assert np.array_equal(got, vh), (i, np.abs(got - vh).max())
        assert r[2] == r_ref and r[0] / r[1] == r[2]


def test_noalloc_ddt_project_matches_allocating_entry():
    import c_channel as cc
    try:
        _run(cc, LARGE, 3, (2, 0))
        _run(cc, SMALLER, 2, (1,))
    finally:
        cc.c_destroy()


def test_noalloc_refuses_invalid_calls():
    import c_channel as cc
    c = SMALLER
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     1) == 0
    try:
        with pytest.raises(AssertionError):
            cc.sc_ddt_project_noalloc(2, 1, 1)          # step beyond the window
        with pytest.raises(AssertionError):
            cc.sc_ddt_project_noalloc(0, 0, 0)          # null arrays
    finally:
        cc.c_destroy()
