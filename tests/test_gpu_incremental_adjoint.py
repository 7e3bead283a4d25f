"""GPU parity of the incremental adjoint stage (sc_increadjproject, the HIP
kernels of increadjproject0 / increadjproject, channel/increadjproject0.c:
14-357 and channel/increadjproject.c:10-645) against the oracle's
restatement on the same seeded inputs, after the same primal and tangent
runs (whose stored tangent states MIC[count], MIC[count - 1] enter the
stage, each side's own).

Tolerance: relative L2 per output array, a few times the stage's own
rounding floor on these inputs (white noise in every mode, so the stiffest
modes carry weight): the oracle built with FMA contraction differs from the
plain oracle by at most 1.1e-11 (16x33x12), 2.0e-12 (12x25x10), 1.7e-9
(32x65x32) and 1.8e-9 (64x97x64, the benchmark grid), the largest in the
Gauss-point values and wall values of stage 0, on the same inputs, each
after its own primal and tangent runs.
"""
import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu

TOL = {"16x33x12": 1e-10, "12x25x10": 1e-10, "32x65x32": 5e-9, "64x97x64": 5e-9}
GRIDS = {
    "16x33x12": dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0,
                     dt=0.005),
    "12x25x10": dict(Nx=12, Ny=25, Nz=10, Lx=2.0, Lz=1.3, Re=800.0, flux=4.0 / 3.0, dt=0.01),
    "32x65x32": dict(Nx=32, Ny=65, Nz=32, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0,
                     dt=0.005),
    "64x97x64": dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0,
                     dt=0.004),
}
NSTEPS = 2


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.c_destroy()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def setup(cc, name):
    c = GRIDS[name]
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], NSTEPS)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     NSTEPS) == 0
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=3, amp=0.02)
    D0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=4, unit=True)
    Co, Cg = C0.copy(), C0.copy()
    o.primal(0, Co)
    cc.c_primal(0, Cg)
    Io, Ig = D0.copy(), D0.copy()
    o.tangent(0, NSTEPS, Io, 0)
    cc.c_tangent(0, NSTEPS, Ig, 0)
    # the stored tangent states of the two sides differ at the tangent's own
    # rounding floor (1.1e-9 after two steps at 32x65x32); they enter the
    # stage scaled by re dt
    assert rel(Ig, Io) <= 1e-8
    return o


def inputs(o, seed, force):
    rng = np.random.default_rng(seed)
    Nz, Mx = o.cshape[0], o.cshape[3]

    def rc(shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    IAC, IU, LIU = rc(o.cshape), rc(o.ushape), rc(o.ushape)
    W = rc((6, Nz, Mx))
    F = rc((2, 2, Nz, Mx, o.dimR)) if force else None
    return IAC, IU, LIU, W, F


@pytest.mark.parametrize("name", list(GRIDS))
def test_stages_match_the_oracle(cc, name):
    o = setup(cc, name)
    for k in (0, 1, 2):
        for force in (False, True):
            print("case: increadjproject k=%d force=%d %s" % (k, force, name), flush=True)
            count = 5 - k
            IAC, IU, LIU, W, F = inputs(o, 10 * k + force, force)
            ref, IAUo, IAUbo = o.increadjproject(k, count, IAC, IU, LIU, W, F)
            got = IAC.copy()
            IAU, IAUb = cc.sc_increadjproject(k, count, got, IU, None if k == 2 else LIU, W, F)
            tol = TOL[name]
            assert rel(got, ref) <= tol, (k, force, rel(got, ref))
            assert rel(IAU, IAUo) <= tol, (k, force, rel(IAU, IAUo))
            assert rel(IAUb, IAUbo) <= tol, (k, force, rel(IAUb, IAUbo))


def test_three_stages_in_a_row(cc):
    """Stages 0, 1, 2 of one step (count 6, 5, 4), the IAC of each stage the
    next one's input, the products of stage k + 1 as LIU of stage k."""
    o = setup(cc, "16x33x12")
    print("case: increadjproject stages 0..2 chained, 16x33x12", flush=True)
    IAC, _, _, W, F = inputs(o, 99, True)
    rng = np.random.default_rng(5)
    prods = [rng.standard_normal(o.ushape) + 1j * rng.standard_normal(o.ushape) for _ in range(3)]
    ro, rg = IAC.copy(), IAC.copy()
    for k in range(3):
        LIU = prods[k + 1] if k < 2 else None
        ro, IAUo, IAUbo = o.increadjproject(k, 6 - k, ro, prods[k], LIU, W, F)
        IAU, IAUb = cc.sc_increadjproject(k, 6 - k, rg, prods[k], LIU, W, F)
        tol = TOL["16x33x12"]
        assert rel(rg, ro) <= tol
        assert rel(IAU, IAUo) <= tol
        assert rel(IAUb, IAUbo) <= tol


def test_arguments_are_checked(cc):
    o = setup(cc, "12x25x10")
    IAC, IU, LIU, W, _ = inputs(o, 1, False)
    for k, count, liu in ((3, 3, LIU), (0, 0, LIU), (1, 3 * NSTEPS + 1, LIU), (0, 3, None)):
        with pytest.raises(Exception):
            cc.sc_increadjproject(k, count, IAC.copy(), IU, liu, W)
    # count 0 is allowed at the last stage (no MIC[count - 1] is read)
    got = IAC.copy()
    cc.sc_increadjproject(2, 0, got, IU, None, W)
    ref, _, _ = o.increadjproject(2, 0, IAC, IU, None, W)
    assert rel(got, ref) <= TOL["12x25x10"]
    # more than 128 omega_y coefficients (Ny = 133) is refused
    assert cc.c_init(8, 133, 8, 2.0, 1.3, 800.0, 4.0 / 3.0, 0.01, 1) == 0
    nR, q = cc.c_dimR(), cc.c_qpts()
    assert nR > 128
    z = np.zeros((8, 2, nR, 4), complex)
    u = np.zeros((8, 5, q, 4), complex)
    with pytest.raises(Exception):
        cc.sc_increadjproject(2, 0, z, u, None, np.zeros((6, 8, 4), complex))


@pytest.mark.parametrize("name", ["16x33x12", "32x65x32"])
def test_incremental_products_match_the_oracle(cc, name):
    """sc_pass2_incremental (the incremental side of channel/pass2.c) against
    the oracle's restatement on Gauss-point fields rebuilt from seeded
    coefficient arrays; then those products through a stage on both sides."""
    o = setup(cc, name)
    print("case: pass2 incremental products, %s" % name, flush=True)
    Nz = o.cshape[0]

    def field(seed):
        r = np.random.default_rng(seed)
        C = r.standard_normal(o.cshape) + 1j * r.standard_normal(o.cshape)
        C[Nz // 2] = 0
        C[0, :, :, 0] = C[0, :, :, 0].real
        return o.coef_to_gauss(C)

    U, IU, AU, IAU = field(1), field(2), field(3), field(4)
    ref = o.pass2_incremental(U, IU, AU, IAU)
    got = cc.sc_pass2_incremental(U, IU, AU, IAU)
    assert rel(got, ref) <= 1e-12, rel(got, ref)
    assert not np.any(got[Nz // 2])
    IAC, _, _, W, _ = inputs(o, 7, False)
    r0, IAUo, _ = o.increadjproject(2, 3, IAC, ref, None, W)
    g0 = IAC.copy()
    IAUg, _ = cc.sc_increadjproject(2, 3, g0, got, None, W)
    assert rel(g0, r0) <= TOL[name]
    assert rel(IAUg, IAUo) <= TOL[name]


def test_mixed_precision_context_runs_fp64(cc):
    """In a mixed-precision context both entry points still compute in FP64:
    the products and a last stage (which reads no tangent state) match the
    oracle at the FP64 tolerance."""
    c = GRIDS["16x33x12"]
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], NSTEPS)
    cc.sc_set_precision(1)
    try:
        assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                         NSTEPS) == 0
        assert cc.sc_precision() == 1
        print("case: mixed-precision context, 16x33x12", flush=True)
        Nz = o.cshape[0]

        def field(seed):
            r = np.random.default_rng(seed)
            C = r.standard_normal(o.cshape) + 1j * r.standard_normal(o.cshape)
            C[Nz // 2] = 0
            C[0, :, :, 0] = C[0, :, :, 0].real
            return o.coef_to_gauss(C)

        U, IU, AU, IAU = field(11), field(12), field(13), field(14)
        ref = o.pass2_incremental(U, IU, AU, IAU)
        got = cc.sc_pass2_incremental(U, IU, AU, IAU)
        assert rel(got, ref) <= 1e-12, rel(got, ref)
        IAC, _, _, W, F = inputs(o, 21, True)
        r0, IAUo, IAUbo = o.increadjproject(2, 0, IAC, ref, None, W, F)
        g0 = IAC.copy()
        IAUg, IAUbg = cc.sc_increadjproject(2, 0, g0, got, None, W, F)
        tol = TOL["16x33x12"]
        assert rel(g0, r0) <= tol
        assert rel(IAUg, IAUo) <= tol
        assert rel(IAUbg, IAUbo) <= tol
    finally:
        cc.sc_set_precision(0)


def test_device_entries_equal_host_entries(cc):
    """sc_pass2_incremental_dev and sc_increadjproject_dev on resident arrays
    give the host-array entries' results bit for bit, chained: the products
    formed on the device feed the stage without a host copy."""
    o = setup(cc, "16x33x12")
    print("case: device-array entries, 16x33x12", flush=True)
    Nz = o.cshape[0]

    def field(seed):
        r = np.random.default_rng(seed)
        C = r.standard_normal(o.cshape) + 1j * r.standard_normal(o.cshape)
        C[Nz // 2] = 0
        C[0, :, :, 0] = C[0, :, :, 0].real
        return o.coef_to_gauss(C)

    U, IU, AU, IAU = field(31), field(32), field(33), field(34)
    LIU = field(35)
    IAC, _, _, W, F = inputs(o, 36, True)
    IH = cc.sc_pass2_incremental(U, IU, AU, IAU)
    ref_c = IAC.copy()
    ref_u, ref_b = cc.sc_increadjproject(0, 4, ref_c, IH, LIU, W, F)
    dev = {n: cc.DeviceArray(a) for n, a in (("U", U), ("IU", IU), ("AU", AU), ("IAU", IAU),
                                             ("LIU", LIU), ("IAC", IAC), ("W", W), ("F", F),
                                             ("IH", np.zeros_like(U)), ("OU", np.zeros_like(U)),
                                             ("OB", np.zeros((2,) + W.shape[1:], complex)))}
    try:
        cc.sc_pass2_incremental_dev(dev["U"], dev["IU"], dev["AU"], dev["IAU"], dev["IH"])
        np.testing.assert_array_equal(dev["IH"].download(), IH)
        cc.sc_increadjproject_dev(0, 4, dev["IAC"], dev["IH"], dev["LIU"], dev["W"], dev["F"],
                                  dev["OU"], dev["OB"])
        np.testing.assert_array_equal(dev["IAC"].download(), ref_c)
        np.testing.assert_array_equal(dev["OU"].download(), ref_u)
        np.testing.assert_array_equal(dev["OB"].download(), ref_b)
    finally:
        for a in dev.values():
            a.free()
