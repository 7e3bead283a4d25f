"""Wall-shear gradients of many trajectories of their own lengths, advanced
together as members (channel.gradient.WallShearMembers: sc_primal_members,
sc_tangent_members with the mean-velocity source, and the objective adjoint
sc_adjoint_objective_members with member windows).

* 16x33x12: three members of 6, 4 and 10 steps against the single
  trajectory's checkpointed gradient (WallShearWindow) of each: J, dJ/dmeanU
  and dJ/dC0, with windows of 2 steps (every length a multiple) and of 4
  steps (members ending inside a window).  The runs differ only in where
  the weight 1/L is applied (inside the adjoint or after it): 1e-12.
* The member windows of the objective adjoint: a member with window 0 is
  zero throughout.
* 128x128x129 in mixed precision against FP64 (Lx 2 pi, Lz pi, Re 5000,
  dt 0.002): 16 seeded trajectories of their own lengths (multiples of 25,
  drawn once), checkpoints every 25 steps, each followed by its adjoint
  sweep: the states at steps 1 and 100, each member's time-averaged mean
  velocity profile, J, dJ/dmeanU and dJ/dC0.  The lengths are 50 to 200
  steps (every member runs the longest's windows, so this takes about a
  minute); with SC_FULL_TRAJECTORIES=1 they are 200 to 800 steps.  The
  state bounds are 3x the single-trajectory measurements at this grid
  (tests/test_gpu_mixed.py), the gradients are held to 1e-4."""
import os

import numpy as np
import pytest

from channel.synthetic import synthetic_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cc():
    import c_channel
    yield c_channel
    c_channel.c_destroy()
    c_channel.sc_set_members(1)
    c_channel.sc_set_precision(0)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


SMALL = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, dt=0.005, meanU=2.0 / 3.0,
             Re=2000.0)


@pytest.mark.parametrize("every", [2, 4])
def test_members_gradient_equals_single_windows(cc, every):
    from channel.gradient import WallShearMembers, WallShearWindow
    c = SMALL
    lengths = [6, 4, 10]
    print("case: member wall-shear gradients, lengths %s, windows of %d" % (lengths, every),
          flush=True)
    C0 = [synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], 2 * c["meanU"],
                          seed=s, amp=0.1) for s in range(len(lengths))]
    ws = WallShearMembers(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["dt"], lengths, every)
    try:
        J, cps, dJ_u, _ = ws.forward(np.stack(C0), c["meanU"], c["Re"])
        g = ws.state_gradient(cps, c["meanU"], c["Re"])
    finally:
        ws.release()
    for m, n in enumerate(lengths):
        win = WallShearWindow(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["dt"], n, every=2)
        try:
            J1, cps1, du1, _ = win.forward(C0[m], c["meanU"], c["Re"], params=True)
            g1 = win.state_gradient(cps1, c["meanU"], c["Re"])
        finally:
            win.release()
        e = (abs(J[m] - J1) / abs(J1), abs(dJ_u[m] - du1) / abs(du1), rel(g[m], g1))
        print("member %d (%d steps): J %.1e dJ/dmeanU %.1e dJ/dC0 %.1e" % ((m, n) + e), flush=True)
        assert max(e) <= 1e-12, (m, e)


def test_objective_adjoint_member_windows(cc):
    c = SMALL
    n = 4
    print("case: objective adjoint with member windows 4, 2, 0", flush=True)
    cc.sc_set_members(3)
    assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], 2 * c["meanU"],
                     c["dt"], n) == 0
    C0 = np.stack([synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], 2 * c["meanU"],
                                   seed=s, amp=0.1) for s in range(3)])
    A0 = np.stack([synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], 2 * c["meanU"],
                                   seed=10 + s, unit=True) for s in range(3)])
    d = cc.DeviceArray(C0)
    cc.sc_primal_members(0, d)
    d.free()
    cc.sc_set_member_windows([4, 2, 0])
    a = cc.DeviceArray(A0)
    try:
        cc.sc_adjoint_objective_members(n, 0, a, 0.25, False)
        Aw = a.download()
    finally:
        a.free()
        cc.sc_set_member_windows(None)
    assert not np.any(Aw[2])
    for m, start in ((0, 4), (1, 2)):
        # the same member's trajectory as member 0 of a single run
        cc.sc_set_members(1)
        assert cc.c_init(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], 2 * c["meanU"],
                         c["dt"], n) == 0
        C = C0[m].copy()
        cc.c_primal(0, C)
        A = A0[m].copy()
        cc.sc_adjoint_objective(start, 0, A, 0.25)
        np.testing.assert_array_equal(Aw[m], A)
    cc.c_destroy()
    cc.sc_set_members(1)


def test_config4_mixed_trajectories(cc):
    from channel.gradient import WallShearMembers
    Nx, Ny, Nz, Lx, Lz, Re, dt = 128, 129, 128, 2 * np.pi, np.pi, 5000.0, 0.002
    meanU, every = 2.0 / 3.0, 25
    full = os.environ.get("SC_FULL_TRAJECTORIES") == "1"
    lo, hi = (8, 33) if full else (2, 9)
    lengths = [int(x) for x in np.random.default_rng(4).integers(lo, hi, size=16) * 25]
    print("case: 16 trajectories at 128x128x129, lengths %s, mixed against FP64" % lengths,
          flush=True)
    C0 = np.stack([synthetic_state(Nx, Ny, Nz, Lx, Lz, 2 * meanU, seed=s)
                   for s in range(len(lengths))])
    out = {}
    for mixed in (0, 1):
        cc.sc_set_precision(mixed)
        ws = WallShearMembers(Nx, Ny, Nz, Lx, Lz, dt, lengths, every)
        states = {}

        def grab(w):
            if w == 0:
                states[1] = np.stack([cc.sc_member_state(0, m, 1) for m in range(ws.n)])
            if w * every == 100:
                states[100] = np.stack([cc.sc_member_state(0, m, 0) for m in range(ws.n)])

        try:
            J, cps, dJ_u, prof = ws.forward(C0, meanU, Re, after_window=grab)
            g = ws.state_gradient(cps, meanU, Re)
        finally:
            ws.release()
            cc.sc_set_precision(0)
        out[mixed] = dict(s1=states[1], s100=states[100], prof=prof, J=J, dJ_u=dJ_u, g=g)
    r, x = out[0], out[1]
    errs = {
        "state@1": max(rel(x["s1"][m], r["s1"][m]) for m in range(16)),
        "state@100": max(rel(x["s100"][m], r["s100"][m]) for m in range(16)),
        "profile": max(rel(x["prof"][m], r["prof"][m]) for m in range(16)),
        "J": float(np.max(np.abs(x["J"] - r["J"]) / np.abs(r["J"]))),
        "dJ/dmeanU": float(np.max(np.abs(x["dJ_u"] - r["dJ_u"]) / np.abs(r["dJ_u"]))),
        "dJ/dC0": max(rel(x["g"][m], r["g"][m]) for m in range(16)),
    }
    print(" ".join("%s %.2e" % kv for kv in errs.items()), flush=True)
    # measured 6.1e-7, 6.2e-7, 1.0e-11 (4.3e-11 at full length), 2.9e-10,
    # 2.2e-11, 1.8e-7 (J, dJ/dmeanU, dJ/dC0; 4.9e-10, 4.0e-11, 2.7e-7 at full
    # length)
    bounds = {"state@1": 2e-6, "state@100": 2e-6, "profile": 1.5e-10, "J": 1e-4,
              "dJ/dmeanU": 1e-4, "dJ/dC0": 1e-4}
    for k, b in bounds.items():
        assert errs[k] <= b, (k, errs[k])
