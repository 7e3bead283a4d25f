"""Pin the two parts of the oracle that the composite runs alone leave open.

1. The forced adjoint (channel/channel.c:622-631, 682-689): at both ends of
   every step n the adjoint subtracts 1/2 strength dt P_n(MIC[3n]), with P_n
   the projection of ddt_project at step n.  The unforced adjoint over steps
   j -> 0 is the transpose of the tangent 0 -> j in the Gauss-weighted
   physical inner product <a, b> (drivers/testAdj.py:55-56), so an adjoint
   from AC = 0 over the tangent's window satisfies the identity that
   drivers/testAdj.py:58-76 prints:
       <AC_0, IC_0> = -1/2 strength dt sum_{n=1}^{N} (<P_n IC_n, IC_n>
                                                      + <P_{n-1} IC_{n-1}, IC_{n-1}>)
   with IC_n the tangent at step n.  A wrong sign, a missing 1/2 or dt, or
   the forcing read from the wrong stored step breaks it.

2. The mean-mode flux correction (channel/project0.c:85-98, 153-168,
   channel/increproject0.c:131-144, 209-214, channel/adjproject0.c:82-95,
   153-167): after the mean solve, alpha += 1/2 e M^{-1} w1 with
   e = target - w1^T alpha and w1_i = sum_j Rw_ij (the flux of basis
   function i).  The flux error after the correction is therefore
       e' = e (1 - 1/2 w1^T M^{-1} w1)
   with M = Rs (primal and tangent: mass only) and M = Rs + re b_k dt Rps
   (adjoint: its left-hand-side matrix, b = {2/3, 1/6, 0}).  Checked on every
   stage of primal, tangent (inhomo 0 and 1) and adjoint runs against M built
   from the pinned Gram bands.

3. Both checks fail on oracles built with a planted mistake (a copy of
   oracle/refcpu.c with one edit, compiled into a temporary directory and
   loaded in a subprocess).
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = dict(Nx=8, Ny=17, Nz=8, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, dt=0.005)
FLUX = 4.0 / 3.0
BB = [2 / 3, 1 / 6, 0.0]


def _oracle(n):
    c = CFG
    return pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], FLUX, c["dt"], n)


def _state(seed, **kw):
    c = CFG
    return synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], FLUX, seed=seed, **kw)


def forced_identity(n=4, strength=1.0, refresh=True):
    """Both sides of the forced-adjoint identity on an n-step window.  With
    refresh, the oracle rebuilds the adjoint velocities after each forcing
    (its test hook), which makes the identity exact; without it the forcing
    reaches the convective term one stage late, as in the reference."""
    o = _oracle(n)
    o.L.orc_set_adjoint_refresh(1 if refresh else 0)
    o.primal(0, _state(0, amp=0.2))
    y, w = o.quad()

    def ip(a, b):
        return float(np.sum(o.spec2phys(a) * o.spec2phys(b) * w[None, None, :, None]))

    IC0 = _state(2, unit=True)
    o.tangent(0, n, IC0.copy(), 0)
    ICs = [o.getincresoln(i) for i in range(n + 1)]
    AC = np.zeros_like(IC0)
    o.adjoint(n, 0, AC, 0, strength)
    left = ip(AC, IC0)
    terms = []
    for i in range(n + 1):
        p = ICs[i].copy()
        o.ddt_project(i, p)
        terms.append(ip(p, ICs[i]))
    right = -0.5 * strength * CFG["dt"] * sum(terms[i] + terms[i - 1] for i in range(1, n + 1))
    o.L.orc_set_adjoint_refresh(0)
    o.destroy()
    return left, right


def flux_factors(o, M_kind, k):
    """1 - 1/2 w1^T M^{-1} w1 for the mean-mode matrix of the stage."""
    nR, q = o.dimR, o.qpts
    Rw = o.array("Rw").reshape(nR, q)
    w1 = Rw.sum(axis=1)

    def dense(band, dl):
        A = np.zeros((nR, nR))
        B = band.reshape(nR, 2 * dl + 1)
        for i in range(nR):
            for m in range(-dl, dl + 1):
                j = i + 2 * m
                if 0 <= j < nR:
                    A[i, j] = B[i, dl + m]
        return A
    Rs, Rps = dense(o.array("Rs"), 2), dense(o.array("Rps"), 2)
    if M_kind == "mass":
        M = Rs
    else:
        M = Rs + o.scalar("re") * BB[k] * CFG["dt"] * Rps
    return 1.0 - 0.5 * w1 @ np.linalg.solve(M, w1)


def flux_records(n=2):
    """Flux-correction records of primal, tangent (inhomo 0, 1) and adjoint
    runs, with the factor each should show."""
    o = _oracle(n)
    o.flux_trace_start()
    o.primal(0, _state(0, amp=0.2))
    for inhomo in (0, 1):
        o.tangent(0, n, _state(1, unit=True), inhomo)
    o.adjoint(n, 0, _state(3, unit=True), 0, 1.0)
    recs = o.flux_trace_stop()
    fac = {}
    for k in range(3):
        fac[(0, k)] = fac[(1, k)] = flux_factors(o, "mass", k)
        fac[(2, k)] = flux_factors(o, "lhs", k)
    o.destroy()
    return recs, fac


def check_flux(recs, fac):
    """Largest deviation of the recorded errors from the closed form, in units
    of the bound 1e-13 max(1, |target|) + 1e-10 |e| (rounding of the flux
    sums), and whether the targets are those of the reference."""
    worst, targets_ok = 0.0, True
    for mode, k, target, before, after in recs:
        mode, k = int(mode), int(k)
        want = before * fac[(mode, k)]
        bound = 1e-13 * max(1.0, abs(target)) + 1e-10 * abs(before)
        worst = max(worst, abs(after - want) / bound)
        if mode == 0:
            targets_ok &= target == FLUX
        elif mode == 2:
            targets_ok &= target == 0.0
    return worst, targets_ok


def test_forced_adjoint_matches_trapezoid_sum_of_projected_tangents():
    left, right = forced_identity()
    assert abs(right) > 1e-3
    assert abs(left - right) <= 1e-10 * abs(right), (left, right)


def test_forced_adjoint_lag_of_the_reference():
    """As in the reference, the forcing enters pass2 one stage late: the
    identity then holds to O(dt) only (here 1.6e-4 relative), and the
    unforced adjoint is the same bits in both modes."""
    left, right = forced_identity(refresh=False)
    dev = abs(left - right) / abs(right)
    assert 1e-7 < dev < 1e-3, dev
    n = 2
    o = _oracle(n)
    o.primal(0, _state(0, amp=0.2))
    o.tangent(0, n, _state(2, unit=True), 0)
    out = []
    for refresh in (0, 1):
        o.L.orc_set_adjoint_refresh(refresh)
        A = _state(3, unit=True)
        o.adjoint(n, 0, A, 0, 0.0)
        out.append(A)
    o.L.orc_set_adjoint_refresh(0)
    o.destroy()
    np.testing.assert_array_equal(out[0], out[1])


def test_forced_adjoint_is_linear_in_strength():
    for refresh in (True, False):
        l1, r1 = forced_identity(n=2, strength=1.0, refresh=refresh)
        l3, r3 = forced_identity(n=2, strength=3.0, refresh=refresh)
        assert abs(l3 - 3 * l1) <= 1e-11 * abs(l3)


def test_flux_correction_closed_form():
    recs, fac = flux_records()
    modes = recs[:, 0].astype(int)
    # 2 primal steps, two 2-step tangents (each re-integrating the base, as
    # the reference's tangent does), a 2-step adjoint: 3 stages each
    assert np.count_nonzero(modes == 0) == 18
    assert np.count_nonzero(modes == 1) == 12
    assert np.count_nonzero(modes == 2) == 6
    # tangent targets: 0 (homogeneous), then 2 (inhomo: d/d meanU of 2 meanU)
    np.testing.assert_array_equal(recs[modes == 1, 2], [0.0] * 6 + [2.0] * 6)
    worst, targets_ok = check_flux(recs, fac)
    assert targets_ok
    assert worst <= 1.0, worst
    # the correction is partial (the R basis, zero at the walls, cannot
    # represent a constant): the factor is small but nonzero, so the errors
    # after the correction are a definite fraction of those before it
    assert all(1e-4 < f < 0.1 for f in fac.values())
    assert np.all(np.abs(recs[:, 3]) > 1e-7)


# ---------------------------------------------------------------- mutants
MUTANTS = {
    # forced adjoint: sign of the first forcing
    "forcing_sign": ("orc_ddt_project(n, tmp);\n        for (long e = 0; e < cs; ++e) {\n"
                     "            G.AC[e].re -= 0.5 * strength * tmp[e].re;\n"
                     "            G.AC[e].im -= 0.5 * strength * tmp[e].im;",
                     "orc_ddt_project(n, tmp);\n        for (long e = 0; e < cs; ++e) {\n"
                     "            G.AC[e].re += 0.5 * strength * tmp[e].re;\n"
                     "            G.AC[e].im += 0.5 * strength * tmp[e].im;"),
    # forced adjoint: the 1/2 of the trapezoid at the step's start
    "forcing_half": ("orc_ddt_project(n, tmp);\n        for (long e = 0; e < cs; ++e) {\n"
                     "            G.AC[e].re -= 0.5 * strength * tmp[e].re;\n"
                     "            G.AC[e].im -= 0.5 * strength * tmp[e].im;",
                     "orc_ddt_project(n, tmp);\n        for (long e = 0; e < cs; ++e) {\n"
                     "            G.AC[e].re -= strength * tmp[e].re;\n"
                     "            G.AC[e].im -= strength * tmp[e].im;"),
    # forced adjoint: strength not scaled by dt
    "forcing_dt": ("strength *= G.dt;", "strength *= 1.0;"),
    # forced adjoint: MIC of the previous step at the step's start
    "forcing_index": ("memcpy(tmp, G.MIC + cs * (3L * n), sizeof(cx) * cs);\n"
                      "        orc_ddt_project(n, tmp);",
                      "memcpy(tmp, G.MIC + cs * (3L * (n - 1)), sizeof(cx) * cs);\n"
                      "        orc_ddt_project(n, tmp);"),
    # flux correction without its 1/2
    "flux_half": ("al[j * s].re = al[j * s].re + flux_t * add[j] * 0.5;",
                  "al[j * s].re = al[j * s].re + flux_t * add[j];"),
    # flux correction with the error's sign flipped
    "flux_sign": ("flux_t = target - flux_t;\n    const double err_before",
                  "flux_t = flux_t - target;\n    const double err_before"),
    # adjoint flux correction through the mass matrix instead of its LHS
    "flux_adjoint_matrix": ("if (adj) op_mean(M, re * b[k] * dt);",
                            "if (0) op_mean(M, re * b[k] * dt);"),
}

_CHILD = r"""
import json, sys
sys.path[:0] = {paths!r}
import test_oracle_adjoint_forcing as t
left, right = t.forced_identity()
recs, fac = t.flux_records()
worst, ok = t.check_flux(recs, fac)
print(json.dumps(dict(forced=abs(left - right) <= 1e-10 * abs(right),
                      flux=bool(worst <= 1.0 and ok))))
"""


def _run_mutant(src, name, tmp):
    old, new = MUTANTS[name]
    assert src.count(old) == 1, name
    d = os.path.join(tmp, name)
    os.makedirs(d)
    c = os.path.join(d, "refcpu.c")
    with open(c, "w") as f:
        f.write(src.replace(old, new))
    with open(os.path.join(d, "refcpu.h"), "w") as f:
        f.write(open(os.path.join(ROOT, "oracle", "refcpu.h")).read())
    so = os.path.join(d, "liborc.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=gnu99", "-shared", "-o", so, c, "-lm"],
                   check=True)
    paths = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"),
             os.path.join(ROOT, "shadow-channel_amd")]
    env = dict(os.environ, ORC_LIB=so)
    out = subprocess.run([sys.executable, "-c", _CHILD.format(paths=paths)], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_pins_fail_on_planted_mistakes():
    src = open(os.path.join(ROOT, "oracle", "refcpu.c")).read()
    with tempfile.TemporaryDirectory() as tmp:
        # the unmodified sources pass both checks through the same harness
        MUTANTS["none"] = ("strength *= G.dt;", "strength *= G.dt;")
        try:
            base = _run_mutant(src, "none", tmp)
        finally:
            del MUTANTS["none"]
        assert base == dict(forced=True, flux=True)
        results = {name: _run_mutant(src, name, tmp) for name in MUTANTS}
    for name, r in results.items():
        caught = (not r["forced"]) if name.startswith("forcing") else (not r["flux"])
        assert caught, (name, r)
