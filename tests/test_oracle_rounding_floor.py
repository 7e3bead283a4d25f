"""How far rounding order alone moves the discrete solution.

The oracle built twice from the same source, once with fused multiply-add
contraction, gives two legitimate rounding orders of the same algorithm.
Their difference after one step of the 32x32x65 minimal channel is the floor
below which no reimplementation with a different summation or FFT order can
agree with the reference; the GPU parity tolerances are set from it.  The
time-derivative projection (ddt_project, a difference quotient of stored
states) has a floor of its own, larger than that of a step.
"""
import os

import numpy as np
import pytest

import pyorc
from channel.synthetic import synthetic_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FMA_LIB = os.path.join(ROOT, "oracle", "liborc_fma.so")


def _has_fma():
    try:
        return " fma " in open("/proc/cpuinfo").read().replace("\n", " ")
    except OSError:
        return False


def _both_builds(fn):
    out = []
    for path in (pyorc.lib_path(), FMA_LIB):
        pyorc._LIB = None
        orig = pyorc.lib_path
        pyorc.lib_path = lambda p=path: p
        try:
            out.append(fn())
        finally:
            pyorc.lib_path = orig
            pyorc._LIB = None
    return np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[0])


needs_fma = pytest.mark.skipif(not _has_fma() or not os.path.exists(FMA_LIB),
                               reason="needs an FMA-capable host and oracle/liborc_fma.so")


@needs_fma
def test_rounding_floor_of_projection():
    c = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)

    def run():
        o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                         c["dt"], 3)
        o.primal(0, synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"],
                                    seed=0, amp=0.1))
        v = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=12,
                            unit=True)
        o.ddt_project(2, v)
        o.destroy()
        return v

    floor = _both_builds(run)
    # measured 9.2e-11; the GPU projection tolerance (3e-10) must stay above it
    assert 1e-12 < floor < 1.5e-10


@needs_fma
def test_rounding_floor_of_one_step():
    c = dict(Nx=32, Ny=65, Nz=32, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)
    C0 = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)
    out = []
    for path in (pyorc.lib_path(), FMA_LIB):
        pyorc._LIB = None
        orig = pyorc.lib_path
        pyorc.lib_path = lambda p=path: p
        try:
            o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                             c["dt"], 1)
            C = C0.copy()
            o.primal(0, C)
            out.append(C)
        finally:
            pyorc.lib_path = orig
            pyorc._LIB = None
    floor = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[0])
    # measured 4.9e-11 on this container's Xeon; the GPU tolerance (1e-10)
    # must stay above it
    assert 1e-12 < floor < 1e-10


def _one_step_parts(c):
    """One primal step, a one-step tangent, ddt_project of two directions at
    steps 0 and 1, and a forced adjoint step, from the loaded oracle build."""
    st = lambda seed, **kw: synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"],
                                            c["flux"], seed=seed, **kw)
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"], 1)
    C = st(0)
    o.primal(0, C)
    out = {"step": C}
    I = st(3, unit=True)
    o.tangent(0, 1, I, 0)                 # also makes the adjoint's forcing active
    out["tangent"] = I
    for i in (0, 1):
        v = st(10 + i, unit=True)
        o.ddt_project(i, v)
        out["ddt%d" % i] = v
    A = st(4, unit=True)
    o.adjoint(1, 0, A, 0, 1.0)
    out["adjoint"] = A
    o.destroy()
    return out


def _floors(c):
    outs = []
    for path in (pyorc.lib_path(), FMA_LIB):
        pyorc._LIB = None
        orig = pyorc.lib_path
        pyorc.lib_path = lambda p=path: p
        try:
            outs.append(_one_step_parts(c))
        finally:
            pyorc.lib_path = orig
            pyorc._LIB = None
    return {k: np.linalg.norm(outs[0][k] - outs[1][k]) / np.linalg.norm(outs[0][k])
            for k in outs[0]}


@needs_fma
def test_rounding_floor_at_benchmark_grid():
    """64x64x97 channel (Lx = 2 pi, Lz = pi, Re = 3000, dt = 0.004).  The
    wall-normal resolution makes the solves far less well conditioned than at
    Ny = 65: measured floors 9.7e-11 (step), 1.7e-9 (tangent), 7.3e-10 and
    4.1e-9 (ddt_project at steps 0 and 1), 1.8e-9 (forced adjoint); the GPU
    bounds at this grid (tests/test_gpu_parity.py: 3e-10, 5e-9, 2.2e-9,
    1.2e-8, 5.4e-9) stay at 3x them."""
    c = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0, dt=0.004)
    f = _floors(c)
    gpu = {"step": 3e-10, "tangent": 5e-9, "ddt0": 2.2e-9, "ddt1": 1.2e-8, "adjoint": 5.4e-9}
    for k, b in gpu.items():
        assert 1e-13 < f[k] <= b / 2.5, (k, f[k])


@needs_fma
def test_rounding_floor_at_largest_grid():
    """128x128x129 channel (Lx = 2 pi, Lz = pi, Re = 5000, dt = 0.002), about
    two minutes.  Measured floors 9.8e-10 (step), 1.9e-8 (tangent), 8.1e-9
    and 1.2e-9 (ddt_project at steps 0 and 1), 1.8e-8 (forced adjoint); the
    GPU bounds at this grid (3e-9, 6e-8, 2.5e-8, 4e-9, 6e-8) stay at 3x them."""
    c = dict(Nx=128, Ny=129, Nz=128, Lx=2 * np.pi, Lz=np.pi, Re=5000.0, flux=4.0 / 3.0, dt=0.002)
    f = _floors(c)
    gpu = {"step": 3e-9, "tangent": 6e-8, "ddt0": 2.5e-8, "ddt1": 4e-9, "adjoint": 6e-8}
    for k, b in gpu.items():
        assert 1e-13 < f[k] <= b / 2.5, (k, f[k])


@needs_fma
def test_rounding_floor_of_gradient_parts():
    """16x33x12 channel, 3 steps: the tangents with the Reynolds-number
    source (from zero, and with the mean-velocity source from a direction),
    the mean-velocity tangent from zero and the objective adjoint.  The
    Reynolds-number tangent from zero is a small difference of O(1) stage
    terms (its result is 1e-8 of the state), so its floor is high: measured
    8.1e-10, against 9e-13, 4e-13 and 8e-13 for the others; the GPU bounds
    (tests/test_gpu_gradient.py: 2.5e-9, 1e-10, 1e-10, 1e-10) stay at 3x
    the floor or above."""
    c = dict(Nx=16, Ny=33, Nz=12, Lx=np.pi, Lz=0.3 * np.pi, Re=2000.0, flux=4.0 / 3.0, dt=0.005)
    st = lambda seed, **kw: synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"],
                                            c["flux"], seed=seed, **kw)

    def parts():
        n = 3
        o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                         c["dt"], n)
        C0 = st(0, amp=0.1)
        o.primal(0, C0.copy())
        out = {}
        for key, inhomo, re_src, seed in (("re", 0, 1, None), ("re_mean_dir", 1, 1, 2),
                                          ("mean", 1, 0, None)):
            I = np.zeros_like(C0) if seed is None else st(seed, unit=True)
            o.tangent_src(0, n, I, inhomo, re_src)
            out[key] = I
        A = st(4, unit=True)
        A[0, 0, :, 0] += np.linspace(0.1, -0.2, A.shape[2])
        o.adjoint_objective(n, 0, A, 0.7)
        out["objective_adjoint"] = A
        o.destroy()
        return out

    outs = []
    for path in (pyorc.lib_path(), FMA_LIB):
        pyorc._LIB = None
        orig = pyorc.lib_path
        pyorc.lib_path = lambda p=path: p
        try:
            outs.append(parts())
        finally:
            pyorc.lib_path = orig
            pyorc._LIB = None
    gpu = {"re": 2.5e-9, "re_mean_dir": 1e-10, "mean": 1e-10, "objective_adjoint": 1e-10}
    for k, b in gpu.items():
        f = np.linalg.norm(outs[0][k] - outs[1][k]) / np.linalg.norm(outs[0][k])
        assert 1e-14 < f <= b / 2.5, (k, f)
