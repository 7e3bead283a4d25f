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


@needs_fma
def test_rounding_floor_at_benchmark_grid():
    """64x64x97 channel (Lx = 2 pi, Lz = pi, Re = 3000, dt = 0.004): one primal
    step and one adjoint step (strength 1).  The wall-normal resolution makes
    the solves far less well conditioned than at Ny = 65: measured floors
    9.7e-11 (step) and 1.8e-9 (adjoint); the GPU tolerances at this grid
    (3e-10, 5e-9) stay above them."""
    c = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0, dt=0.004)

    def run(part):
        def f():
            o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                             c["dt"], 1)
            C = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)
            o.primal(0, C)
            if part == "primal":
                o.destroy()
                return C
            # a tangent first, so that the adjoint's forcing P(MIC[3n]) is active
            o.tangent(0, 1, synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"],
                                            c["flux"], seed=3, unit=True), 0)
            A = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=4,
                                unit=True)
            o.adjoint(1, 0, A, 0, 1.0)
            o.destroy()
            return A
        return f

    step = _both_builds(run("primal"))
    adj = _both_builds(run("adjoint"))
    assert 1e-12 < step < 2e-10
    assert 1e-11 < adj < 3e-9


@needs_fma
def test_rounding_floor_at_largest_grid():
    """128x128x129 channel (Lx = 2 pi, Lz = pi, Re = 5000, dt = 0.002): one
    primal step and one adjoint step (strength 1), about 100 s.  Measured
    floors 9.8e-10 (step) and 1.75e-8 (adjoint); the GPU tolerances at this
    grid (3e-9, 6e-8) stay above them."""
    c = dict(Nx=128, Ny=129, Nz=128, Lx=2 * np.pi, Lz=np.pi, Re=5000.0, flux=4.0 / 3.0, dt=0.002)

    def run():
        o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"],
                         c["dt"], 1)
        C = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=0)
        o.primal(0, C)
        # a tangent first, so that the adjoint's forcing P(MIC[3n]) is active
        o.tangent(0, 1, synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"],
                                        seed=3, unit=True), 0)
        A = synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["flux"], seed=4,
                            unit=True)
        o.adjoint(1, 0, A, 0, 1.0)
        o.destroy()
        return np.concatenate([C.ravel(), A.ravel()])

    outs = []
    for path in (pyorc.lib_path(), FMA_LIB):
        pyorc._LIB = None
        orig = pyorc.lib_path
        pyorc.lib_path = lambda p=path: p
        try:
            outs.append(run())
        finally:
            pyorc.lib_path = orig
            pyorc._LIB = None
    n = outs[0].size // 2
    step = np.linalg.norm(outs[0][:n] - outs[1][:n]) / np.linalg.norm(outs[0][:n])
    adj = np.linalg.norm(outs[0][n:] - outs[1][n:]) / np.linalg.norm(outs[0][n:])
    assert 1e-12 < step < 1.5e-9, step
    assert 1e-11 < adj < 3e-8, adj
