"""Write the multi-step parity fixture of the 64x64x97 workload (run here,
on the CPU, with the oracle; the GPU test reads the result).

Workload: Lx = 2 pi, Lz = pi, Re = 3000, dt = 0.004, flux 4/3, the seeded
synthetic state (seed 0, amplitude 0.02), a window of NSTEP steps: the
primal, a homogeneous tangent from the unit direction of seed 3, then the
forced adjoint (strength 1) from the unit direction of seed 4.  Stored: the
sketches (tests/golden/sketch.py) of the primal states at steps 1, NSTEP/2
and NSTEP, the tangent at NSTEP, the final adjoint state, and the norms.

    python tests/golden/make_window_fixtures.py            (writes the fixture)
    python tests/golden/make_window_fixtures.py --floor    (FMA/plain floors)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [HERE, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "shadow-channel_amd")]

import pyorc                                   # noqa: E402
from channel.synthetic import synthetic_state  # noqa: E402
from sketch import sketch                      # noqa: E402

CFG = dict(Nx=64, Ny=97, Nz=64, Lx=2 * np.pi, Lz=np.pi, Re=3000.0, flux=4.0 / 3.0, dt=0.004)
NSTEP = 10
OUT = os.path.join(HERE, "window_64x64x97_10steps.npz")


def run_window():
    """The window's output arrays from the loaded oracle build."""
    c = CFG
    t0 = time.time()
    o = pyorc.Oracle(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"], c["Re"], c["flux"], c["dt"],
                     NSTEP)
    st = lambda seed, **kw: synthetic_state(c["Nx"], c["Ny"], c["Nz"], c["Lx"], c["Lz"],
                                            c["flux"], seed=seed, **kw)
    C = st(0)
    o.primal(0, C)
    print("primal %.0f s" % (time.time() - t0), flush=True)
    IC = st(3, unit=True)
    o.tangent(0, NSTEP, IC, 0)
    print("tangent %.0f s" % (time.time() - t0), flush=True)
    AC = st(4, unit=True)
    o.adjoint(NSTEP, 0, AC, 0, 1.0)
    print("adjoint %.0f s" % (time.time() - t0), flush=True)
    arrays = {"primal_1": o.getsoln(1), "primal_mid": o.getsoln(NSTEP // 2),
              "primal_end": o.getsoln(NSTEP), "tangent_end": IC, "adjoint_end": AC}
    o.destroy()
    return arrays


def main():
    t0 = time.time()
    arrays = run_window()
    if "--floor" in sys.argv:
        # the same window with the oracle built with FMA contraction: how far
        # rounding order alone moves each array (sets the GPU test's bounds)
        pyorc._LIB = None
        os.environ["ORC_LIB"] = os.path.join(ROOT, "oracle", "liborc_fma.so")
        fma = run_window()
        for k in arrays:
            d = np.linalg.norm(arrays[k] - fma[k]) / np.linalg.norm(arrays[k])
            print("floor %s %.2e" % (k, d))
        return
    out = {"nstep": NSTEP}
    for k, a in arrays.items():
        out[k] = sketch(a)
        out[k + "_norm"] = np.linalg.norm(a)
    np.savez(OUT, **out)
    print("wrote %s (%.0f s)" % (OUT, time.time() - t0))


if __name__ == "__main__":
    main()
