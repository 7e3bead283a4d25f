"""Seeded random sketches of coefficient arrays, for parity fixtures.

A full 64x64x97 state is 6.2 MB; a fixture keeps instead K = 64 projections
s_k = <g_k, x> of the array x (real and imaginary parts as one real vector)
on Gaussian vectors g_k drawn from numpy's PCG64 with seed SEED + k.  For two
arrays x, y, ||S(x) - S(y)|| / ||S(y)|| estimates ||x - y|| / ||y|| within a
factor of about 1 +- 0.35 for K = 64 (chi-square with 64 degrees of freedom),
so a parity bound b on the arrays is checked as 1.5 b on the sketches.
"""
import numpy as np

K = 64
SEED = 20240


def sketch(x):
    v = np.ascontiguousarray(x).view(np.float64).ravel()
    out = np.empty(K)
    for k in range(K):
        out[k] = np.random.default_rng(SEED + k).standard_normal(v.size) @ v
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
