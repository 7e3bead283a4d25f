"""The symmetric Krylov solvers of the least-squares-shadowing driver
(channel/lss.py: minres, cg; the interface of drivers/pariter.py:9-125),
on the CPU: solutions of small dense systems against numpy's direct solve,
MINRES on symmetric indefinite matrices (where CG does not apply) with a
monotone residual, a user inner product (a weighted one, with the
operator symmetric in it), the iteration count bound of exact arithmetic
(n steps for n distinct eigenvalues), and the stopping flags."""
import numpy as np

from channel.lss import cg, minres


def _sym(n, seed, indefinite):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = rng.uniform(0.5, 3.0, n)
    if indefinite:
        ev[::3] *= -1.0
    return (Q * ev) @ Q.T, ev


def test_minres_solves_symmetric_indefinite_systems():
    for seed in range(3):
        A, ev = _sym(40, seed, True)
        b = np.random.default_rng(10 + seed).standard_normal(40)
        res = []
        x, (flag, it) = minres(lambda v: A @ v, b, tol=1e-12, maxiter=200,
                               callback=lambda x: res.append(np.linalg.norm(b - A @ x)))
        assert flag == 0 and it <= 60         # n + 1 in exact arithmetic
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-9)
        assert all(r1 <= r0 * (1 + 1e-10) for r0, r1 in zip(res, res[1:]))


def test_minres_with_a_weighted_inner_product():
    # A symmetric in <x, y>_W = x^T W y: A = S W with S symmetric
    rng = np.random.default_rng(5)
    n = 30
    W = np.diag(rng.uniform(0.5, 2.0, n))
    S, _ = _sym(n, 6, True)
    A = S @ W
    dot = lambda x, y: float(x @ W @ y)
    b = rng.standard_normal(n)
    x, (flag, it) = minres(lambda v: A @ v, b, dot=dot, tol=1e-12, maxiter=200)
    assert flag == 0
    np.testing.assert_allclose(A @ x, b, rtol=0, atol=1e-9)


def test_minres_few_distinct_eigenvalues():
    # three distinct eigenvalues: the Krylov space is exhausted in 3 steps
    rng = np.random.default_rng(2)
    Q, _ = np.linalg.qr(rng.standard_normal((25, 25)))
    ev = np.repeat([-2.0, 0.5, 3.0], [8, 8, 9])
    A = (Q * ev) @ Q.T
    b = rng.standard_normal(25)
    x, (flag, it) = minres(lambda v: A @ v, b, tol=1e-10, maxiter=50)
    assert flag == 0 and it <= 3
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-9)


def test_cg_and_stopping_flags():
    A, _ = _sym(30, 7, False)
    b = np.random.default_rng(8).standard_normal(30)
    x, (flag, it) = cg(lambda v: A @ v, b, tol=1e-12, maxiter=200)
    assert flag == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-9)
    x, (flag, it) = cg(lambda v: A @ v, b, tol=1e-14, maxiter=2)
    assert (flag, it) == (-1, 2)
    A2, _ = _sym(30, 9, True)
    x, (flag, it) = minres(lambda v: A2 @ v, b, tol=1e-14, maxiter=2)
    assert (flag, it) == (1, 2)
    x, (flag, it) = minres(lambda v: A2 @ v, np.zeros(30))
    assert (flag, it) == (0, 0) and not np.any(x)
