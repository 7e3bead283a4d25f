"""Least-squares shadowing over time segments: the KKT operator and the
symmetric Krylov solvers the drivers use (drivers/wrapper.py:6-54,
drivers/pariter.py:1-125).

The window of the stored trajectory is cut at the step indices `bounds`
(bounds[0] = 0 < ... < bounds[n] = the last step).  For segment i:

    forward(i, v0, inhomo): project v0 at bounds[i] (ddt_project), run the
        tangent to bounds[i+1], project at bounds[i+1]; returns the end state
        and the projection coefficient eta;
    backward(i, w0, strength): project w0 at bounds[i+1], run the forced
        adjoint back to bounds[i] (its forcing reads the tangent states the
        forward of the same segment stored), project at bounds[i];

and the KKT residual of the stacked unknowns x = (v_0 .. v_{n-1},
w_0 .. w_{n-2}) (complex states as interleaved real and imaginary parts,
the layout of drivers/wrapper.py:34-40) is

    R_w[i] = v[i+1] - forward(i, v[i])                       (i < n-1)
    R_v[i] = w[i-1] - backward(i, w[i])   (w[-1] = 0, w[n-1] = 0)

with zeta[i] = -eta_i kept for the gradient (channel.uavgGrad).

Two implementations of the same operator:

* Segments: the reference's loop, one segment after another, on the single
  trajectory of the current run (any bounds; c_tangent, c_adjoint and
  c_ddt_project).
* BatchedSegments: the segments (of equal or uneven length) as members of
  one batched run.  Member s recomputes its segment of the base trajectory
  from the state at step bounds[s] (bit for bit the stored one: the same
  kernels on the same state), so every segment's tangent, forced adjoint
  and projections advance in the same kernel launches.  Every member's
  window is as long as the longest segment; a shorter segment's tangent
  ends inside it and its adjoint starts there (sc_set_member_windows).
  The projection at a segment's start is the one at the previous
  segment's end (the backward difference at step bounds[s] uses steps
  bounds[s] - 1 and bounds[s], both in member s-1's window); segment 0
  starts with the forward difference at step 0.  The operator is the
  reference's bit for bit: each member's time derivative at its window's
  first step is the backward difference from the state one step before
  the window (sc_set_member_previous), as on the long trajectory.

minres (Paige and Saunders' minimal-residual method) and cg take the
operator as a callable and an inner product, like drivers/pariter.py:63-125.
"""
import math

import numpy as np

import c_channel


def _as_states(x, n, shape):
    """Interleaved real vector -> complex array of n states."""
    z = np.asarray(x, dtype=np.float64)
    c = z[0::2] + 1j * z[1::2]
    return c.reshape((n,) + shape)


class Segments:
    """The reference's segment loop over the current run's trajectory
    (channel must hold a primal run covering bounds[-1] steps)."""

    def __init__(self, grid, bounds):
        self.Nx, self.Ny, self.Nz = (int(g) for g in grid)
        self.nb = [int(b) for b in bounds]
        self.n = len(self.nb) - 1
        self.shape = (self.Nz, 2, self.Ny - 2, self.Nx // 2)
        self.m = int(np.prod(self.shape))
        self.zeta = np.zeros(self.n)

    def proj(self, i_step, v):
        return c_channel.c_ddt_project(int(i_step), v)

    def forward(self, i, v0, inhomo):
        v1 = np.ascontiguousarray(v0, dtype=np.complex128).copy()
        self.proj(self.nb[i], v1)
        c_channel.c_tangent(self.nb[i], self.nb[i + 1], v1, int(inhomo))
        eta = self.proj(self.nb[i + 1], v1)
        return v1, eta

    def backward(self, i, w0, strength):
        w1 = np.ascontiguousarray(w0, dtype=np.complex128).copy()
        self.proj(self.nb[i + 1], w1)
        c_channel.c_adjoint(self.nb[i + 1], self.nb[i], w1, 0, float(strength))
        self.proj(self.nb[i], w1)
        return w1

    def matvec(self, x, inhomo=0):
        n = self.n
        assert np.size(x) == 2 * self.m * (2 * n - 1)
        v = _as_states(np.ravel(x)[:2 * self.m * n], n, self.shape)
        w = np.concatenate([_as_states(np.ravel(x)[2 * self.m * n:], n - 1, self.shape),
                            np.zeros((1,) + self.shape, complex)])
        R_v = np.zeros((n,) + self.shape, complex)
        R_w = np.zeros((n - 1,) + self.shape, complex)
        for i in range(n):
            vip, eta = self.forward(i, v[i], inhomo)
            self.zeta[i] = -eta
            wim = self.backward(i, w[i], 1.0)
            if i < n - 1:
                R_w[i] = v[i + 1] - vip
            R_v[i] = (w[i - 1] if i > 0 else 0.0) - wim
        return np.concatenate([R_v.ravel(), R_w.ravel()]).view(np.float64)


class BatchedSegments(Segments):
    """The segments advanced together as members.

    BatchedSegments(grid, box, Re, flux, dt, n_seg, seg_steps) cuts n_seg
    segments of seg_steps steps; bounds=[0, b_1, ..., b_n] gives segments of
    any lengths instead.  setup(C0) runs the base trajectory of b_n steps
    once, keeps the states at the segment starts, and makes a member context
    of n members whose windows are the segments (c_init with nsteps = the
    longest segment)."""

    def __init__(self, grid, box, Re, flux, dt, n_seg=None, seg_steps=None, mixed=False,
                 bounds=None):
        if bounds is None:
            bounds = [int(seg_steps) * s for s in range(int(n_seg) + 1)]
        Segments.__init__(self, grid, bounds)
        self.box = tuple(float(b) for b in box)
        self.Re, self.flux, self.dt = float(Re), float(flux), float(dt)
        self.lens = [self.nb[i + 1] - self.nb[i] for i in range(self.n)]
        if self.nb[0] != 0 or min(self.lens) < 1:
            raise ValueError("bounds must start at 0 and increase")
        self.L = max(self.lens)
        self.uneven = min(self.lens) != self.L
        self.mixed = int(bool(mixed))
        self.starts = None

    def setup(self, C0):
        """Base trajectory from C0, then one member per segment."""
        n, L = self.n, self.L
        grid = (self.Nx, self.Ny, self.Nz)
        c_channel.sc_set_precision(self.mixed)
        c_channel.sc_set_members(1)
        assert c_channel.c_init(*grid, *self.box, self.Re, self.flux, self.dt, self.nb[-1]) == 0
        C = np.ascontiguousarray(C0, dtype=np.complex128).copy()
        c_channel.c_primal(0, C)
        self.starts = np.stack([c_channel.c_getsoln(self.nb[s]).copy() for s in range(n)])
        # the states one step before each window (segment 0 has none)
        prev = np.stack([np.zeros_like(self.starts[0])] +
                        [c_channel.c_getsoln(self.nb[s] - 1).copy() for s in range(1, n)])
        c_channel.c_destroy()
        c_channel.sc_set_members(n)
        assert c_channel.c_init(*grid, *self.box, self.Re, self.flux, self.dt, L) == 0
        d = c_channel.DeviceArray(self.starts)
        try:
            c_channel.sc_primal_members(0, d)
        finally:
            d.free()
        # the members' step-0 time derivatives as on the long trajectory:
        # the backward difference (the adjoint's last forcing uses it)
        c_channel.sc_set_member_previous(prev)
        # segments shorter than the window end inside it
        c_channel.sc_set_member_windows(self.lens if self.uneven else None)

    def release(self):
        c_channel.c_destroy()
        c_channel.sc_set_members(1)
        c_channel.sc_set_precision(0)

    def _proj_start(self, s, v):
        # step bounds[s]: member 0's step 0, or the end of member s - 1's segment
        if s == 0:
            return c_channel.sc_member_ddt_project(0, 0, v)
        return c_channel.sc_member_ddt_project(s - 1, self.lens[s - 1], v)

    def _proj_end(self, s, v):
        return c_channel.sc_member_ddt_project(s, self.lens[s], v)

    def forward_all(self, V, inhomo):
        """forward(i, V[i]) of every segment: (end states, etas)."""
        V1 = np.ascontiguousarray(V, dtype=np.complex128).copy()
        for s in range(self.n):
            self._proj_start(s, V1[s])
        mask = (1 << self.n) - 1 if inhomo else 0
        d = c_channel.DeviceArray(V1)
        try:
            c_channel.sc_tangent_members(0, self.L, d, mask, False)
            if self.uneven:
                # each member's state at the end of its own segment
                V1 = np.stack([c_channel.sc_member_state(1, s, self.lens[s])
                               for s in range(self.n)])
            else:
                V1 = d.download()
        finally:
            d.free()
        eta = np.array([self._proj_end(s, V1[s]) for s in range(self.n)])
        return V1, eta

    def backward_all(self, Wt, strength):
        """backward(i, Wt[i]) of every segment (after forward_all)."""
        W1 = np.ascontiguousarray(Wt, dtype=np.complex128).copy()
        for s in range(self.n):
            self._proj_end(s, W1[s])
        d = c_channel.DeviceArray(W1)
        try:
            c_channel.sc_adjoint_members(self.L, 0, d, float(strength), False)
            W1 = d.download()
        finally:
            d.free()
        for s in range(self.n):
            self._proj_start(s, W1[s])
        return W1

    def matvec(self, x, inhomo=0):
        n = self.n
        assert np.size(x) == 2 * self.m * (2 * n - 1)
        v = _as_states(np.ravel(x)[:2 * self.m * n], n, self.shape)
        w = np.concatenate([_as_states(np.ravel(x)[2 * self.m * n:], n - 1, self.shape),
                            np.zeros((1,) + self.shape, complex)])
        vip, eta = self.forward_all(v, inhomo)
        self.zeta[:] = -eta
        wim = self.backward_all(w, 1.0)
        R_w = v[1:] - vip[:-1]
        R_v = -wim
        R_v[1:] += w[:-1]
        return np.concatenate([R_v.ravel(), R_w.ravel()]).view(np.float64)


# ---------------------------------------------------------------- solvers

def minres(A, b, x0=None, dot=np.dot, tol=1e-5, maxiter=None, callback=None):
    """Minimal residual method for a symmetric operator A (a callable) with
    the inner product `dot` (Paige and Saunders, SIAM J. Numer. Anal. 12,
    1975): Lanczos vectors, the tridiagonal QR factorisation by Givens
    rotations updated one column at a time, and the solution updated along
    the directions of R^-1 V.  Returns (x, (0, iterations)) when the
    residual norm drops below tol times the initial one, (1, maxiter) when
    it does not."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    if maxiter is None:
        maxiter = 10 * b.size
    q = b - A(x)
    beta = math.sqrt(dot(q, q))
    r0 = beta
    if r0 == 0.0:
        return x, (0, 0)
    q_prev = np.zeros_like(b)
    d_prev = np.zeros_like(b)     # search directions of the two previous steps
    d_prev2 = np.zeros_like(b)
    cs, sn = 1.0, 0.0             # the last rotation
    cs_prev, sn_prev = 1.0, 0.0   # the one before
    phi = beta                    # residual norm carried by the rotations
    for k in range(maxiter):
        q = q / beta
        Aq = A(q)
        alpha = dot(q, Aq)
        q_next = Aq - alpha * q - beta * q_prev
        beta_next = math.sqrt(dot(q_next, q_next))
        # column k of the tridiagonal matrix through the two earlier rotations
        e_far = sn_prev * beta                    # entry two rows up
        t = cs_prev * beta
        e_up = cs * t + sn * alpha                # entry one row up
        diag = -sn * t + cs * alpha
        # the new rotation zeroes beta_next below the diagonal
        rho = math.hypot(diag, beta_next)
        cs_new, sn_new = diag / rho, beta_next / rho
        d = (q - e_up * d_prev - e_far * d_prev2) / rho
        x = x + (cs_new * phi) * d
        phi = -sn_new * phi
        d_prev2, d_prev = d_prev, d
        cs_prev, sn_prev, cs, sn = cs, sn, cs_new, sn_new
        q_prev, q, beta = q, q_next, beta_next
        if callback is not None:
            callback(x)
        if abs(phi) / r0 < tol or beta == 0.0:
            return x, (0, k + 1)
    return x, (1, maxiter)


def cg(A, b, x0=None, dot=np.dot, tol=1e-5, maxiter=None, callback=None):
    """Conjugate gradients for a symmetric positive definite operator with
    the inner product `dot`; the Polak-Ribiere form of beta (which allows a
    preconditioner that changes between iterations, drivers/pariter.py:9-59).
    Returns (x, (0, iterations)) or (x, (-1, maxiter))."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    if maxiter is None:
        maxiter = 10 * b.size
    r = b - A(x)
    p = r.copy()
    rr0 = dot(r, r)
    if rr0 == 0.0:
        return x, (0, 0)
    rz = rr0
    for k in range(maxiter):
        Ap = A(p)
        a = rz / dot(p, Ap)
        x = x + a * p
        r_new = r - a * Ap
        rr = dot(r_new, r_new)
        if callback is not None:
            callback(x)
        if rr / rr0 < tol * tol:
            return x, (0, k + 1)
        beta = dot(r_new, r_new - r) / rz
        rz = rr
        r = r_new
        p = r + beta * p
    return x, (-1, maxiter)
