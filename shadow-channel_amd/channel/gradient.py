"""Time-averaged wall shear and its gradient, with checkpointed trajectories.

    J(C0, meanU, Re) = (1/N) sum_{n=0}^{N-1} tau(C_n),
    tau(C) = d<u>/dy at the lower wall = sum_j R_j'(-1) Re alpha_j(0,0)

(the Uxb(0,0) of channel/project0.c:227-234), over an N-step window from the
state C0.  The library keeps a window of `every` steps (its stored
trajectory MC); the forward pass runs the window `every` steps at a time
and keeps only the state at each window's start (a checkpoint, in host
memory), and the backward pass recomputes each window from its checkpoint,
latest first, before running the objective adjoint over it:

* dJ/dC0: the adjoint of J (sc_adjoint_objective: exact transpose of the
  tangent), carried back through the windows, then mapped to the gradient
  in the real and imaginary parts of the coefficients (sc_gradient_euclid);
* dJ/dmeanU and dJ/dRe: tangents from zero with the mean-velocity and the
  Reynolds-number sources (sc_tangent_sources), run in the forward pass
  along each window and carried across windows (two parameters: forward
  mode costs two tangents, no more than an adjoint).

WallShear is the torch.autograd.Function over these (complex128 C0; meanU
and Re as float64 scalar tensors); torch is used in this layer only.

WallShearMembers runs many trajectories of their own lengths together, every
member in each kernel launch (sc_*_members): J_m over member m's L_m steps,
dJ_m/dC0_m by the objective adjoint with member windows (a member whose
trajectory ends inside a window starts its adjoint there), dJ_m/dmeanU by
tangents with the mean-velocity source.  Checkpoints stay on the device.
"""
import numpy as np

import c_channel


class WallShearWindow:
    """One grid and time step; windows of `every` steps."""

    def __init__(self, Nx, Ny, Nz, Lx, Lz, dt, n_steps, every=50):
        if n_steps % every:
            raise ValueError("n_steps must be a multiple of the checkpoint interval")
        self.grid = (int(Nx), int(Ny), int(Nz))
        self.box = (float(Lx), float(Lz))
        self.dt = float(dt)
        self.n_steps, self.every = int(n_steps), int(every)
        self._key = None

    def _init(self, meanU, Re):
        key = (float(meanU), float(Re))
        if key != self._key:
            c_channel.c_init(*self.grid, *self.box, float(Re), 2.0 * float(meanU), self.dt,
                             self.every)
            self._key = key

    def release(self):
        if self._key is not None:
            c_channel.c_destroy()
            self._key = None

    def _wall_sum(self, incre, weight):
        """weight * sum of the lower-wall shear over the window's steps 0 .. every-1."""
        return float(c_channel.sc_mean_profile(incre, 0, self.every, weight)[1][0])

    def forward(self, C0, meanU, Re, params=True):
        """J, the checkpoints, and (params) dJ/dmeanU and dJ/dRe."""
        self._init(meanU, Re)
        w = 1.0 / self.n_steps
        C = np.ascontiguousarray(C0, dtype=np.complex128).copy()
        I_u, I_re = np.zeros_like(C), np.zeros_like(C)
        J = dJ_u = dJ_re = 0.0
        checkpoints = []
        for _ in range(self.n_steps // self.every):
            checkpoints.append(C.copy())
            c_channel.c_primal(0, C)               # C <- the state at the window's end
            J += self._wall_sum(0, w)
            if params:
                c_channel.sc_tangent_sources(0, self.every, I_u, 1, 0)
                dJ_u += self._wall_sum(1, w)
                c_channel.sc_tangent_sources(0, self.every, I_re, 0, 1)
                dJ_re += self._wall_sum(1, w)
        return J, checkpoints, dJ_u, dJ_re

    def state_gradient(self, checkpoints, meanU, Re):
        """dJ/dC0 as the derivative in the real and imaginary parts of C0."""
        self._init(meanU, Re)
        w = 1.0 / self.n_steps
        lam = np.zeros_like(checkpoints[0])
        for C in reversed(checkpoints):
            c_channel.c_primal(0, C.copy())        # the window's trajectory again
            c_channel.sc_adjoint_objective(self.every, 0, lam, w)
        return c_channel.sc_gradient_euclid(lam)


class WallShearMembers:
    """Trajectories of lengths `lengths` (steps) as members, windows of
    `every` steps; each member's J averages the lower-wall shear over its
    own steps 0 .. L_m - 1."""

    def __init__(self, Nx, Ny, Nz, Lx, Lz, dt, lengths, every=25):
        self.grid = (int(Nx), int(Ny), int(Nz))
        self.box = (float(Lx), float(Lz))
        self.dt = float(dt)
        self.lengths = [int(n) for n in lengths]
        if not self.lengths or min(self.lengths) < 1:
            raise ValueError("every member needs at least one step")
        self.n, self.every = len(self.lengths), int(every)
        self.nwin = -(-max(self.lengths) // self.every)
        self._key = None

    def _init(self, meanU, Re):
        key = (float(meanU), float(Re))
        if key != self._key:
            c_channel.sc_set_members(self.n)
            c_channel.c_init(*self.grid, *self.box, float(Re), 2.0 * float(meanU), self.dt,
                             self.every)
            self._key = key

    def release(self):
        if self._key is not None:
            c_channel.c_destroy()
            c_channel.sc_set_members(1)
            self._key = None

    def steps_in_window(self, w):
        """The steps of each member inside window w (0 .. every)."""
        return [max(0, min(self.every, n - w * self.every)) for n in self.lengths]

    def forward(self, C0s, meanU, Re, mean_velocity=True, after_window=None):
        """J per member, the checkpoints (DeviceArrays: free them, or pass
        them to state_gradient, which does), dJ/dmeanU per member (zeros
        without mean_velocity) and each member's time-averaged mean-velocity
        profile.  after_window(w) runs after window w's primal (its stored
        states are then readable with sc_member_state)."""
        self._init(meanU, Re)
        c_channel.sc_set_member_windows(None)
        n = self.n
        C = c_channel.DeviceArray(np.ascontiguousarray(C0s, dtype=np.complex128))
        I = c_channel.DeviceArray(np.zeros(C.shape, complex)) if mean_velocity else None
        J, dJ_u = np.zeros(n), np.zeros(n)
        prof = np.zeros((n, c_channel.c_qpts()))
        cps = []
        try:
            for w in range(self.nwin):
                cps.append(C.clone())
                c_channel.sc_primal_members(0, C)
                k = self.steps_in_window(w)
                for m in range(n):
                    if k[m]:
                        p, sh = c_channel.sc_member_mean_profile(m, 0, 0, k[m],
                                                                 1.0 / self.lengths[m])
                        J[m] += sh[0]
                        prof[m] += p
                if after_window:
                    after_window(w)
                if I is not None:
                    c_channel.sc_tangent_members(0, self.every, I, (1 << n) - 1, False)
                    for m in range(n):
                        if k[m]:
                            dJ_u[m] += c_channel.sc_member_mean_profile(
                                m, 1, 0, k[m], 1.0 / self.lengths[m])[1][0]
        finally:
            C.free()
            if I is not None:
                I.free()
        return J, cps, dJ_u, prof

    def state_gradient(self, checkpoints, meanU, Re):
        """dJ_m/dC0_m for every member (derivatives in the real and imaginary
        parts), latest window first; frees the checkpoints."""
        self._init(meanU, Re)
        lam = c_channel.DeviceArray(np.zeros(checkpoints[0].shape, complex))
        try:
            for w in reversed(range(self.nwin)):
                k = self.steps_in_window(w)
                C = checkpoints[w]
                c_channel.sc_set_member_windows(None)
                c_channel.sc_primal_members(0, C)          # the window's trajectories again
                C.free()
                full = all(x == self.every for x in k)
                c_channel.sc_set_member_windows(None if full else k)
                # weight 1: the adjoint is linear in it, 1/L_m is applied below
                c_channel.sc_adjoint_objective_members(self.every, 0, lam, 1.0, False)
            L = lam.download()
        finally:
            c_channel.sc_set_member_windows(None)
            lam.free()
            for C in checkpoints:
                C.free()
        return np.stack([c_channel.sc_gradient_euclid(L[m]) / self.lengths[m]
                         for m in range(self.n)])


def _torch():
    import torch
    return torch


def wall_shear(window, C0, meanU, Re):
    """J through WallShear for torch tensors (C0 complex128, meanU and Re
    float64 scalars)."""
    return WallShear.apply(C0, meanU, Re, window)


def _make_function():
    torch = _torch()

    class WallShearFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, C0, meanU, Re, window):
            mu, re = float(meanU), float(Re)
            need_p = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
            J, cps, dJ_u, dJ_re = window.forward(C0.detach().cpu().numpy(), mu, re, need_p)
            ctx.window, ctx.cps, ctx.mu, ctx.re = window, cps, mu, re
            ctx.dJ = (dJ_u, dJ_re)
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            g = float(gJ)
            gC = None
            if ctx.needs_input_grad[0]:
                gC = torch.from_numpy(g * ctx.window.state_gradient(ctx.cps, ctx.mu, ctx.re))
            gu = torch.tensor(g * ctx.dJ[0], dtype=torch.float64) if ctx.needs_input_grad[1] else None
            gr = torch.tensor(g * ctx.dJ[1], dtype=torch.float64) if ctx.needs_input_grad[2] else None
            return gC, gu, gr, None

    return WallShearFn


class _Lazy:
    """Builds the autograd Function on first use (torch loads only then)."""

    _fn = None

    @classmethod
    def apply(cls, *args):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn.apply(*args)


WallShear = _Lazy
