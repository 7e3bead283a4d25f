"""channel: Python interface of the channel DNS with tangent and adjoint.

Python-3 counterpart of the reference wrapper (channel/__init__.py:1-300),
with the same module parameters, functions, argument meanings and asserts,
running on the MI355X library through ``c_channel``.  Parameters are read at
init() time:

    Nx, Ny, Nz   Fourier modes in x and z, Legendre modes in y
    Lx, Lz       box size
    dt           time step
    meanU        mean velocity; the solver holds the flux 2*meanU constant
    Re           Reynolds number
"""
import numpy as np

import c_channel

__is_initialized__ = False

Nx = 16
Ny = 33
Nz = 16

Lx = 1.6
Lz = 1.6

dt = .01

meanU = 1.

Re = 5000


def _shape():
    return (c_channel.c_Nz(), 2, c_channel.c_dimR(), c_channel.c_Nx() // 2)


def quad():
    """Gauss-Legendre points and weights in y (channel/__init__.py:28-30)."""
    y, w = c_channel.c_quad()
    return y.copy(), w.copy()


def save_solution(filename, C):
    """Save C (shape (Nz, 2, dimR, Nx/2)) and the last tangent state to a
    solution file; '.hd5' is appended when missing (channel/__init__.py:32-42).
    """
    if not filename.endswith('.hd5'):
        filename += '.hd5'
    assert C.shape == _shape()
    c_channel.c_save_solution(filename, C)


def read_solution(filename):
    """Read a solution file written by save_solution (channel/__init__.py:44-55)."""
    if not filename.endswith('.hd5'):
        filename += '.hd5'
    C = np.zeros(_shape(), complex)
    c_channel.c_read_solution(str(filename), C)
    return C


def get_solution(i_step, copy=True):
    """Primal solution at step i_step (channel/__init__.py:57-69)."""
    C = c_channel.c_getsoln(int(i_step))
    assert C.shape == _shape()
    if copy:
        C = C.copy()
    return C


def init(n_steps, ru_steps=0, restart=None):
    """Initialise and run the primal over ru_steps (not stored) + n_steps
    (stored).  restart None: perturbed laminar start; otherwise a solution
    file (channel/__init__.py:71-94)."""
    global __is_initialized__
    assert not __is_initialized__
    __is_initialized__ = True

    assert n_steps >= 0 and ru_steps >= 0
    c_channel.c_init(int(Nx), int(Ny), int(Nz),
                     float(Lx), float(Lz), float(Re),
                     float(meanU * 2), float(dt), int(n_steps))
    if restart is None:
        c_channel.c_primal(int(ru_steps), np.zeros([0, 0, 0, 0], complex))
    else:
        C = read_solution(restart)
        assert C.shape == _shape()
        c_channel.c_primal(int(ru_steps), C)


def para_init(n_steps, C):
    """Like init, with the initial condition given as an array and no run-up
    (channel/__init__.py:96-112)."""
    global __is_initialized__
    assert not __is_initialized__
    __is_initialized__ = True

    assert n_steps >= 0
    c_channel.c_init(int(Nx), int(Ny), int(Nz),
                     float(Lx), float(Lz), float(Re),
                     float(meanU * 2), float(dt), int(n_steps))
    assert C.shape == _shape()
    c_channel.c_primal(0, C)


def destroy():
    """Release the simulation (channel/__init__.py:114-122)."""
    global __is_initialized__
    assert __is_initialized__
    __is_initialized__ = False
    c_channel.c_destroy()


def primal(C_init):
    """Run nsteps primal steps from C_init; C_init receives the final state
    (channel/__init__.py:124-133)."""
    assert C_init.shape == _shape()
    c_channel.c_primal(0, C_init)


def tangent(start_step, end_step, IC_init, inhomo):
    """Tangent from start_step to end_step in place on IC_init; inhomo=1 adds
    the mean-velocity (meanU) source (channel/__init__.py:135-145)."""
    assert IC_init.shape == _shape()
    assert 0 <= start_step <= end_step <= c_channel.c_nsteps()
    c_channel.c_tangent(int(start_step), int(end_step), IC_init, int(inhomo))


def adjoint(start_step, end_step, AC_init, inhomo, strength):
    """Adjoint from start_step back to end_step in place on AC_init, forced
    by strength * (projected tangent) (channel/__init__.py:147-152)."""
    assert AC_init.shape == _shape()
    assert 0 <= end_step <= start_step <= c_channel.c_nsteps()
    c_channel.c_adjoint(int(start_step), int(end_step), AC_init, int(inhomo),
                        float(strength))


def spec2phys(solution):
    """Physical velocity flow[3, 3Nx/2, qpts, 3Nz/2] of a solution given as an
    array, a file name or a step index (channel/__init__.py:154-167)."""
    if isinstance(solution, str):
        solution = read_solution(solution)
    if isinstance(solution, (int, np.integer)):
        solution = get_solution(int(solution), copy=False)

    assert solution.dtype == complex
    assert solution.shape == _shape()

    flow = np.zeros([3, int(c_channel.c_Nx() * 3 // 2), c_channel.c_qpts(),
                     int(c_channel.c_Nz() * 3 // 2)], np.float64)
    c_channel.c_spec2phys(np.ascontiguousarray(solution), flow)
    return flow


def ddt_project(i_step, v):
    """Make v orthogonal to the primal time derivative at i_step in the
    Gauss-weighted physical inner product; returns the removed coefficient
    (channel/__init__.py:199-205)."""
    return c_channel.c_ddt_project(int(i_step), v)


def vel(i_step):
    """x velocity of the primal at i_step (channel/__init__.py:207-215)."""
    C = get_solution(i_step, copy=True)
    u = spec2phys(C)
    return u[0]


def velAvg(start_step, end_step, T, profile=True):
    """Time-averaged mean x-velocity profile, or its centreline value
    (channel/__init__.py:219-239)."""
    if profile:
        uavg = np.zeros(c_channel.c_qpts())
    else:
        uavg = 0.0
    assert 0 <= start_step <= end_step <= c_channel.c_nsteps()
    for i in range(start_step, end_step):
        u = vel(i)
        u = (u.mean(2)).mean(0)
        if not profile:
            u = u[c_channel.c_qpts() // 2 + 1]
        uavg = uavg + (dt / T) * u
    return uavg


def Ivel(i_step, project=False):
    """x velocity of the tangent stored at i_step (channel/__init__.py:243-257)."""
    IC = c_channel.c_getincresoln(int(i_step))
    IC = IC.copy()
    assert IC.shape == _shape()
    if project:
        ddt_project(i_step, IC)
    v = spec2phys(IC)
    return v[0]


def IvelAvg(start_step, end_step, T, profile=True):
    """Time-averaged mean tangent x-velocity profile (channel/__init__.py:259-273)."""
    if profile:
        vavg = np.zeros(c_channel.c_qpts())
    else:
        vavg = 0.0
    assert 0 <= start_step <= end_step <= c_channel.c_nsteps()
    for i in range(start_step, end_step):
        v = Ivel(i, project=False)
        v = (v.mean(2)).mean(0)
        if not profile:
            v = v[c_channel.c_qpts() // 2 + 1]
        vavg = vavg + (dt / T) * v
    return vavg


def uavgGrad(n_chunk, chunk_bounds, T, uavg, zeta, profile=True):
    """LSS sensitivity of the time-averaged mean x-velocity to meanU, with
    time-dilation terms zeta (channel/__init__.py:275-299)."""
    if profile:
        grad = np.zeros(c_channel.c_qpts())
    else:
        grad = 0.0
    for i in range(n_chunk):
        start_step = chunk_bounds[i]
        end_step = chunk_bounds[i + 1]
        vavg = IvelAvg(start_step, end_step, T, profile)
        uEnd = (vel(end_step).mean(2)).mean(0)
        if not profile:
            uEnd = uEnd[c_channel.c_qpts() // 2 + 1]
        grad = grad + vavg + (1. / T) * zeta[i] * (uEnd - uavg)
    return grad
