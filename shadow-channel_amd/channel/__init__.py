"""channel: the Python interface the least-squares-shadowing drivers call.

Drop-in for the reference package (channel/__init__.py): the same module
parameters and function names with the same arguments and asserts
(drivers/wrapper.py, drivers/mpi_wrapper.py and drivers/testAdj.py call
them), running on the MI355X library through ``c_channel``.

Parameters, read when ``init`` / ``para_init`` start a run:

    Nx, Ny, Nz   Fourier modes in x and z, Legendre modes in y
    Lx, Lz       box size
    dt           time step
    meanU        mean velocity; the solver holds the flux 2 meanU constant
    Re           Reynolds number

The objectives (time-averaged mean profiles, their tangents, the LSS
gradient) are reductions on the device over the stored trajectory
(``c_channel.sc_mean_profile``): the plane mean of u at a Gauss point is
the real part of the mean Fourier column, so no physical field is formed
or copied to the host.
"""
import numpy as np

import c_channel

# module parameters (the reference's defaults)
Nx, Ny, Nz = 16, 33, 16
Lx, Lz = 1.6, 1.6
dt = .01
meanU = 1.
Re = 5000

__is_initialized__ = False

# the profile functions' scalar form reads this Gauss point: one above the
# middle of the ascending points (the reference's qpts/2 + 1)
_CENTRE_ABOVE = 1


def _coef_shape():
    return (c_channel.c_Nz(), 2, c_channel.c_dimR(), c_channel.c_Nx() // 2)


def _solution_file(name):
    return name if name.endswith('.hd5') else name + '.hd5'


def _start(n_steps):
    """Claim the process's single simulation and size it from the module
    parameters (the solver keeps the flux 2 meanU fixed)."""
    global __is_initialized__
    assert not __is_initialized__
    __is_initialized__ = True
    grid = (int(Nx), int(Ny), int(Nz))
    box = (float(Lx), float(Lz))
    c_channel.c_init(*grid, *box, float(Re), 2.0 * float(meanU), float(dt), int(n_steps))


def _scalar_or_profile(prof, profile):
    return prof if profile else prof[c_channel.c_qpts() // 2 + _CENTRE_ABOVE]


def _window_mean(incre, first, last, weight):
    """weight * sum of the plane-mean u profiles of stored steps first .. last-1."""
    return c_channel.sc_mean_profile(incre, first, last, weight)[0]


# ------------------------------------------------------------- runs

def quad():
    """Gauss-Legendre points and weights in y."""
    y, w = c_channel.c_quad()
    return y.copy(), w.copy()


def init(n_steps, ru_steps=0, restart=None):
    """Start a run of n_steps stored steps after ru_steps unstored run-up
    steps, from the perturbed laminar profile (restart None) or from a
    solution file."""
    assert n_steps >= 0 and ru_steps >= 0
    _start(n_steps)
    if restart is None:
        start_state = np.zeros((0, 0, 0, 0), complex)     # the library's laminar start
    else:
        start_state = read_solution(restart)
        assert start_state.shape == _coef_shape()
    c_channel.c_primal(int(ru_steps), start_state)


def para_init(n_steps, C):
    """Start a run of n_steps stored steps from the coefficient array C
    (no run-up)."""
    assert n_steps >= 0
    _start(n_steps)
    assert C.shape == _coef_shape()
    c_channel.c_primal(0, C)


def destroy():
    """Release the run."""
    global __is_initialized__
    assert __is_initialized__
    __is_initialized__ = False
    c_channel.c_destroy()


def primal(C_init):
    """Re-run the stored window from C_init, which receives the final state."""
    assert C_init.shape == _coef_shape()
    c_channel.c_primal(0, C_init)


def tangent(start_step, end_step, IC_init, inhomo):
    """Tangent from start_step to end_step in place on IC_init; inhomo 1
    adds the mean-velocity (meanU) source."""
    assert IC_init.shape == _coef_shape()
    assert 0 <= start_step <= end_step <= c_channel.c_nsteps()
    c_channel.c_tangent(int(start_step), int(end_step), IC_init, int(inhomo))


def adjoint(start_step, end_step, AC_init, inhomo, strength):
    """Adjoint from start_step back to end_step in place on AC_init, forced
    by strength times the projected tangent stored in the window."""
    assert AC_init.shape == _coef_shape()
    assert 0 <= end_step <= start_step <= c_channel.c_nsteps()
    c_channel.c_adjoint(int(start_step), int(end_step), AC_init, int(inhomo), float(strength))


def ddt_project(i_step, v):
    """Remove from v its component along the primal's time derivative at
    i_step (Gauss-weighted physical inner product); returns the coefficient
    removed."""
    return c_channel.c_ddt_project(int(i_step), v)


# ------------------------------------------------------------- fields and files

def get_solution(i_step, copy=True):
    """Stored primal state at step i_step (a library-owned view unless copied)."""
    view = c_channel.c_getsoln(int(i_step))
    assert view.shape == _coef_shape()
    return view.copy() if copy else view


def save_solution(filename, C):
    """Write C (and the run's last tangent state) to filename (.hd5 added)."""
    assert C.shape == _coef_shape()
    c_channel.c_save_solution(_solution_file(filename), C)


def read_solution(filename):
    """Coefficient array read from a file written by save_solution."""
    out = np.zeros(_coef_shape(), complex)
    c_channel.c_read_solution(str(_solution_file(filename)), out)
    return out


def spec2phys(solution):
    """Physical velocity [3, 3Nx/2, qpts, 3Nz/2] of a coefficient array, of
    a solution file (str) or of a stored step (int)."""
    if isinstance(solution, str):
        coef = read_solution(solution)
    elif isinstance(solution, (int, np.integer)):
        coef = get_solution(int(solution), copy=False)
    else:
        coef = solution
    assert coef.dtype == complex
    assert coef.shape == _coef_shape()
    shape = (3, 3 * c_channel.c_Nx() // 2, c_channel.c_qpts(), 3 * c_channel.c_Nz() // 2)
    flow = np.zeros(shape, np.float64)
    c_channel.c_spec2phys(np.ascontiguousarray(coef), flow)
    return flow


def vel(i_step):
    """Streamwise velocity of the primal at stored step i_step."""
    return spec2phys(int(i_step))[0]


def Ivel(i_step, project=False):
    """Streamwise velocity of the tangent stored at i_step, optionally made
    orthogonal to the primal's time derivative first."""
    coef = c_channel.c_getincresoln(int(i_step)).copy()
    assert coef.shape == _coef_shape()
    if project:
        ddt_project(i_step, coef)
    return spec2phys(coef)[0]


# ------------------------------------------------------------- objectives

def velAvg(start_step, end_step, T, profile=True):
    """Mean streamwise velocity profile averaged over steps start_step ..
    end_step - 1 with weight dt / T per step (its value one Gauss point
    above the middle when profile is False)."""
    assert 0 <= start_step <= end_step <= c_channel.c_nsteps()
    return _scalar_or_profile(_window_mean(0, start_step, end_step, dt / T), profile)


def IvelAvg(start_step, end_step, T, profile=True):
    """The tangent of velAvg: the same average over the stored tangent."""
    assert 0 <= start_step <= end_step <= c_channel.c_nsteps()
    return _scalar_or_profile(_window_mean(1, start_step, end_step, dt / T), profile)


def uavgGrad(n_chunk, chunk_bounds, T, uavg, zeta, profile=True):
    """Least-squares-shadowing derivative of the averaged mean profile with
    respect to meanU: per segment, the averaged tangent profile plus the
    time-dilation term zeta_i / T (profile at the segment's end - uavg)."""
    total = 0.0
    for i in range(n_chunk):
        first, last = chunk_bounds[i], chunk_bounds[i + 1]
        tangent_part = IvelAvg(first, last, T, profile)
        end_profile = _scalar_or_profile(_window_mean(0, last, last + 1, 1.0), profile)
        total = total + tangent_part + (zeta[i] / T) * (end_profile - uavg)
    return total


def wallShear(start_step, end_step, T, incre=0):
    """Time-averaged wall shear d<u>/dy of the primal (incre 0) or of the
    tangent (incre 1) over steps start_step .. end_step - 1 (weight dt / T):
    (lower wall at y = -1, upper wall at y = +1)."""
    assert 0 <= start_step <= end_step <= c_channel.c_nsteps()
    return c_channel.sc_mean_profile(incre, start_step, end_step, dt / T)[1]


def statistics(i_step, incre=0):
    """Plane statistics [10, qpts] of the stored primal (or tangent) state at
    i_step: means of u, v, w, of uv, uw, vw, uu, vv, ww, and of du/dy."""
    return c_channel.sc_step_statistics(incre, i_step)
