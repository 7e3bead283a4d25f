"""c_channel: the reference's SWIG module (channel/channel.i) over the C ABI of
libshadowchannel.so (include/shadow_channel.h).

Same function names and Python-visible behaviour as the SWIG build with
numpy.i typemaps:
  * INPLACE_ARRAY4 arguments must be C-contiguous complex128 (float64 for the
    flow array) numpy arrays of rank 4, else TypeError, as numpy.i raises;
  * c_getsoln / c_getincresoln / c_quad return numpy views of library-owned
    memory (valid until c_destroy; c_getsoln refreshes the view of that step);
  * where the C side of the reference asserts (and would abort the process),
    an AssertionError is raised instead.
The library is the only compute path: importing this module without a built
libshadowchannel.so raises ImportError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshadowchannel.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libshadowchannel.so is not built (run __graft_entry__.build() "
                      "or make -C shadow-channel_amd/csrc)")

_L = ctypes.CDLL(LIB_PATH)
_c = ctypes
_i, _d, _vp = _c.c_int, _c.c_double, _c.c_void_p
_ip = _c.POINTER(_c.c_int)
_dpp = _c.POINTER(_c.POINTER(_c.c_double))

_L.c_init.argtypes = [_i, _i, _i, _d, _d, _d, _d, _d, _i]
_L.c_init.restype = _i
_L.c_destroy.argtypes = []
_L.c_destroy.restype = None
_L.c_primal.argtypes = [_i, _vp, _i, _i, _i, _i]
_L.c_primal.restype = _i
_L.c_tangent.argtypes = [_i, _i, _vp, _i, _i, _i, _i, _i]
_L.c_tangent.restype = _i
_L.c_adjoint.argtypes = [_i, _i, _vp, _i, _i, _i, _i, _i, _d]
_L.c_adjoint.restype = _i
_L.c_ddt_project.argtypes = [_i, _vp, _i, _i, _i, _i]
_L.c_ddt_project.restype = _d
_L.c_getsoln.argtypes = [_i, _c.POINTER(_vp), _ip, _ip, _ip, _ip]
_L.c_getsoln.restype = _i
_L.c_getincresoln.argtypes = [_i, _c.POINTER(_vp), _ip, _ip, _ip, _ip]
_L.c_getincresoln.restype = _i
_L.c_spec2phys.argtypes = [_vp, _i, _i, _i, _i, _vp, _i, _i, _i, _i]
_L.c_spec2phys.restype = _i
_L.c_read_solution.argtypes = [_c.c_char_p, _vp, _i, _i, _i, _i]
_L.c_read_solution.restype = _i
_L.c_save_solution.argtypes = [_c.c_char_p, _vp, _i, _i, _i, _i]
_L.c_save_solution.restype = _i
_L.c_quad.argtypes = [_dpp, _ip, _dpp, _ip]
_L.c_quad.restype = _i
for _f in ("c_qpts", "c_dimR", "c_dimQ", "c_Nx", "c_Nz", "c_nsteps"):
    getattr(_L, _f).argtypes = []
    getattr(_L, _f).restype = _i
_L.sc_last_error.restype = _c.c_char_p
_L.sc_last_status.restype = _i
_L.sc_srand.argtypes = [_c.c_uint]
_L.sc_prof_enable.argtypes = [_i]
_L.sc_prof_enable.restype = _i
_L.sc_prof_read.argtypes = [_c.POINTER(_d)]
_L.sc_prof_read.restype = _i
_L.sc_prof_bytes.argtypes = []
_L.sc_prof_bytes.restype = _d
_L.sc_prof_replay.argtypes = [_i, _c.POINTER(_d), _c.POINTER(_d)]
_L.sc_prof_replay.restype = _i
_L.sc_prof_replay_cold.argtypes = [_i, _d, _c.POINTER(_d), _c.POINTER(_d)]
_L.sc_prof_replay_cold.restype = _i
_L.sc_synchronize.argtypes = []
_L.sc_synchronize.restype = _i
_L.sc_set_field_cache.argtypes = [_i]
_L.sc_set_field_cache.restype = _i
_L.sc_field_cache_active.argtypes = []
_L.sc_field_cache_active.restype = _i
_L.sc_tangent_forcing.argtypes = [_i, _c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p]
_L.sc_tangent_forcing.restype = _i
_L.sc_wall_plane.argtypes = [_c.c_void_p, _c.c_void_p, _c.c_void_p]
_L.sc_wall_plane.restype = _i
_L.sc_incre_boundary.argtypes = [_c.c_void_p, _c.c_void_p]
_L.sc_incre_boundary.restype = _i
_L.sc_increadjproject.argtypes = [_i, _i] + [_c.c_void_p] * 7
_L.sc_increadjproject.restype = _i
_L.sc_pass2_incremental.argtypes = [_c.c_void_p] * 5
_L.sc_pass2_incremental.restype = _i
_L.sc_increadjproject_dev.argtypes = [_i, _i] + [_c.c_void_p] * 7
_L.sc_increadjproject_dev.restype = _i
_L.sc_pass2_incremental_dev.argtypes = [_c.c_void_p] * 5
_L.sc_pass2_incremental_dev.restype = _i
_L.sc_primal_dev.argtypes = [_i, _c.c_void_p]
_L.sc_primal_dev.restype = _i
_L.sc_adjoint_dev.argtypes = [_i, _i, _c.c_void_p, _d]
_L.sc_adjoint_dev.restype = _i

_L.sc_ddt_project_noalloc.argtypes = [_i, _vp, _vp]
_L.sc_ddt_project_noalloc.restype = _i
_L.sc_set_members.argtypes = [_i]
_L.sc_set_members.restype = _i
_L.sc_members.argtypes = []
_L.sc_members.restype = _i
_L.sc_primal_members.argtypes = [_i, _vp]
_L.sc_primal_members.restype = _i
_L.sc_tangent_members.argtypes = [_i, _i, _vp, _c.c_uint, _i]
_L.sc_tangent_members.restype = _i
_L.sc_adjoint_members.argtypes = [_i, _i, _vp, _d, _i]
_L.sc_adjoint_members.restype = _i
_L.sc_member_state.argtypes = [_i, _i, _i, _vp]
_L.sc_member_state.restype = _i
_L.sc_mean_profile.argtypes = [_i, _i, _i, _d, _vp, _vp]
_L.sc_mean_profile.restype = _i
_L.sc_member_mean_profile.argtypes = [_i, _i, _i, _i, _d, _vp, _vp]
_L.sc_member_mean_profile.restype = _i
_L.sc_step_statistics.argtypes = [_i, _i, _vp]
_L.sc_step_statistics.restype = _i
_L.sc_tangent_sources.argtypes = [_i, _i, _vp, _i, _i]
_L.sc_tangent_sources.restype = _i
_L.sc_adjoint_objective.argtypes = [_i, _i, _vp, _d]
_L.sc_adjoint_objective.restype = _i
_L.sc_adjoint_objective_members.argtypes = [_i, _i, _vp, _d, _i]
_L.sc_adjoint_objective_members.restype = _i
_L.sc_gradient_euclid.argtypes = [_vp, _vp]
_L.sc_gradient_euclid.restype = _i
_L.sc_h5_dataset_shape.argtypes = [_c.c_char_p, _c.c_char_p, _c.POINTER(_c.c_longlong), _i,
                                   _ip]
_L.sc_h5_dataset_shape.restype = _i
_L.sc_h5_read_doubles.argtypes = [_c.c_char_p, _c.c_char_p, _vp, _c.c_longlong]
_L.sc_h5_read_doubles.restype = _c.c_longlong
_L.sc_h5_write_solution.argtypes = [_c.c_char_p, _i, _i, _i, _d, _d, _vp, _vp]
_L.sc_h5_write_solution.restype = _i
_L.sc_set_precision.argtypes = [_i]
_L.sc_set_precision.restype = _i
_L.sc_precision.argtypes = []
_L.sc_precision.restype = _i
_L.sc_set_tuning.argtypes = [_c.c_char_p, _i]
_L.sc_set_tuning.restype = _i
_L.sc_get_tuning.argtypes = [_c.c_char_p]
_L.sc_get_tuning.restype = _i
_L.sc_member_ddt_project.argtypes = [_i, _i, _vp]
_L.sc_member_ddt_project.restype = _d
_L.sc_set_member_previous.argtypes = [_vp]
_L.sc_set_member_previous.restype = _i
_L.sc_set_member_windows.argtypes = [_vp]
_L.sc_set_member_windows.restype = _i
_L.sc_set_graphs.argtypes = [_i]
_L.sc_set_graphs.restype = _i

SC_EINVAL = -3
SC_ENOTREADY = -2


def _inplace(a, dtype, name):
    if not isinstance(a, np.ndarray):
        raise TypeError("%s must be a numpy array" % name)
    if a.dtype != dtype:
        raise TypeError("%s must have dtype %s" % (name, np.dtype(dtype).name))
    if a.ndim != 4:
        raise TypeError("%s must have 4 dimensions" % name)
    if not a.flags.c_contiguous or not a.flags.writeable:
        raise TypeError("%s must be C-contiguous and writeable" % name)
    return a


def _check(rc, what):
    if rc == 0:
        return
    msg = "%s failed (%d): %s" % (what, rc, last_error())
    if rc in (SC_EINVAL, SC_ENOTREADY):
        raise AssertionError(msg)
    raise RuntimeError(msg)


def last_error():
    e = _L.sc_last_error()
    return e.decode() if e else ""


def last_status():
    """0 after a clean run, -1 after a CFL violation truncated the last run."""
    return _L.sc_last_status()


def c_init(Nx, Ny, Nz, Lx, Lz, Re, flux, dt, nstep):
    return _L.c_init(int(Nx), int(Ny), int(Nz), float(Lx), float(Lz), float(Re),
                     float(flux), float(dt), int(nstep))


def c_destroy():
    _L.c_destroy()


def c_primal(ru_steps, C_given):
    C = _inplace(C_given, np.complex128, "C_given")
    ptr = C.ctypes.data if C.size else None
    _check(_L.c_primal(int(ru_steps), ptr, *C.shape), "c_primal")


def c_tangent(start_step, end_step, IC_given, inhomo):
    IC = _inplace(IC_given, np.complex128, "IC_given")
    _check(_L.c_tangent(int(start_step), int(end_step), IC.ctypes.data, *IC.shape, int(inhomo)),
           "c_tangent")


def c_adjoint(start_step, end_step, AC_given, inhomo, strength):
    AC = _inplace(AC_given, np.complex128, "AC_given")
    _check(_L.c_adjoint(int(start_step), int(end_step), AC.ctypes.data, *AC.shape, int(inhomo),
                        float(strength)), "c_adjoint")


class DeviceArray:
    """Device copy of a complex128 numpy array, allocated through the HIP
    runtime the library itself uses (for sc_primal_dev / sc_adjoint_dev)."""

    _hip = None

    def __init__(self, host):
        if DeviceArray._hip is None:
            h = _c.CDLL("libamdhip64.so.7")      # the runtime already loaded by the library
            h.hipMalloc.argtypes = [_c.POINTER(_vp), _c.c_size_t]
            h.hipMemcpy.argtypes = [_vp, _vp, _c.c_size_t, _i]
            h.hipFree.argtypes = [_vp]
            DeviceArray._hip = h
        host = np.ascontiguousarray(host, dtype=np.complex128)
        self.shape, self.nbytes = host.shape, host.nbytes
        self.ptr = _vp()
        if DeviceArray._hip.hipMalloc(_c.byref(self.ptr), self.nbytes) != 0:
            raise MemoryError("hipMalloc of %d bytes failed" % self.nbytes)
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.complex128)
        assert host.nbytes == self.nbytes
        if DeviceArray._hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy host to device failed")

    def clone(self):
        """A second device array with the same contents (device-to-device)."""
        out = DeviceArray.__new__(DeviceArray)
        out.shape, out.nbytes, out.ptr = self.shape, self.nbytes, _vp()
        if DeviceArray._hip.hipMalloc(_c.byref(out.ptr), self.nbytes) != 0:
            raise MemoryError("hipMalloc of %d bytes failed" % self.nbytes)
        if DeviceArray._hip.hipMemcpy(out.ptr, self.ptr, self.nbytes, 3) != 0:
            raise RuntimeError("hipMemcpy device to device failed")
        return out

    def download(self):
        out = np.empty(self.shape, dtype=np.complex128)
        if DeviceArray._hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy device to host failed")
        return out

    def free(self):
        if self.ptr:
            DeviceArray._hip.hipFree(self.ptr)
            self.ptr = _vp()


def _devptr(x):
    return x.ptr if isinstance(x, DeviceArray) else _vp(int(x))


def sc_primal_dev(ru_steps, C_dev):
    """c_primal on a device-resident state: C_dev is a DeviceArray (or the
    device address) of a [Nz][2][dimR][Nx/2] complex128 array, overwritten
    with the final state."""
    _check(_L.sc_primal_dev(int(ru_steps), _devptr(C_dev)), "sc_primal_dev")


def sc_adjoint_dev(start_step, end_step, AC_dev, strength):
    """c_adjoint on a device-resident adjoint state (see sc_primal_dev)."""
    _check(_L.sc_adjoint_dev(int(start_step), int(end_step), _devptr(AC_dev), float(strength)),
           "sc_adjoint_dev")


def c_ddt_project(i_step, IC_given):
    IC = _inplace(IC_given, np.complex128, "IC_given")
    r = _L.c_ddt_project(int(i_step), IC.ctypes.data, *IC.shape)
    if np.isnan(r):
        # NaN with an error text is a failed call (an invalid one raises
        # AssertionError, as the reference's assert would abort); NaN
        # without one is the projection's own 0/0 (a constant trajectory)
        msg = last_error()
        if msg.startswith("ddt_project requires"):
            raise AssertionError("c_ddt_project: " + msg)
        if msg:
            raise RuntimeError("c_ddt_project: " + msg)
    return r


def _view(fn, i_step):
    p = _vp()
    dims = [_i() for _ in range(4)]
    _check(fn(int(i_step), _c.byref(p), *[_c.byref(x) for x in dims]), fn.__name__)
    shape = tuple(x.value for x in dims)
    n = int(np.prod(shape))
    buf = (_c.c_double * (2 * n)).from_address(p.value)
    return np.frombuffer(buf, dtype=np.complex128).reshape(shape)


def c_getsoln(i_step):
    return _view(_L.c_getsoln, i_step)


def c_getincresoln(i_step):
    return _view(_L.c_getincresoln, i_step)


def c_spec2phys(C_ptr, flow_ptr):
    C = _inplace(C_ptr, np.complex128, "C")
    F = _inplace(flow_ptr, np.float64, "flow")
    _check(_L.c_spec2phys(C.ctypes.data, *C.shape, F.ctypes.data, *F.shape), "c_spec2phys")


def c_read_solution(filename, C_ptr):
    C = _inplace(C_ptr, np.complex128, "C")
    _check(_L.c_read_solution(str(filename).encode(), C.ctypes.data, *C.shape),
           "c_read_solution")


def c_save_solution(filename, C_ptr):
    C = _inplace(C_ptr, np.complex128, "C")
    _check(_L.c_save_solution(str(filename).encode(), C.ctypes.data, *C.shape),
           "c_save_solution")


def c_quad():
    y = _c.POINTER(_c.c_double)()
    w = _c.POINTER(_c.c_double)()
    n1, n2 = _i(), _i()
    _check(_L.c_quad(_c.byref(y), _c.byref(n1), _c.byref(w), _c.byref(n2)), "c_quad")
    return (np.ctypeslib.as_array(y, shape=(n1.value,)),
            np.ctypeslib.as_array(w, shape=(n2.value,)))


def c_qpts():
    return _L.c_qpts()


def c_dimR():
    return _L.c_dimR()


def c_dimQ():
    return _L.c_dimQ()


def c_Nx():
    return _L.c_Nx()


def c_Nz():
    return _L.c_Nz()


def c_nsteps():
    return _L.c_nsteps()


def sc_srand(seed):
    _L.sc_srand(int(seed))


def sc_prof_enable(kernel_class):
    return _L.sc_prof_enable(int(kernel_class))


def sc_prof_read():
    """(launches, milliseconds, algorithmic bytes) recorded since sc_prof_enable."""
    nbytes = _L.sc_prof_bytes()
    ms = _d(0.0)
    n = _L.sc_prof_read(_c.byref(ms))
    return n, ms.value, nbytes


def sc_prof_replay(reps):
    """(average milliseconds, algorithmic bytes) of the last recorded launch,
    re-issued reps times back to back (overwrites that kernel's outputs)."""
    ms, nb = _d(0.0), _d(0.0)
    _check(_L.sc_prof_replay(int(reps), _c.byref(ms), _c.byref(nb)), "sc_prof_replay")
    return ms.value, nb.value


def sc_prof_replay_cold(reps, flush_mb=320.0):
    """Like sc_prof_replay, each launch preceded by a cache-evicting fill of
    flush_mb MB whose own time is subtracted: (avg_ms, bytes)."""
    ms, nb = _c.c_double(), _c.c_double()
    _check(_L.sc_prof_replay_cold(int(reps), float(flush_mb), _c.byref(ms), _c.byref(nb)),
           "sc_prof_replay_cold")
    return ms.value, nb.value


def sc_set_field_cache(enable):
    """Enable (default) or disable the transformed-field cache of the next
    c_init (results are the same either way)."""
    _L.sc_set_field_cache(int(bool(enable)))


def sc_field_cache_active():
    return bool(_L.sc_field_cache_active())


def sc_synchronize():
    _check(_L.sc_synchronize(), "sc_synchronize")


def sc_tangent_forcing(count):
    """Reynolds-number sensitivity forcing of the stored state MC[count]
    (channel/tangent_forcing.c, channel/tangent_forcing0.c): returns
    (f_a [Nz, dimQ, Nx/2], f_b [Nz, dimR, Nx/2], f0_a [dimR], f0_b [dimR])."""
    Nz, Mx, nQ, nR = _L.c_Nz(), _L.c_Nx() // 2, _L.c_dimQ(), _L.c_dimR()
    fa = np.zeros((Nz, nQ, Mx), complex)
    fb = np.zeros((Nz, nR, Mx), complex)
    f0a = np.zeros(nR, complex)
    f0b = np.zeros(nR, complex)
    _check(_L.sc_tangent_forcing(int(count), fa.ctypes.data, fb.ctypes.data, f0a.ctypes.data,
                                 f0b.ctypes.data), "sc_tangent_forcing")
    return fa, fb, f0a, f0b


def sc_wall_plane(C_ptr):
    """Wall plane of a coefficient array (the Uxb, Uzb of initAlphaBeta2,
    channel/initAlphaBeta2.c:70-81, 210-223): returns (Uxb, Uzb), each
    [Nz, Nx/2] complex128."""
    C = _inplace(C_ptr, np.complex128, "C")
    Nz, Mx, nR = _L.c_Nz(), _L.c_Nx() // 2, _L.c_dimR()
    assert C.shape == (Nz, 2, nR, Mx), "C must be (Nz, 2, dimR, Nx/2)"
    ub = np.zeros((Nz, Mx), complex)
    wb = np.zeros((Nz, Mx), complex)
    _check(_L.sc_wall_plane(C.ctypes.data, ub.ctypes.data, wb.ctypes.data), "sc_wall_plane")
    return ub, wb


def sc_incre_boundary(Uxb, Uzb):
    """increBoundary (channel/increBoundary.c:20-162) in place on two wall
    planes, each a C-contiguous complex128 [Nz, Nx/2] array."""
    shape = (_L.c_Nz(), _L.c_Nx() // 2)
    for a, name in ((Uxb, "Uxb"), (Uzb, "Uzb")):
        if not isinstance(a, np.ndarray) or a.dtype != np.complex128:
            raise TypeError("%s must be a complex128 numpy array" % name)
        if a.shape != shape or not a.flags.c_contiguous or not a.flags.writeable:
            raise TypeError("%s must be C-contiguous, writeable, of shape (Nz, Nx/2)" % name)
    ub, wb = Uxb, Uzb
    _check(_L.sc_incre_boundary(ub.ctypes.data, wb.ctypes.data), "sc_incre_boundary")


def sc_pass2_incremental(U, IU, AU, IAU):
    """The incremental side of pass2 (channel/pass2.c:120-566) from the
    Gauss-point arrays [Nz, 5, qpts, Nx/2] of the state, tangent, adjoint and
    incremental adjoint: returns IH (H slots 0..2, IS_x, IS_z in slots 3, 4),
    the products sc_increadjproject takes."""
    Nz, Mx, q = _L.c_Nz(), _L.c_Nx() // 2, _L.c_qpts()
    arrs = []
    for a, name in ((U, "U"), (IU, "IU"), (AU, "AU"), (IAU, "IAU")):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == (Nz, 5, q, Mx), "%s must be (Nz, 5, qpts, Nx/2)" % name
        arrs.append(a)
    IH = np.zeros((Nz, 5, q, Mx), complex)
    _check(_L.sc_pass2_incremental(*[a.ctypes.data for a in arrs], IH.ctypes.data),
           "sc_pass2_incremental")
    return IH


def sc_pass2_incremental_dev(U_dev, IU_dev, AU_dev, IAU_dev, IH_dev):
    """sc_pass2_incremental on DeviceArray operands (IH_dev written)."""
    _check(_L.sc_pass2_incremental_dev(U_dev.ptr, IU_dev.ptr, AU_dev.ptr, IAU_dev.ptr, IH_dev.ptr),
           "sc_pass2_incremental_dev")


def sc_increadjproject_dev(k, count, IAC_dev, IU_dev, LIU_dev, walls_dev, force_dev, IAU_dev,
                           IAUb_dev):
    """sc_increadjproject on DeviceArray operands (LIU_dev and force_dev may
    be None); IAC_dev in place, IAU_dev and IAUb_dev written."""
    ptr = lambda a: None if a is None else a.ptr
    _check(_L.sc_increadjproject_dev(k, count, IAC_dev.ptr, IU_dev.ptr, ptr(LIU_dev),
                                     walls_dev.ptr, ptr(force_dev), IAU_dev.ptr, IAUb_dev.ptr),
           "sc_increadjproject_dev")


def sc_increadjproject(k, count, IAC, IU, LIU, walls, force=None):
    """One Runge-Kutta stage k of the incremental adjoint over every mode
    (increadjproject0 and increadjproject, channel/increadjproject0.c:14-357,
    channel/increadjproject.c:10-645): IAC [Nz, 2, dimR, Nx/2] is updated in
    place (the next stage's right-hand side for k < 2, the solution at k = 2);
    IU, LIU [Nz, 5, qpts, Nx/2] are the products of this stage and the next
    (LIU may be None at k = 2); walls [6, Nz, Nx/2] holds Uxb, Uzb, Uxbt, Uzbt,
    AUxb, AUzb; force [2, 2, Nz, Nx/2, dimR] or None.  Uses the tangent states
    MIC[count], MIC[count - 1] of the last tangent run.  Returns (IAU
    [Nz, 5, qpts, Nx/2], IAUb [2, Nz, Nx/2])."""
    Nz, Mx, nR, q = _L.c_Nz(), _L.c_Nx() // 2, _L.c_dimR(), _L.c_qpts()
    IAC = _inplace(IAC, np.complex128, "IAC")
    assert IAC.shape == (Nz, 2, nR, Mx), "IAC must be (Nz, 2, dimR, Nx/2)"

    def arr(a, shape, name):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == shape, "%s must have shape %s" % (name, shape)
        return a

    IU = arr(IU, (Nz, 5, q, Mx), "IU")
    LIU = None if LIU is None else arr(LIU, (Nz, 5, q, Mx), "LIU")
    walls = arr(walls, (6, Nz, Mx), "walls")
    force = None if force is None else arr(force, (2, 2, Nz, Mx, nR), "force")
    IAU = np.zeros((Nz, 5, q, Mx), complex)
    IAUb = np.zeros((2, Nz, Mx), complex)
    _check(_L.sc_increadjproject(k, count, IAC.ctypes.data, IU.ctypes.data,
                                 None if LIU is None else LIU.ctypes.data, walls.ctypes.data,
                                 None if force is None else force.ctypes.data,
                                 IAU.ctypes.data, IAUb.ctypes.data), "sc_increadjproject")
    return IAU, IAUb


def sc_mean_profile(incre, start_step, end_step, weight):
    """weight * (sum over stored steps start_step .. end_step - 1 of the x-z
    plane-mean streamwise velocity profile at the Gauss points) of the primal
    (incre 0) or tangent (incre 1); returns (profile[qpts], shear[2]) with the
    same sums of d<u>/dy at y = -1 and y = +1."""
    prof = np.zeros(c_qpts())
    shear = np.zeros(2)
    _check(_L.sc_mean_profile(int(incre), int(start_step), int(end_step), float(weight),
                              prof.ctypes.data, shear.ctypes.data), "sc_mean_profile")
    return prof, shear


def sc_member_mean_profile(member, incre, start_step, end_step, weight):
    """sc_mean_profile of one member's stored trajectory."""
    prof = np.zeros(c_qpts())
    shear = np.zeros(2)
    _check(_L.sc_member_mean_profile(int(member), int(incre), int(start_step), int(end_step),
                                     float(weight), prof.ctypes.data, shear.ctypes.data),
           "sc_member_mean_profile")
    return prof, shear


def sc_step_statistics(incre, i_step):
    """Plane statistics [10, qpts] of the stored state at i_step: means of u,
    v, w, then of uv, uw, vw, uu, vv, ww, then of du/dy."""
    st = np.zeros((10, c_qpts()))
    _check(_L.sc_step_statistics(int(incre), int(i_step), st.ctypes.data), "sc_step_statistics")
    return st


def sc_ddt_project_noalloc(i_step, v_dev, r_dev):
    """c_ddt_project on device memory without any allocation: v_dev (a
    DeviceArray or device address of a [Nz][2][dimR][Nx/2] complex128
    array) is projected in place, r_dev (device address of 3 doubles)
    receives num, den and num/den.  Asynchronous on the library's stream
    (sc_synchronize waits for it)."""
    _check(_L.sc_ddt_project_noalloc(int(i_step), _devptr(v_dev), _devptr(r_dev)),
           "sc_ddt_project_noalloc")


def sc_set_members(n):
    """Members the next c_init allocates (1 .. 32)."""
    _check(_L.sc_set_members(int(n)), "sc_set_members")


def sc_members():
    return _L.sc_members()


def sc_primal_members(ru_steps, C_dev):
    """Primal runs of all members from their states in C_dev (a DeviceArray
    of shape (members, Nz, 2, dimR, Nx/2)), overwritten with the final states."""
    _check(_L.sc_primal_members(int(ru_steps), _devptr(C_dev)), "sc_primal_members")


def sc_tangent_members(start_step, end_step, IC_dev, inhomo_mask=0, shared_base=False):
    """Tangents of all members in one pass of every kernel; bit m of
    inhomo_mask adds the mean-velocity source to member m; shared_base: all
    along member 0's trajectory."""
    _check(_L.sc_tangent_members(int(start_step), int(end_step), _devptr(IC_dev),
                                 int(inhomo_mask), int(bool(shared_base))), "sc_tangent_members")


def sc_adjoint_members(start_step, end_step, AC_dev, strength, shared_base=False):
    """Adjoints of all members (forced by their own stored tangents)."""
    _check(_L.sc_adjoint_members(int(start_step), int(end_step), _devptr(AC_dev),
                                 float(strength), int(bool(shared_base))), "sc_adjoint_members")


def sc_member_state(incre, member, i_step):
    """MC[3 i_step] (incre 0) or MIC[3 i_step] (incre 1) of one member."""
    out = np.zeros((_L.c_Nz(), 2, _L.c_dimR(), _L.c_Nx() // 2), complex)
    _check(_L.sc_member_state(int(incre), int(member), int(i_step), out.ctypes.data),
           "sc_member_state")
    return out


def _coef_array(a, name):
    a = _inplace(a, np.complex128, name)
    shape = (_L.c_Nz(), 2, _L.c_dimR(), _L.c_Nx() // 2)
    if a.shape != shape:
        raise TypeError("%s must have shape %s" % (name, shape))
    return a


def sc_tangent_sources(start_step, end_step, IC, inhomo, re_source):
    """c_tangent with the mean-velocity source (inhomo) and/or the
    Reynolds-number source (re_source), in place on IC."""
    IC = _coef_array(IC, "IC")
    _check(_L.sc_tangent_sources(int(start_step), int(end_step), IC.ctypes.data, int(inhomo),
                                 int(re_source)), "sc_tangent_sources")


def sc_adjoint_objective_members(start_step, end_step, AC_dev, weight, shared_base=False):
    """sc_adjoint_objective of all members in every launch (AC_dev: a
    DeviceArray of the members' states, in place); member windows apply."""
    _check(_L.sc_adjoint_objective_members(int(start_step), int(end_step), _devptr(AC_dev),
                                           float(weight), int(bool(shared_base))),
           "sc_adjoint_objective_members")


def sc_adjoint_objective(start_step, end_step, AC, weight):
    """Adjoint of weight * (sum of the lower-wall shear over steps end_step ..
    start_step - 1), in place on AC (the exact transpose of the tangent)."""
    AC = _coef_array(AC, "AC")
    _check(_L.sc_adjoint_objective(int(start_step), int(end_step), AC.ctypes.data,
                                   float(weight)), "sc_adjoint_objective")


def sc_gradient_euclid(lam):
    """g with sum Re(conj(g) d) = <lam, d> in the physical inner product."""
    lam = np.ascontiguousarray(lam, dtype=np.complex128)
    g = np.zeros_like(lam)
    _check(_L.sc_gradient_euclid(lam.ctypes.data, g.ctypes.data), "sc_gradient_euclid")
    return g


def sc_set_graphs(enable):
    """Record each stepping loop once as a HIP graph and replay it on later
    calls with the same arguments (1), or launch kernel by kernel (0)."""
    _check(_L.sc_set_graphs(int(enable)), "sc_set_graphs")


def h5_read(path, dataset):
    """A root-group dataset of an HDF5 file as a numpy array of its shape
    (complex128 for the reference's {real, imaginary} compound, float64
    otherwise), read by the library without libhdf5."""
    dims = (_c.c_longlong * 32)()
    kind = _c.c_int()
    rank = _L.sc_h5_dataset_shape(str(path).encode(), dataset.encode(), dims, 32,
                                  _c.byref(kind))
    _check(rank if rank < 0 else 0, "sc_h5_dataset_shape")
    shape = tuple(int(d) for d in dims[:rank])
    n = int(np.prod(shape)) * (2 if kind.value == 2 else 1)
    out = np.zeros(n)
    got = _L.sc_h5_read_doubles(str(path).encode(), dataset.encode(), out.ctypes.data, n)
    _check(int(got) if got < 0 else 0, "sc_h5_read_doubles")
    if kind.value == 2:
        return out.view(np.complex128).reshape(shape)
    return out.reshape(shape)


def h5_write_solution(path, Nx, Ny, Nz, dt, Re, C, IC):
    """save_solution's HDF5 file from host arrays C and IC [Nz][2][Ny-2][Nx/2]
    (no c_init needed)."""
    shape = (int(Nz), 2, int(Ny) - 2, int(Nx) // 2)
    C = np.ascontiguousarray(C, dtype=np.complex128)
    IC = np.ascontiguousarray(IC, dtype=np.complex128)
    if C.shape != shape or IC.shape != shape:
        raise TypeError("C and IC must have shape %s" % (shape,))
    _check(_L.sc_h5_write_solution(str(path).encode(), int(Nx), int(Ny), int(Nz), float(dt),
                                   float(Re), C.ctypes.data, IC.ctypes.data),
           "sc_h5_write_solution")


def sc_set_precision(mixed):
    """Precision of the next c_init: 0 FP64, 1 mixed (transforms and products
    of the time steps in FP32; solves, y-transforms and states FP64)."""
    _check(_L.sc_set_precision(int(mixed)), "sc_set_precision")


def sc_precision():
    return _L.sc_precision()


def sc_set_tuning(name, value):
    """Launch shape `name` (solve_members_per_block, rebuild_blocks,
    force_blocks) for later launches; 0 restores the built-in rule."""
    _check(_L.sc_set_tuning(str(name).encode(), int(value)), "sc_set_tuning")


def sc_get_tuning(name):
    v = _L.sc_get_tuning(str(name).encode())
    _check(v if v < 0 else 0, "sc_get_tuning")
    return v


def sc_member_ddt_project(member, i_step, v):
    """c_ddt_project of v (in place) on member `member`'s trajectory at
    i_step; returns num/den."""
    v = _coef_array(v, "v")
    r = _L.sc_member_ddt_project(int(member), int(i_step), v.ctypes.data)
    if np.isnan(r) and last_error():
        raise AssertionError("sc_member_ddt_project failed: %s" % last_error())
    return r


def sc_set_member_previous(prev):
    """The state one step before each member's window (members x [Nz][2]
    [dimR][Nx/2]; member 0's entry unused), or None to switch it off."""
    if prev is None:
        _check(_L.sc_set_member_previous(None), "sc_set_member_previous")
        return
    prev = np.ascontiguousarray(prev, dtype=np.complex128)
    _check(_L.sc_set_member_previous(prev.ctypes.data), "sc_set_member_previous")


def sc_set_member_windows(lengths):
    """Windows of their own for the members (the last step of member m's
    window, one per member), or None: every member over the whole run."""
    if lengths is None:
        _check(_L.sc_set_member_windows(None), "sc_set_member_windows")
        return
    n = _L.sc_members()
    w = np.ascontiguousarray(lengths, dtype=np.int32)
    if w.shape != (n,):
        raise TypeError("sc_set_member_windows needs one length per member (%d)" % n)
    _check(_L.sc_set_member_windows(w.ctypes.data), "sc_set_member_windows")
