"""ctypes binding of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The functions mirror the reference's
``c_channel`` calls (channel/channel.i) on numpy arrays.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib_path():
    """The oracle library; ORC_LIB names another build of the same sources
    (the planted-mistake tests load mutated copies this way)."""
    return os.environ.get("ORC_LIB") or os.path.join(_HERE, "liborc.so")


def build():
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    c = ctypes
    dp = c.POINTER(c.c_double)
    L.orc_init.argtypes = [c.c_int, c.c_int, c.c_int, c.c_double, c.c_double,
                           c.c_double, c.c_double, c.c_double, c.c_int]
    L.orc_init.restype = c.c_int
    L.orc_primal.argtypes = [c.c_int, c.c_void_p]
    L.orc_primal.restype = c.c_int
    L.orc_tangent.argtypes = [c.c_int, c.c_int, c.c_void_p, c.c_int]
    L.orc_tangent.restype = c.c_int
    L.orc_adjoint.argtypes = [c.c_int, c.c_int, c.c_void_p, c.c_int, c.c_double]
    L.orc_adjoint.restype = c.c_int
    L.orc_ddt_project.argtypes = [c.c_int, c.c_void_p]
    L.orc_ddt_project.restype = c.c_double
    L.orc_spec2phys.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_getsoln.argtypes = [c.c_int]
    L.orc_getsoln.restype = c.c_void_p
    L.orc_getincresoln.argtypes = [c.c_int]
    L.orc_getincresoln.restype = c.c_void_p
    L.orc_getincrestage.argtypes = [c.c_int]
    L.orc_getincrestage.restype = c.c_void_p
    L.orc_quad.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_increadjproject.argtypes = [c.c_int, c.c_int] + [c.c_void_p] * 7
    L.orc_increadjproject.restype = c.c_int
    L.orc_adjproject_stage.argtypes = [c.c_int] + [c.c_void_p] * 4
    L.orc_adjproject_stage.restype = c.c_int
    L.orc_pass2_incremental.argtypes = [c.c_void_p] * 5
    L.orc_pass2_incremental.restype = c.c_int
    L.orc_get_array.argtypes = [c.c_char_p, dp, c.c_int]
    L.orc_get_array.restype = c.c_int
    L.orc_get_scalar.argtypes = [c.c_char_p]
    L.orc_get_scalar.restype = c.c_double
    L.orc_inv2d.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_fwd2d.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_band_solve.argtypes = [c.c_void_p, c.c_void_p, c.c_int, c.c_int]
    L.orc_pass1_state.argtypes = [c.c_void_p, c.c_int]
    L.orc_pass1_state.restype = c.c_int
    L.orc_pass2_state.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_coef_to_gauss.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_get_state.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_set_faithful_cost.argtypes = [c.c_int]
    L.orc_srand.argtypes = [c.c_uint]
    L.orc_tangent_forcing.argtypes = [c.c_int, c.c_void_p, c.c_void_p, c.c_void_p, c.c_void_p]
    L.orc_wall_plane.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p]
    L.orc_incre_boundary.argtypes = [c.c_void_p, c.c_void_p]
    L.orc_incre_boundary.restype = c.c_int
    L.orc_rand.restype = c.c_int
    L.orc_flux_trace.argtypes = [c.c_void_p, c.c_int]
    L.orc_flux_trace.restype = c.c_int
    L.orc_set_adjoint_refresh.argtypes = [c.c_int]
    L.orc_tangent_src.argtypes = [c.c_int, c.c_int, c.c_void_p, c.c_int, c.c_int]
    L.orc_tangent_src.restype = c.c_int
    L.orc_adjoint_objective.argtypes = [c.c_int, c.c_int, c.c_void_p, c.c_double]
    L.orc_adjoint_objective.restype = c.c_int
    L.orc_set_exact_mean_adjoint.argtypes = [c.c_int]
    for f in ("orc_qpts", "orc_dimR", "orc_dimQ", "orc_Nx", "orc_Nz",
              "orc_nsteps", "orc_last_status"):
        getattr(L, f).restype = c.c_int
    _LIB = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check_c(a):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a


class Oracle:
    """One simulation (the oracle, like the reference, is process-global)."""

    def __init__(self, Nx, Ny, Nz, Lx, Lz, Re, flux, dt, nsteps):
        self.L = load()
        rc = self.L.orc_init(Nx, Ny, Nz, Lx, Lz, Re, flux, dt, nsteps)
        if rc != 0:
            raise ValueError("oracle init failed")
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.qpts = self.L.orc_qpts()
        self.dimR = self.L.orc_dimR()
        self.dimQ = self.L.orc_dimQ()
        self.nsteps = nsteps
        self.cshape = (Nz, 2, self.dimR, Nx // 2)
        self.ushape = (Nz, 5, self.qpts, Nx // 2)

    def primal(self, ru_steps, C=None):
        if C is None:
            return self.L.orc_primal(ru_steps, None)
        return self.L.orc_primal(ru_steps, _ptr(_check_c(C)))

    def tangent(self, start, end, IC, inhomo):
        return self.L.orc_tangent(start, end, _ptr(_check_c(IC)), inhomo)

    def adjoint(self, start, end, AC, inhomo, strength):
        return self.L.orc_adjoint(start, end, _ptr(_check_c(AC)), inhomo, strength)

    def tangent_src(self, start, end, IC, inhomo, re_src):
        return self.L.orc_tangent_src(start, end, _ptr(_check_c(IC)), inhomo, re_src)

    def adjoint_objective(self, start, end, AC, weight):
        return self.L.orc_adjoint_objective(start, end, _ptr(_check_c(AC)), weight)

    def ddt_project(self, i, v):
        return self.L.orc_ddt_project(i, _ptr(_check_c(v)))

    def spec2phys(self, C):
        flow = np.zeros((3, 3 * self.Nx // 2, self.qpts, 3 * self.Nz // 2))
        self.L.orc_spec2phys(_ptr(_check_c(np.ascontiguousarray(C))), _ptr(flow))
        return flow

    def _view(self, p):
        n = int(np.prod(self.cshape))
        buf = (ctypes.c_double * (2 * n)).from_address(p)
        return np.frombuffer(buf, dtype=np.complex128).reshape(self.cshape)

    def getsoln(self, i):
        p = self.L.orc_getsoln(i)
        assert p
        return self._view(p).copy()

    def getincrestage(self, count):
        p = self.L.orc_getincrestage(count)
        assert p
        return self._view(p).copy()

    def getincresoln(self, i):
        p = self.L.orc_getincresoln(i)
        assert p
        return self._view(p).copy()

    def increadjproject(self, k, count, IAC, IU, LIU, walls, force=None):
        """One stage k of the incremental adjoint (channel/increadjproject0.c,
        channel/increadjproject.c) on copies of the inputs; returns
        (IAC after the stage, IAU [Nz][5][qpts][Nx/2], IAUb [2][Nz][Nx/2])."""
        IAC = _check_c(np.ascontiguousarray(IAC, dtype=np.complex128)).copy()
        IU = np.ascontiguousarray(IU, dtype=np.complex128)
        LIU = None if LIU is None else np.ascontiguousarray(LIU, dtype=np.complex128)
        walls = np.ascontiguousarray(walls, dtype=np.complex128)
        F = None if force is None else np.ascontiguousarray(force, dtype=np.complex128)
        Nz, _, _, Mx = IAC.shape
        IAU = np.zeros((Nz, 5, self.qpts, Mx), np.complex128)
        IAUb = np.zeros((2, Nz, Mx), np.complex128)
        rc = self.L.orc_increadjproject(k, count, _ptr(IAC), _ptr(IU),
                                        None if LIU is None else _ptr(LIU), _ptr(walls),
                                        None if F is None else _ptr(F), _ptr(IAU), _ptr(IAUb))
        assert rc == 0, rc
        return IAC, IAU, IAUb

    def pass2_incremental(self, U, IU, AU, IAU):
        """The incremental side of pass2 (channel/pass2.c:120-566): returns
        IH [Nz][5][qpts][Nx/2] (H slots 0..2, IS_x, IS_z in slots 3, 4)."""
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (U, IU, AU, IAU)]
        IH = np.zeros(self.ushape, np.complex128)
        rc = self.L.orc_pass2_incremental(*[_ptr(a) for a in arrs], _ptr(IH))
        assert rc == 0, rc
        return IH

    def adjproject_stage(self, k, AC, U, LU):
        """One adjoint stage k (channel/adjproject0.c, channel/adjproject.c)
        on copies; returns (AC after the stage, AU)."""
        AC = _check_c(np.ascontiguousarray(AC, dtype=np.complex128)).copy()
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        LU = None if LU is None else np.ascontiguousarray(LU, dtype=np.complex128)
        AU = np.zeros_like(U)
        rc = self.L.orc_adjproject_stage(k, _ptr(AC), _ptr(U), None if LU is None else _ptr(LU),
                                         _ptr(AU))
        assert rc == 0, rc
        return AC, AU

    def quad(self):
        y = np.zeros(self.qpts)
        w = np.zeros(self.qpts)
        self.L.orc_quad(_ptr(y), _ptr(w))
        return y, w

    def array(self, name):
        out = np.zeros(1 << 20)
        n = self.L.orc_get_array(name.encode(), out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), out.size)
        assert n >= 0, name
        return out[:n].copy()

    def scalar(self, name):
        return self.L.orc_get_scalar(name.encode())

    def inv2d(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128)
        phys = np.zeros((3 * self.Nz // 2, 3 * self.Nx // 2))
        self.L.orc_inv2d(_ptr(spec), _ptr(phys))
        return phys

    def fwd2d(self, phys):
        phys = np.ascontiguousarray(phys, dtype=np.float64)
        spec = np.zeros((self.Nz, self.Nx // 2), complex)
        self.L.orc_fwd2d(_ptr(phys), _ptr(spec))
        return spec

    def pass1(self, U, check_cfl=0):
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        rc = self.L.orc_pass1_state(_ptr(U), check_cfl)
        return rc, U

    def pass2(self, U, AU):
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        AU = np.ascontiguousarray(AU, dtype=np.complex128)
        self.L.orc_pass2_state(_ptr(U), _ptr(AU))
        return U

    def coef_to_gauss(self, C):
        U = np.zeros(self.ushape, complex)
        self.L.orc_coef_to_gauss(_ptr(_check_c(np.ascontiguousarray(C))), _ptr(U))
        return U

    def state(self):
        C = np.zeros(self.cshape, complex)
        U = np.zeros(self.ushape, complex)
        self.L.orc_get_state(_ptr(C), _ptr(U))
        return C, U

    def tangent_forcing(self, count):
        """Re-sensitivity forcing of MC[count]: (f_a [Nz][dimQ][Nx/2],
        f_b [Nz][dimR][Nx/2], f0_a [dimR], f0_b [dimR])."""
        nQ, nR = self.dimQ, self.dimR
        fa = np.zeros((self.Nz, nQ, self.Nx // 2), complex)
        fb = np.zeros((self.Nz, nR, self.Nx // 2), complex)
        f0a = np.zeros(nR, complex)
        f0b = np.zeros(nR, complex)
        assert self.L.orc_tangent_forcing(count, _ptr(fa), _ptr(fb), _ptr(f0a), _ptr(f0b)) == 0
        return fa, fb, f0a, f0b

    def wall_plane(self, C):
        """Wall plane of C (initAlphaBeta2): (Uxb, Uzb), each [Nz][Nx/2]."""
        C = _check_c(np.ascontiguousarray(C))
        ub = np.zeros((self.Nz, self.Nx // 2), complex)
        wb = np.zeros((self.Nz, self.Nx // 2), complex)
        self.L.orc_wall_plane(_ptr(C), _ptr(ub), _ptr(wb))
        return ub, wb

    def incre_boundary(self, Uxb, Uzb):
        """increBoundary on copies of two wall planes; returns (Uxb, Uzb)."""
        ub = np.array(Uxb, dtype=np.complex128, order="C")
        wb = np.array(Uzb, dtype=np.complex128, order="C")
        assert ub.shape == wb.shape == (self.Nz, self.Nx // 2)
        assert self.L.orc_incre_boundary(_ptr(ub), _ptr(wb)) == 0
        return ub, wb

    def flux_trace_start(self, cap=4096):
        """Record the mean-mode flux correction of every later stage."""
        self._ftrace = np.zeros((cap, 5))
        self.L.orc_flux_trace(_ptr(self._ftrace), cap)

    def flux_trace_stop(self):
        """Records {mode, k, target, error before, error after} since
        flux_trace_start (mode 0 primal, 1 tangent, 2 adjoint)."""
        n = self.L.orc_flux_trace(None, 0)
        out = self._ftrace[:n].copy()
        self._ftrace = None
        return out

    def last_status(self):
        return self.L.orc_last_status()

    def destroy(self):
        self.L.orc_destroy()


def band_solve(A, b, dl):
    """Solve the stride-2 banded system (A: N x (2dl+1)) with the oracle."""
    L = load()
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.complex128).copy()
    L.orc_band_solve(_ptr(A), _ptr(b), b.shape[0], dl)
    return b


def glibc_rand_sequence(n, seed=1):
    L = load()
    L.orc_srand(seed)
    return [L.orc_rand() for _ in range(n)]
