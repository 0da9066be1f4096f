"""ctypes binding of the FP64 oracle (oracle/_build/libfhd_oracle.so) -- test infrastructure."""
import ctypes as C
import os
import subprocess

import numpy as np

from fluam_amd._abi import FhdParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "libfhd_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(FhdParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        for name in ("orc_set_velocity", "orc_get_velocity", "orc_set_particles", "orc_get_particles",
                     "orc_get_particle_velocities", "orc_get_forces", "orc_divergence"):
            getattr(L, name).argtypes = [C.c_void_p, dp]
        for name in ("orc_solve_incompressible", "orc_solve_stokes", "orc_project", "orc_wca", "orc_bonded"):
            getattr(L, name).argtypes = [C.c_void_p, dp, dp]
        L.orc_interp.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_spread.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_spread_drift.argtypes = [C.c_void_p, dp, C.c_double, dp]
        L.orc_spread_quirk.argtypes = [C.c_void_p, dp, dp, dp, C.c_double, dp]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_step.restype = C.c_longlong
        L.orc_get_step.argtypes = [C.c_void_p]
        L.orc_get_np.argtypes = [C.c_void_p]
        L.orc_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_normals4.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_uint32, dp]
        L.orc_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int]
        for name in ("orc_delta3", "orc_delta4", "orc_ddelta3", "orc_ddelta4"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        L.orc_initial_positions.argtypes = [C.POINTER(FhdParams), dp]
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def normals4(seed, step, stream, sub, index):
    o = np.zeros(4)
    lib().orc_normals4(seed, step, stream, sub, index, _p(o))
    return o


def fft3d(a, sign=-1):
    """a: complex array shaped (mz, my, mx); returns unnormalised transform."""
    mz, my, mx = a.shape
    buf = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().orc_fft3d(mx, my, mz, buf.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), sign)
    return buf


def delta3(r):
    return lib().orc_delta3(float(r))


def delta4(r):
    return lib().orc_delta4(float(r))


def ddelta3(r):
    return lib().orc_ddelta3(float(r))


def ddelta4(r):
    return lib().orc_ddelta4(float(r))


def initial_positions(params):
    q = np.zeros(3 * params.np)
    lib().orc_initial_positions(C.byref(params), _p(q))
    return q


class Oracle:
    def __init__(self, params: FhdParams):
        self.p = params.copy()
        self.h = lib().orc_create(C.byref(self.p))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.p.np = lib().orc_get_np(self.h)  # a coordinates file sets the count
        self.N = self.p.mx * self.p.my * self.p.mz

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1):
        assert lib().orc_step(self.h, n) == 1

    @property
    def nstep(self):
        return lib().orc_get_step(self.h)

    def velocity(self):
        v = np.zeros(3 * self.N)
        lib().orc_get_velocity(self.h, _p(v))
        return v

    def set_velocity(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_set_velocity(self.h, _p(v))

    def particles(self):
        q = np.zeros(3 * self.p.np)
        lib().orc_get_particles(self.h, _p(q))
        return q

    def set_particles(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().orc_set_particles(self.h, _p(q))

    def particle_velocities(self):
        u = np.zeros(3 * self.p.np)
        lib().orc_get_particle_velocities(self.h, _p(u))
        return u

    def forces(self):
        f = np.zeros(3 * self.p.np)
        lib().orc_get_forces(self.h, _p(f))
        return f

    def wca(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        f = np.zeros_like(q)
        lib().orc_wca(self.h, _p(q), _p(f))
        return f

    def bonded(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        f = np.zeros_like(q)
        lib().orc_bonded(self.h, _p(q), _p(f))
        return f

    def spread(self, q, F, field=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64)
        out = np.zeros(3 * self.N) if field is None else np.ascontiguousarray(field, dtype=np.float64).copy()
        lib().orc_spread(self.h, _p(q), _p(F), _p(out))
        return out

    def spread_drift(self, q, scale):
        """field_c = scale * sum_p d/dq_c [phi phi phi] (thermal-drift spread)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(3 * self.N)
        lib().orc_spread_drift(self.h, _p(q), float(scale), _p(out))
        return out

    def spread_quirk(self, q, qd, F, scale):
        """scale * sum_p F_p phi(q - node) over the nodes the displaced qd selects."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        qd = np.ascontiguousarray(qd, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64)
        out = np.zeros(3 * self.N)
        lib().orc_spread_quirk(self.h, _p(q), _p(qd), _p(F), float(scale), _p(out))
        return out

    def interp(self, q, field):
        q = np.ascontiguousarray(q, dtype=np.float64)
        field = np.ascontiguousarray(field, dtype=np.float64)
        u = np.zeros(3 * self.p.np)
        lib().orc_interp(self.h, _p(q), _p(field), _p(u))
        return u

    def divergence(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        d = np.zeros(self.N)
        lib().orc_divergence(self.h, _p(v), _p(d))
        return d

    def solve_incompressible(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64)
        v = np.zeros_like(W)
        lib().orc_solve_incompressible(self.h, _p(W), _p(v))
        return v

    def project(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64)
        v = np.zeros_like(W)
        lib().orc_project(self.h, _p(W), _p(v))
        return v

    def solve_stokes(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64)
        v = np.zeros_like(W)
        lib().orc_solve_stokes(self.h, _p(W), _p(v))
        return v


def _bind_terms():
    L = lib()
    dp = C.POINTER(C.c_double)
    L.orc_laplacian_term.argtypes = [C.c_void_p, dp, dp]
    L.orc_advection_term.argtypes = [C.c_void_p, dp, dp]
    L.orc_noise_divergence.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_int, dp]
    L.orc_fluid_noise.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int, dp]
    return L


def laplacian_term(o, v):
    out = np.zeros(3 * o.N)
    _bind_terms().orc_laplacian_term(o.h, _p(np.ascontiguousarray(v, dtype=np.float64)), _p(out))
    return out


def advection_term(o, v):
    out = np.zeros(3 * o.N)
    _bind_terms().orc_advection_term(o.h, _p(np.ascontiguousarray(v, dtype=np.float64)), _p(out))
    return out


def noise_divergence(o, Z6, fact1, fact4, trace_free):
    out = np.zeros(3 * o.N)
    _bind_terms().orc_noise_divergence(o.h, _p(np.ascontiguousarray(Z6, dtype=np.float64)), fact1, fact4,
                                       int(trace_free), _p(out))
    return out


def kspace_stress(o, step, nset, f1, f4):
    """The Stokes step's stress drawn in Fourier space: (sigma6, w3, max imaginary part)."""
    L = lib()
    dp = C.POINTER(C.c_double)
    L.orc_kspace_stress.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_double, dp, dp, dp]
    sig = np.zeros(6 * o.N)
    w = np.zeros(3 * o.N)
    mi = C.c_double(0.0)
    L.orc_kspace_stress(o.h, step, nset, f1, f4, _p(sig), _p(w), C.byref(mi))
    return sig, w, mi.value


def kspace_noise(o, step, nset, f4):
    """The Stokes step's noise source as the solve keeps it, drawn in Fourier space: (3 x N real
    space values, largest imaginary part of the inverse transform)."""
    L = lib()
    dp = C.POINTER(C.c_double)
    L.orc_kspace_noise.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_double, dp, dp]
    u = np.zeros(3 * o.N)
    mi = C.c_double(0.0)
    L.orc_kspace_noise(o.h, step, nset, f4, _p(u), C.byref(mi))
    return u, mi.value


def fluid_noise(seed, step, stream, N):
    Z = np.zeros(6 * N)
    _bind_terms().orc_fluid_noise(seed, step, stream, N, _p(Z))
    return Z
