"""fluam_amd -- MI355X (gfx950) implementation of fluam's incompressible fluctuating-
hydrodynamics time step (schemeIncompressible, schemeQuasiNeutrallyBuoyant,
schemeStokesLimit) behind a C ABI (include/fhd.h, fluam_amd/lib/libfhd.so).

This module is a thin ctypes layer over that ABI.  It never computes anything itself: if the
HIP library is missing it raises, it does not fall back to a CPU path.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (FhdParams, make_params, SCHEME_INCOMPRESSIBLE, SCHEME_QUASI_NEUTRALLY_BUOYANT,
                   SCHEME_STOKES_LIMIT, RFD_DISPLACED, RFD_REFERENCE)

__all__ = ["FhdParams", "make_params", "Simulation", "lib", "load_datamain", "get_params", "main",
           "scheme_incompressible", "scheme_quasi_neutrally_buoyant", "scheme_stokes_limit",
           "SCHEME_INCOMPRESSIBLE", "SCHEME_QUASI_NEUTRALLY_BUOYANT", "SCHEME_STOKES_LIMIT",
           "RFD_DISPLACED", "RFD_REFERENCE", "LIB_PATH", "Batch", "Group", "bytes_needed", "mem_info",
           "read_fluid_file", "write_fluid_file"]

PKG = os.path.dirname(os.path.abspath(__file__))
# FLUAM_AMD_LIB names another build of the library (for example one installed elsewhere)
LIB_PATH = os.environ.get("FLUAM_AMD_LIB") or os.path.join(PKG, "lib", "libfhd.so")
_lib = None


class FhdError(RuntimeError):
    pass


def lib():
    """Load libfhd.so from the package tree (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FhdError(f"{LIB_PATH} is missing: build it with `make -C fluam_amd/csrc` "
                       "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.fhd_default_params.argtypes = [C.POINTER(FhdParams)]
    L.fhd_default_params.restype = None
    L.fhd_load_datamain.argtypes = [C.c_char_p]
    L.fhd_get_params.argtypes = [C.POINTER(FhdParams)]
    L.fhd_set_params.argtypes = [C.POINTER(FhdParams)]
    for n in ("fhd_scheme_incompressible", "fhd_scheme_quasi_neutrally_buoyant", "fhd_scheme_stokes_limit"):
        getattr(L, n).argtypes = []
    L.fhd_main.argtypes = [C.c_char_p]
    L.fhd_read_fluid_file.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
    L.fhd_last_error.restype = C.c_char_p
    L.fhd_create.argtypes = [C.POINTER(FhdParams)]
    L.fhd_create.restype = vp
    L.fhd_destroy.argtypes = [vp]
    L.fhd_destroy.restype = None
    L.fhd_ctx_error.argtypes = [vp]
    L.fhd_ctx_error.restype = C.c_char_p
    for n in ("fhd_set_velocity", "fhd_get_velocity", "fhd_set_particles", "fhd_get_particles",
              "fhd_get_particle_velocities", "fhd_get_forces", "fhd_max_divergence"):
        getattr(L, n).argtypes = [vp, dp]
    L.fhd_step.argtypes = [vp, C.c_int]
    L.fhd_get_schedule.argtypes = [vp, C.POINTER(C.c_int)]
    L.fhd_sync.argtypes = [vp]
    L.fhd_set_graph.argtypes = [vp, C.c_int]
    L.fhd_set_noise_overlap.argtypes = [vp, C.c_int]
    L.fhd_sk_reset.argtypes = [vp]
    L.fhd_sk_sample.argtypes = [vp]
    L.fhd_sk_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    L.fhd_bytes_needed.argtypes = [C.POINTER(FhdParams)]
    L.fhd_bytes_needed.restype = C.c_longlong
    L.fhd_mem_info.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
    L.fhd_record.argtypes = [vp]
    L.fhd_query.argtypes = [vp]
    L.fhd_batch_create.argtypes = [C.POINTER(FhdParams), C.c_int]
    L.fhd_batch_create.restype = vp
    L.fhd_batch_destroy.argtypes = [vp]
    L.fhd_batch_adopt.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    L.fhd_batch_adopt.restype = vp
    L.fhd_batch_destroy.restype = None
    L.fhd_batch_size.argtypes = [vp]
    L.fhd_batch_replica.argtypes = [vp, C.c_int]
    L.fhd_batch_replica.restype = vp
    L.fhd_batch_error.argtypes = [vp]
    L.fhd_batch_error.restype = C.c_char_p
    L.fhd_batch_step.argtypes = [vp, C.c_int]
    L.fhd_batch_set_mode.argtypes = [vp, C.c_int]
    L.fhd_batch_sync.argtypes = [vp]
    L.fhd_batch_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.fhd_get_step.argtypes = [vp]
    L.fhd_get_step.restype = C.c_longlong
    L.fhd_get_np.argtypes = [vp]
    L.fhd_time_steps.argtypes = [vp, C.c_int, dp]
    L.fhd_profile_kernels.argtypes = [vp, C.c_int, dp, C.c_char_p, C.c_int]
    L.fhd_fft_r2c.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    L.fhd_fft_c2r.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    L.fhd_normals.argtypes = [C.c_ulonglong, C.c_longlong, C.c_int, C.c_int, C.c_int, dp]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _err(ctx=None):
    L = lib()
    msg = L.fhd_ctx_error(ctx) if ctx else L.fhd_last_error()
    return (msg or b"").decode(errors="replace")


def default_params():
    p = FhdParams()
    lib().fhd_default_params(C.byref(p))
    return p


def load_datamain(path):
    """Parse a fluam data.main file into the library's global parameters (loadDataMain)."""
    if not lib().fhd_load_datamain(os.fsencode(path)):
        raise FhdError(_err())
    return get_params()


def get_params():
    p = FhdParams()
    if not lib().fhd_get_params(C.byref(p)):
        raise FhdError(_err())
    return p


def set_params(p):
    lib().fhd_set_params(C.byref(p))


def read_fluid_file(path, cells):
    """Velocity (3N, component-major, x fastest) and density of a .fluidFinalConfiguration file
    for a grid of `cells` = (mx, my, mz); the restart input of fluam's `fluid` keyword."""
    mx, my, mz = cells
    v = np.zeros(3 * mx * my * mz)
    d = C.c_double(0.0)
    if not lib().fhd_read_fluid_file(str(path).encode(), mx, my, mz, _p(v), C.byref(d)):
        raise FhdError(_err())
    return v, d.value


def write_fluid_file(path, cells, box, v, densfluid=1.0, temperature=0.0):
    """Writes a velocity field in fluam's .fluidFinalConfiguration layout
    (src/saveFluidFinalConfiguration.cu:30-75), 17 significant digits so FP32 values survive."""
    mx, my, mz = cells
    N = mx * my * mz
    v = np.asarray(v, dtype=np.float64).reshape(3, N)
    with open(path, "w") as f:
        f.write(f"Number of cells {N}\nmx my mz\n{mx} {my} {mz}\nlx ly lz\n{box[0]} {box[1]} {box[2]}\n")
        f.write(f"Density {densfluid!r}\nPressure parameters 0 0 0\nThermostat 0\nTemperature {temperature}\n")
        f.write("Cells properties \nDensity \nVelocity\n")
        for i in range(N):
            f.write(f"{densfluid!r}\n{v[0, i]:.17g} {v[1, i]:.17g} {v[2, i]:.17g}\n")


def scheme_incompressible():
    return bool(lib().fhd_scheme_incompressible())


def scheme_quasi_neutrally_buoyant():
    return bool(lib().fhd_scheme_quasi_neutrally_buoyant())


def scheme_stokes_limit():
    return bool(lib().fhd_scheme_stokes_limit())


def main(path="data.main"):
    """`fluam data.main`: load the file and run the scheme it selects (src/main.cpp order)."""
    return bool(lib().fhd_main(os.fsencode(path)))


class Simulation:
    """One re-entrant step context (all fields and particles resident on the GPU)."""

    def __init__(self, params: FhdParams, handle=None):
        self.p = params.copy()
        self._owned = handle is None
        self.h = lib().fhd_create(C.byref(self.p)) if handle is None else handle
        if not self.h:
            raise FhdError(_err())
        self.p.np = lib().fhd_get_np(self.h)  # a coordinates file sets the count
        self.N = self.p.mx * self.p.my * self.p.mz

    def close(self):
        if getattr(self, "h", None) and self._owned:
            lib().fhd_destroy(self.h)
        self.h = None

    def record(self):
        """Mark the work enqueued so far; done() reports when it has completed."""
        self._check(lib().fhd_record(self.h))

    def done(self):
        r = lib().fhd_query(self.h)
        if r < 0:
            raise FhdError(_err(self.h))
        return r == 1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, ok):
        if not ok:
            raise FhdError(_err(self.h))

    def step(self, n=1):
        self._check(lib().fhd_step(self.h, int(n)))

    def set_graph(self, on=True):
        """Replay the steady-state step from captured HIP graphs (identical results)."""
        self._check(lib().fhd_set_graph(self.h, 1 if on else 0))

    def sk_reset(self):
        """Clear the structure-factor accumulator."""
        self._check(lib().fhd_sk_reset(self.h))

    def sk_sample(self):
        """Add the current velocity's |v_hat_a(k)|^2 to the structure-factor accumulator (GPU)."""
        self._check(lib().fhd_sk_sample(self.h))

    def structure_factor(self):
        """(S, samples): S[a, z, ky, kx] = rho dV / kT <|v_hat_a(k)|^2> / N over the half
        spectrum (kx = 0..mx/2); the diagonal of the discrete projector at equilibrium."""
        p = self.p
        s = np.zeros((3, p.mz, p.my, p.mx // 2 + 1))
        n = C.c_longlong(0)
        self._check(lib().fhd_sk_read(self.h, s.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
        return s, n.value

    def set_noise_overlap(self, on=True):
        """Accepted for callers of earlier versions; no effect (the Stokes-limit noise is drawn
        in Fourier space inside the solve)."""
        self._check(lib().fhd_set_noise_overlap(self.h, 1 if on else 0))

    def sync(self):
        self._check(lib().fhd_sync(self.h))

    @property
    def nstep(self):
        return lib().fhd_get_step(self.h)

    def velocity(self):
        v = np.zeros(3 * self.N)
        self._check(lib().fhd_get_velocity(self.h, _p(v)))
        return v

    def set_velocity(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._check(lib().fhd_set_velocity(self.h, _p(v)))

    def particles(self):
        q = np.zeros(3 * self.p.np)
        self._check(lib().fhd_get_particles(self.h, _p(q)))
        return q

    def set_particles(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._check(lib().fhd_set_particles(self.h, _p(q)))

    def particle_velocities(self):
        u = np.zeros(3 * self.p.np)
        self._check(lib().fhd_get_particle_velocities(self.h, _p(u)))
        return u

    def forces(self):
        f = np.zeros(3 * self.p.np)
        self._check(lib().fhd_get_forces(self.h, _p(f)))
        return f

    def max_divergence(self):
        d = C.c_double(0)
        self._check(lib().fhd_max_divergence(self.h, C.byref(d)))
        return d.value

    def time_steps(self, n):
        ms = C.c_double(0)
        self._check(lib().fhd_time_steps(self.h, int(n), C.byref(ms)))
        return ms.value

    def schedule(self):
        """The autotuned schedule: z-pass column tile, xy (forward, inverse) and z radix plans,
        W-kernel z chunks."""
        out = (C.c_int * 6)()
        if not lib().fhd_get_schedule(self.h, out):
            raise FhdError(_err(self.h))
        return {"z_column_tile": out[0], "xy_radix_plan": out[1], "z_radix_plan": out[2],
                "w_z_chunk": out[3], "w_z_chunk_batched": out[4], "xy_radix_plan_inverse": out[5]}

    def profile(self, n, nkern=32):
        kms = np.zeros(2 * nkern)
        names = C.create_string_buffer(48 * nkern)
        k = lib().fhd_profile_kernels(self.h, int(n), _p(kms), names, nkern)
        if k <= 0:
            raise FhdError(_err(self.h))
        out = {}
        raw = names.raw
        for j in range(k):
            nm = raw[48 * j:48 * (j + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (kms[2 * j], int(kms[2 * j + 1]))
        return out


def fft_r2c(field):
    """GPU forward real FFT of a (mz, my, mx) field -> (mz, my, mx//2+1) complex (unnormalised)."""
    mz, my, mx = field.shape
    out = np.zeros(mz * my * (mx // 2 + 1) * 2)
    a = np.ascontiguousarray(field, dtype=np.float64)
    if not lib().fhd_fft_r2c(mx, my, mz, _p(a), _p(out)):
        raise FhdError(_err())
    return out.view(np.complex128).reshape(mz, my, mx // 2 + 1)


def fft_c2r(spec, mx):
    """GPU inverse of fft_r2c (unnormalised): half spectrum -> real (mz, my, mx)."""
    mz, my, P = spec.shape
    s = np.ascontiguousarray(spec, dtype=np.complex128).view(np.float64)
    out = np.zeros(mz * my * mx)
    if not lib().fhd_fft_c2r(mx, my, mz, _p(s), _p(out)):
        raise FhdError(_err())
    return out.reshape(mz, my, mx)


def normals(seed, step, stream, sub, n):
    out = np.zeros(4 * n)
    if not lib().fhd_normals(seed, step, stream, sub, n, _p(out)):
        raise FhdError(_err())
    return out.reshape(n, 4)


def mem_info():
    """(free, total) device memory in bytes."""
    f, t = C.c_ulonglong(0), C.c_ulonglong(0)
    if not lib().fhd_mem_info(C.byref(f), C.byref(t)):
        raise FhdError(_err())
    return f.value, t.value


def bytes_needed(params: FhdParams) -> int:
    """Device bytes a Simulation with these parameters allocates."""
    p = params.copy()
    return int(lib().fhd_bytes_needed(C.byref(p)))


class Batch:
    """Independent replicas stepped together: each on its own stream with graph-replayed steps,
    so their kernels overlap on the GPU.  Each replica's results equal those of a Simulation
    with the same parameters run alone."""

    def __init__(self, params_list):
        n = len(params_list)
        arr = (FhdParams * n)(*[p.copy() for p in params_list])
        self.h = lib().fhd_batch_create(arr, n)
        if not self.h:
            raise FhdError(_err())
        self.replicas = [Simulation(params_list[i], handle=lib().fhd_batch_replica(self.h, i)) for i in range(n)]

    def _check(self, ok):
        if not ok:
            raise FhdError(lib().fhd_batch_error(self.h).decode())

    def set_mode(self, mode):
        """0: per-replica graph replay on each replica's stream; 1: one graph for the batch step
        with a branch per replica; 2: merged stages, one batched dispatch per stage and kernel
        instance covering every replica."""
        self._check(lib().fhd_batch_set_mode(self.h, mode))

    def step(self, n=1):
        self._check(lib().fhd_batch_step(self.h, n))

    def sync(self):
        self._check(lib().fhd_batch_sync(self.h))

    def time_steps(self, n):
        ms = C.c_double(0.0)
        self._check(lib().fhd_batch_time_steps(self.h, n, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            for r in self.replicas:
                r.h = None
            lib().fhd_batch_destroy(self.h)
            self.h = None


class Group(Batch):
    """A batch over existing Simulations (fhd_batch_adopt), for a serving loop that regroups its
    running jobs: closing the group leaves the Simulations alive."""

    def __init__(self, sims):
        n = len(sims)
        arr = (C.c_void_p * n)(*[s.h for s in sims])
        self.h = lib().fhd_batch_adopt(arr, n)
        if not self.h:
            raise FhdError(_err())
        self.replicas = list(sims)

    def close(self):
        if getattr(self, "h", None):
            lib().fhd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
