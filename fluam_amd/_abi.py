"""ctypes mirror of include/fhd.h (the C ABI of libfhd.so).

Only plain C types cross the boundary.  The struct layout must match ``fhd_params``
field for field; ``tests/test_abi.py`` checks the size against the library.
"""
import ctypes as C

SCHEME_INCOMPRESSIBLE = 0
SCHEME_QUASI_NEUTRALLY_BUOYANT = 1
SCHEME_STOKES_LIMIT = 2
RFD_DISPLACED = 0
RFD_REFERENCE = 1


class FhdParams(C.Structure):
    _fields_ = [
        ("scheme", C.c_int),
        ("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("densfluid", C.c_double),
        ("shearviscosity", C.c_double),
        ("temperature", C.c_double),
        ("dt", C.c_double),
        ("seed", C.c_ulonglong),
        ("numsteps", C.c_longlong),
        ("numsteps_relaxation", C.c_longlong),
        ("samplefreq", C.c_longlong),
        ("initfluid", C.c_int),
        ("background", C.c_double * 3),
        ("np", C.c_int),
        ("kernel_points", C.c_int),
        ("nonbonded", C.c_int),
        ("cutoff", C.c_double),
        ("wca_sigma", C.c_double),
        ("wca_epsilon", C.c_double),
        ("mass", C.c_double),
        ("volume_particle", C.c_double),
        ("extra_mobility", C.c_double),
        ("rfd_delta", C.c_double),
        ("rfd_mode", C.c_int),
        ("max_part_in_cell", C.c_int),
        ("max_part_in_cell_nonbonded", C.c_int),
        ("initial_positions", C.c_int),
        ("outputname", C.c_char * 512),
        ("coordinates", C.c_char * 512),
        ("fluid", C.c_char * 512),
        ("initial_step", C.c_longlong),
        ("bonded_forces", C.c_char * 512),
        ("structure_factor", C.c_int),
        ("velocities", C.c_char * 512),
    ]

    def copy(self):
        out = FhdParams()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(FhdParams))
        return out

    @property
    def ncells(self):
        return self.mx * self.my * self.mz


def make_params(scheme=SCHEME_INCOMPRESSIBLE, cells=(32, 32, 32), box=None, **kw):
    """Parameters with fluam's defaults (src/loadDataMain.cu:177-299) restated on the Python
    side, overridden by keyword arguments.  ``box`` defaults to dx = 1."""
    p = FhdParams()
    p.scheme = scheme
    p.mx, p.my, p.mz = cells
    if box is None:
        box = tuple(float(m) for m in cells)
    p.lx, p.ly, p.lz = box
    p.densfluid = 1.0
    p.shearviscosity = 1.0
    p.temperature = 1e-3
    p.dt = 0.05
    p.seed = 1
    p.numsteps = 0
    p.numsteps_relaxation = 0
    p.samplefreq = 0
    p.initfluid = 0
    p.np = 0
    p.kernel_points = 3
    p.nonbonded = 1
    p.cutoff = 0.0
    dx = p.lx / p.mx
    p.wca_sigma = 2.0 * dx
    p.wca_epsilon = p.temperature
    p.mass = 0.0
    p.volume_particle = 8.0
    p.extra_mobility = 0.0
    p.rfd_delta = 1e-3
    p.rfd_mode = RFD_DISPLACED
    p.max_part_in_cell = 6
    p.max_part_in_cell_nonbonded = 6
    p.initial_positions = 0
    for k, v in kw.items():
        if k == "background":
            for a in range(3):
                p.background[a] = v[a]
        elif k in ("outputname", "coordinates", "fluid", "bonded_forces", "velocities"):
            setattr(p, k, v.encode() if isinstance(v, str) else v)
        else:
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
    if "wca_epsilon" not in kw:
        p.wca_epsilon = p.temperature
    return p
