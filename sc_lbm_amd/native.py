"""ctypes front end to libsclbm.so (the C-ABI of include/sc_lbm_amd.h).

This is the path every test, smoke check and benchmark calls.  It loads the
in-tree library and raises if it is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SCLBM_LIBRARY selects another build of the library (e.g. to compare builds)
LIB_PATH = os.environ.get("SCLBM_LIBRARY") or os.path.join(HERE, "libsclbm.so")

(SCL_OK, SCL_ERR_ARG, SCL_ERR_HIP, SCL_ERR_RANGE, SCL_ERR_UNSUPPORTED, SCL_ERR_BUSY, SCL_ERR_TIMEOUT,
 SCL_ERR_INEXACT) = range(8)
FP64, FP32_DEV, FP16_DEV = 0, 1, 2
F, FEQ, RHO, UX, UY, UEQX, UEQY, FRX, FRY, FSX, FSY = range(11)
PRECISION = {"fp64": FP64, "fp32dev": FP32_DEV, "fp16dev": FP16_DEV}

# every entry point declared in include/sc_lbm_amd.h
EXPORTS = [
    "scl_last_error", "scl_init", "scl_synchronize", "scl_geometry_create", "scl_geometry_destroy",
    "scl_field_create", "scl_field_destroy", "scl_field_set_interactions", "scl_field_upload",
    "scl_field_download", "scl_compute_density", "scl_compute_velocities",
    "scl_compute_f_equilibrium_geom", "scl_compute_f_equilibrium", "scl_solid_surface_force",
    "scl_fluid_repulsion", "scl_equilibrium_velocities", "scl_collide1", "scl_collide2",
    "scl_volume_force", "scl_stream1", "scl_stream2", "scl_step1", "scl_step2", "scl_moments", "scl_moments_noalloc",
    "scl_set_tile", "scl_set_lookahead", "scl_get_tile", "scl_autotune", "scl_timing", "scl_probe_copy",
    "scl_set_step_fusion",
    "scl_set_fp64_fusion_watch",
    "scl_get_step_fusion",
    "scl_laplace_observables", "scl_contact_angle",
    "scl_field_checkpoint_size", "scl_field_checkpoint", "scl_field_restore",
    # serving layer (bound in serve.py)
    "scl_server_default_config", "scl_server_create", "scl_server_submit", "scl_server_status",
    "scl_server_wait", "scl_server_cancel", "scl_server_suspend", "scl_server_resume", "scl_server_result",
    "scl_server_release", "scl_server_drain",
    "scl_server_get_stats", "scl_server_destroy", "scl_job_mask",
    # the geometry pipeline on the device (bound in serve.py)
    "scl_mask_raster", "scl_job_mask_bits",
]


class SclError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sclbm error %d: %s" % (code, msg))
        self.code = code


class RangeError(SclError, IndexError):
    """The reference would throw std::out_of_range here."""


class InexactError(SclError):
    """A fused step reused an equilibrium velocity older than the previous
    step's at a zero-density site (the data were still delivered)."""


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libsclbm.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.scl_last_error.restype = ctypes.c_char_p
    sig = {
        "scl_init": [ctypes.c_int],
        "scl_synchronize": [],
        "scl_geometry_create": [ctypes.c_int, ctypes.c_int, _ip, ctypes.POINTER(_vp)],
        "scl_geometry_destroy": [_vp],
        "scl_field_create": [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_vp)],
        "scl_field_destroy": [_vp],
        "scl_field_set_interactions": [_vp, ctypes.c_double, ctypes.c_double],
        "scl_field_upload": [_vp, ctypes.c_int, _dp, ctypes.c_size_t],
        "scl_field_download": [_vp, ctypes.c_int, _dp, ctypes.c_size_t],
        "scl_compute_density": [_vp],
        "scl_compute_velocities": [_vp, _vp],
        "scl_compute_f_equilibrium_geom": [_vp, _vp],
        "scl_compute_f_equilibrium": [_vp],
        "scl_solid_surface_force": [_vp, _vp, _vp],
        "scl_fluid_repulsion": [_vp, _vp, _vp],
        "scl_equilibrium_velocities": [_vp, _vp, _vp],
        "scl_collide1": [_vp, _vp],
        "scl_collide2": [_vp, _vp],
        "scl_volume_force": [_vp, _vp, _dp],
        "scl_stream1": [_vp, _vp],
        "scl_stream2": [_vp, _vp, _vp],
        "scl_step1": [_vp, _vp, _dp, ctypes.c_int],
        "scl_step2": [_vp, _vp, _vp, _dp, _dp, ctypes.c_int],
        "scl_moments": [_vp, _vp, _dp],
        "scl_moments_noalloc": [_vp, _vp, _dp],
        "scl_laplace_observables": [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp],
        "scl_contact_angle": [_vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _dp],
        "scl_field_checkpoint_size": [_vp, ctypes.POINTER(ctypes.c_size_t)],
        "scl_field_checkpoint": [_vp, ctypes.c_void_p, ctypes.c_size_t],
        "scl_field_restore": [_vp, ctypes.c_void_p, ctypes.c_size_t],
        "scl_set_tile": [_vp, ctypes.c_int, ctypes.c_int],
        "scl_set_lookahead": [_vp, ctypes.c_int],
        "scl_set_step_fusion": [_vp, ctypes.c_int, ctypes.c_int],
        "scl_set_fp64_fusion_watch": [_vp, ctypes.c_double],
        "scl_get_step_fusion": [_vp, ctypes.c_int, _ip, _ip],
        "scl_get_tile": [_vp, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip],
        "scl_autotune": [_vp, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip],
        "scl_timing": [ctypes.c_int, _dp, _ip],
        "scl_probe_copy": [ctypes.c_size_t, ctypes.c_int, _dp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc):
    if rc != SCL_OK:
        msg = load().scl_last_error().decode(errors="replace")
        if rc == SCL_ERR_RANGE:
            raise RangeError(rc, msg)
        if rc == SCL_ERR_INEXACT:
            raise InexactError(rc, msg)
        raise SclError(rc, msg)


def _arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _force(v):
    if v is None:
        return None
    a = _arr(v)
    assert a.shape == (9,)
    return a


def init(device=0):
    check(load().scl_init(device))


def synchronize():
    check(load().scl_synchronize())


class Geometry:
    def __init__(self, mask):
        lib = load()
        m = np.ascontiguousarray(mask, dtype=np.int32)
        assert m.ndim == 2
        self.ny, self.nx = m.shape
        self.mask = m
        h = _vp()
        check(lib.scl_geometry_create(self.nx, self.ny, m.ctypes.data_as(_ip), ctypes.byref(h)))
        self.h = h

    def set_tile(self, cols, rows):
        check(load().scl_set_tile(self.h, cols, rows))

    def set_lookahead(self, rows_ahead):
        check(load().scl_set_lookahead(self.h, int(rows_ahead)))

    def set_step_fusion(self, steps_per_launch, rows=0):
        """Time steps per launch (0 automatic, 1, 2 or 3) and rows per band of
        the multi-step kernels (0 automatic)."""
        check(load().scl_set_step_fusion(self.h, int(steps_per_launch), int(rows)))

    def set_fp64_fusion_watch(self, rho):
        """Density below which FP64 steps hand over from two per launch to one
        (at least 1e-8)."""
        check(load().scl_set_fp64_fusion_watch(self.h, float(rho)))

    def step_fusion(self, precision=FP32_DEV):
        """(steps per launch, band rows) two-component steps of `precision` use."""
        if isinstance(precision, str):
            precision = PRECISION[precision]
        n, r = ctypes.c_int(), ctypes.c_int()
        check(load().scl_get_step_fusion(self.h, precision, ctypes.byref(n), ctypes.byref(r)))
        return n.value, r.value

    def tile(self, ncomp=2, precision=FP64):
        """(cols, rows, lookahead) the fused kernel uses on this geometry."""
        if isinstance(precision, str):
            precision = PRECISION[precision]
        c, r, la = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(load().scl_get_tile(self.h, ncomp, precision, ctypes.byref(c), ctypes.byref(r), ctypes.byref(la)))
        return c.value, r.value, la.value

    def autotune(self, ncomp=2, precision=FP64):
        """Time every launch shape and keep the fastest; returns it."""
        if isinstance(precision, str):
            precision = PRECISION[precision]
        c, r, la = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(load().scl_autotune(self.h, ncomp, precision, ctypes.byref(c), ctypes.byref(r), ctypes.byref(la)))
        return c.value, r.value, la.value

    def close(self):
        if getattr(self, "h", None):
            load().scl_geometry_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    def __init__(self, geom, tau, precision=FP64, rho0=1.0, g_solid=0.0, g_fluid=0.0):
        lib = load()
        if isinstance(precision, str):
            precision = PRECISION[precision]
        if getattr(geom, "h", None) is None:
            raise SclError(SCL_ERR_ARG, "geometry is closed")
        self.geom = geom
        self.precision = precision
        h = _vp()
        check(lib.scl_field_create(geom.h, precision, float(tau), float(rho0), ctypes.byref(h)))
        self.h = h
        check(lib.scl_field_set_interactions(h, float(g_solid), float(g_fluid)))
        self.n = geom.nx * geom.ny

    def set_interactions(self, g_solid, g_fluid):
        check(load().scl_field_set_interactions(self.h, float(g_solid), float(g_fluid)))

    def _size(self, which):
        return 9 * self.n if which in (F, FEQ) else self.n

    def upload(self, which, values):
        a = _arr(values).ravel()
        check(load().scl_field_upload(self.h, which, a.ctypes.data_as(_dp), a.size))

    def download(self, which=F):
        out = np.empty(self._size(which), dtype=np.float64)
        check(load().scl_field_download(self.h, which, out.ctypes.data_as(_dp), out.size))
        return out

    def set_f(self, f):
        self.upload(F, f)

    def get_f(self):
        return self.download(F)

    def checkpoint(self):
        """Exact checkpoint of the field (populations as stored, their phase,
        the zero-density rule's equilibrium velocities) as bytes."""
        n = ctypes.c_size_t()
        check(load().scl_field_checkpoint_size(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(load().scl_field_checkpoint(self.h, buf, n.value))
        return buf.raw

    def restore(self, data):
        """Continue from a checkpoint of a field of the same size, precision
        and rho0."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        check(load().scl_field_restore(self.h, buf, len(data)))

    def moments(self):
        out = np.zeros(3)
        check(load().scl_moments(self.geom.h, self.h, out.ctypes.data_as(_dp)))
        return out

    def moments_noalloc(self):
        """scl_moments_noalloc: the same sums without allocating device memory."""
        out = np.zeros(3)
        check(load().scl_moments_noalloc(self.geom.h, self.h, out.ctypes.data_as(_dp)))
        return out

    def close(self):
        if getattr(self, "h", None):
            load().scl_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def step1(geom, f, force=None, nsteps=1):
    fa = _force(force)
    check(load().scl_step1(geom.h, f.h, _fp(fa), int(nsteps)))


def step2(geom, a, b, force_a=None, force_b=None, nsteps=1):
    fa, fb = _force(force_a), _force(force_b)
    check(load().scl_step2(geom.h, a.h, b.h, _fp(fa), _fp(fb), int(nsteps)))


def solid_surface_force(geom, a, b):
    check(load().scl_solid_surface_force(geom.h, a.h, b.h))


# per-operator API (reference method names)
def compute_density(f):
    check(load().scl_compute_density(f.h))


def compute_velocities(geom, f):
    check(load().scl_compute_velocities(geom.h, f.h))


def compute_f_equilibrium_geom(geom, f):
    check(load().scl_compute_f_equilibrium_geom(geom.h, f.h))


def compute_f_equilibrium(f):
    check(load().scl_compute_f_equilibrium(f.h))


def fluid_repulsion(geom, a, b):
    check(load().scl_fluid_repulsion(geom.h, a.h, b.h))


def equilibrium_velocities(geom, a, b):
    check(load().scl_equilibrium_velocities(geom.h, a.h, b.h))


def collide1(geom, f):
    check(load().scl_collide1(geom.h, f.h))


def collide2(a, b):
    check(load().scl_collide2(a.h, b.h))


def volume_force(geom, f, force):
    fa = _force(force)
    check(load().scl_volume_force(geom.h, f.h, _fp(fa)))


def stream1(geom, f):
    check(load().scl_stream1(geom.h, f.h))


def stream2(geom, a, b):
    check(load().scl_stream2(geom.h, a.h, b.h))


def timing(reset=0):
    avg, cnt = ctypes.c_double(), ctypes.c_int()
    check(load().scl_timing(int(reset), ctypes.byref(avg), ctypes.byref(cnt)))
    return avg.value, cnt.value


def probe_copy(nbytes, reps=20):
    """Mean milliseconds of one device copy of nbytes (k_probe_copy)."""
    ms = ctypes.c_double()
    check(load().scl_probe_copy(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def gpu_available():
    """True when a gfx950 device answers scl_init (never raises)."""
    try:
        lib = load()
    except ImportError:
        return False
    return lib.scl_init(0) == SCL_OK


def laplace_observables(geom, droplet, bulk, G, p_in, p_out, rho_bulk, rho_disp):
    """(pressure difference, radius) of a droplet, on the device: the
    arithmetic of the reference's laplace_law/compute_surface_tension.py
    (lines 63-70).  p_in / p_out are (row, column) = (y, x) as the script's
    p_tot[p_in[0], p_in[1]]."""
    out = np.zeros(2)
    check(load().scl_laplace_observables(geom.h, droplet.h, bulk.h, float(G), int(p_in[1]), int(p_in[0]),
                                         int(p_out[1]), int(p_out[0]), float(rho_bulk), float(rho_disp),
                                         out.ctypes.data_as(_dp)))
    return float(out[0]), float(out[1])


def contact_angle(geom, droplet, bulk, rho_bulk, rho_disp):
    """(theta in degrees, base length, height, radius) of a droplet on a solid
    below it, on the device: the arithmetic of the reference's
    contact_angles/compute_contact_angles.py (lines 46-100)."""
    out = np.zeros(4)
    check(load().scl_contact_angle(geom.h, droplet.h, bulk.h, float(rho_bulk), float(rho_disp),
                                   out.ctypes.data_as(_dp)))
    return tuple(float(v) for v in out)
