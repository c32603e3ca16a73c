"""ctypes front end to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the compiled reference (oracle/_ref/libscref.so).

ORACLE / TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module, and only as the checker.

Both libraries expose the same system model, so a test can run one case
through either and compare the raw doubles.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libscref.so")

F, FEQ, RHO, UX, UY, UEQX, UEQY, FRX, FRY, FSX, FSY = range(11)
FIELD_SIZE9 = {F, FEQ}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_szp = ctypes.POINTER(ctypes.c_size_t)


def _d(a):
    return a.ctypes.data_as(_dp)


def _opt9(v):
    if v is None:
        return None
    a = np.ascontiguousarray(v, dtype=np.float64)
    assert a.shape == (9,)
    return a


_lib_cache = {}


def _load(path, prefix):
    if path in _lib_cache:
        return _lib_cache[path]
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    lib = ctypes.CDLL(path)
    create = getattr(lib, prefix + "create")
    create.restype = ctypes.c_void_p
    create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp]
    getattr(lib, prefix + "destroy").argtypes = [ctypes.c_void_p]
    for nm in ("set", "get"):
        fn = getattr(lib, prefix + nm)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_size_t]
        fn.restype = ctypes.c_int
    fn = getattr(lib, prefix + "run")
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    fn = getattr(lib, prefix + "solid_surface_force")
    fn.argtypes = [ctypes.c_void_p]
    for nm in ("init_droplet", "init_rectangle"):
        fn = getattr(lib, prefix + nm)
        fn.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * (7 if nm == "init_droplet" else 8)
    _lib_cache[path] = lib
    return lib


def oracle_lib():
    lib = _load(ORACLE_SO, "oracle_")
    if not hasattr(lib, "_extra"):
        for nm in ("density", "velocities", "feq_geom", "feq_mc"):
            getattr(lib, "oracle_" + nm).argtypes = [ctypes.c_void_p, ctypes.c_int]
        for nm in ("repulsion", "eq_velocities", "collide1", "collide2", "stream"):
            getattr(lib, "oracle_" + nm).argtypes = [ctypes.c_void_p]
        lib.oracle_volume_force.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        lib.oracle_moments.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        lib.oracle_add_walls.argtypes = [_ip, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int]
        lib.oracle_add_circle.argtypes = [_ip, ctypes.c_int, ctypes.c_int] + [ctypes.c_size_t] * 3
        lib.oracle_add_rectangle.argtypes = [_ip, ctypes.c_int, ctypes.c_int] + [ctypes.c_size_t] * 4
        lib._extra = True
    return lib


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    lib = _load(REF_SO, "ref_")
    if not hasattr(lib, "_extra"):
        lib.ref_geometry.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_int, _szp, _szp, _szp, _ip]
        lib.ref_add_rectangle.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_size_t] * 4 + [_ip]
        lib.ref_write_f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        lib.ref_geometry_object.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                            _szp, _szp, ctypes.c_size_t, ctypes.c_size_t, _ip]
        lib._extra = True
    return lib


class System:
    """One lattice with one or two components, driven through either the
    oracle (kind="oracle") or the compiled reference (kind="ref")."""

    def __init__(self, mask, tau, g_fluid=None, g_solid=None, kind="oracle"):
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        assert mask.ndim == 2
        self.ny, self.nx = mask.shape
        self.n = self.nx * self.ny
        self.ncomp = len(tau)
        self.kind = kind
        self.lib = oracle_lib() if kind == "oracle" else ref_lib()
        self.p = "oracle_" if kind == "oracle" else "ref_"
        self.mask = mask
        t = np.ascontiguousarray(tau, dtype=np.float64)
        gf = np.ascontiguousarray(g_fluid if g_fluid is not None else [0.0] * self.ncomp, dtype=np.float64)
        gs = np.ascontiguousarray(g_solid if g_solid is not None else [0.0] * self.ncomp, dtype=np.float64)
        self.h = getattr(self.lib, self.p + "create")(self.nx, self.ny, self.ncomp,
                                                     mask.ctypes.data_as(_ip), _d(t), _d(gf), _d(gs))
        if not self.h:
            raise ValueError("create failed")

    def close(self):
        if self.h:
            getattr(self.lib, self.p + "destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _size(self, field):
        return 9 * self.n if field in FIELD_SIZE9 else self.n

    def set(self, comp, field, values):
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        rc = getattr(self.lib, self.p + "set")(self.h, comp, field, _d(a), a.size)
        if rc:
            raise ValueError("set failed")

    def get(self, comp, field):
        out = np.empty(self._size(field), dtype=np.float64)
        rc = getattr(self.lib, self.p + "get")(self.h, comp, field, _d(out), out.size)
        if rc:
            raise ValueError("get failed (field %d)" % field)
        return out

    def set_f(self, comp, f):
        self.set(comp, F, f)

    def get_f(self, comp):
        return self.get(comp, F)

    def solid_surface_force(self):
        getattr(self.lib, self.p + "solid_surface_force")(self.h)

    def run(self, nsteps, force0=None, force1=None):
        f0, f1 = _opt9(force0), _opt9(force1)
        rc = getattr(self.lib, self.p + "run")(self.h, nsteps,
                                               _d(f0) if f0 is not None else None,
                                               _d(f1) if f1 is not None else None)
        if self.kind == "ref" and rc:
            raise RuntimeError("reference run raised")

    def init_droplet(self, *args):
        getattr(self.lib, self.p + "init_droplet")(self.h, *[float(a) for a in args])

    def init_rectangle(self, *args):
        return getattr(self.lib, self.p + "init_rectangle")(self.h, *[float(a) for a in args])

    def moments(self, comp):
        assert self.kind == "oracle"
        out = np.zeros(3)
        self.lib.oracle_moments(self.h, comp, _d(out))
        return out


def add_walls(mask, dh, where):
    m = np.ascontiguousarray(mask, dtype=np.int32).copy()
    ny, nx = m.shape
    rc = oracle_lib().oracle_add_walls(m.ctypes.data_as(_ip), nx, ny, dh, 1 if where == "x" else 0)
    if rc:
        raise RuntimeError("wall thickness larger than domain")
    return m


def add_circle(mask, d, xc, yc):
    m = np.ascontiguousarray(mask, dtype=np.int32).copy()
    ny, nx = m.shape
    rc = oracle_lib().oracle_add_circle(m.ctypes.data_as(_ip), nx, ny, d, xc, yc)
    if rc:
        raise RuntimeError("circle does not fit the domain")
    return m


def add_rectangle(mask, lx, ly, xc, yc):
    m = np.ascontiguousarray(mask, dtype=np.int32).copy()
    ny, nx = m.shape
    rc = oracle_lib().oracle_add_rectangle(m.ctypes.data_as(_ip), nx, ny, lx, ly, xc, yc)
    if rc:
        raise RuntimeError("rectangle does not fit the domain")
    return m


def ref_geometry_object(nx, ny, kind, props, obj_type="", bounds=(0, 0, 0, 0), u=0, v=0):
    """Mask built by the compiled reference's Geometry from one object call:
    kind 0 add_ellipse, 1 add_square, 2 add_array by gaps (u, v), 3 add_array
    by numbers of objects (u, v).  Returns (status, mask); status 0 ok,
    1 runtime_error, 2 invalid_argument, 3 out_of_range, 4 other."""
    lib = ref_lib()
    p = np.array(list(props) or [0], dtype=np.uint64)
    b = np.array(bounds, dtype=np.uint64)
    out = np.zeros(nx * ny, dtype=np.int32)
    rc = lib.ref_geometry_object(nx, ny, kind, obj_type.encode(), len(props), p.ctypes.data_as(_szp),
                                 b.ctypes.data_as(_szp), u, v, out.ctypes.data_as(_ip))
    return rc, out.reshape(ny, nx)


def ref_geometry(nx, ny, dh=0, where="x", circles=()):
    """Mask built by the compiled reference's Geometry (add_walls + add_circle)."""
    lib = ref_lib()
    k = len(circles)
    d = np.array([c[0] for c in circles] or [0], dtype=np.uint64)
    xc = np.array([c[1] for c in circles] or [0], dtype=np.uint64)
    yc = np.array([c[2] for c in circles] or [0], dtype=np.uint64)
    out = np.zeros(nx * ny, dtype=np.int32)
    rc = lib.ref_geometry(nx, ny, dh, 1 if where == "x" else 0, k,
                          d.ctypes.data_as(_szp), xc.ctypes.data_as(_szp), yc.ctypes.data_as(_szp),
                          out.ctypes.data_as(_ip))
    return rc, out.reshape(ny, nx)


def body_force(g, axis="x"):
    """The reference drivers' per-direction volume force for a body force g:
    (g/6)*c_axis (benchmarks/.../laminar_flow_x_dir.cpp:21-24)."""
    c = [0, 1, 0, -1, 0, 1, -1, -1, 1] if axis == "x" else [0, 0, 1, 0, -1, 1, 1, -1, -1]
    beta = 1.0 / 6.0
    return np.array([float(v) * (g * beta) for v in c])
