"""ctypes front end to the serving layer of libsclbm.so (scl_server_* in
include/sc_lbm_amd.h): independent two-component Shan-Chen jobs kept
resident on the device and advanced together by a persistent kernel.

    from sc_lbm_amd import serve
    with serve.Server(precision="fp32dev", budget_gib=4) as srv:
        jid = srv.submit(nx=256, ny=256, steps=2000, porosity=0.8, geom_seed=1,
                         ic=serve.IC_DROPLET, rho_hi=1.0, rho_lo=0.06, radius=40,
                         tau=(1.0, 1.0), g_fluid=(0.9, 0.9), deliver=serve.DELIVER_RHO)
        st = srv.wait(jid)
        rho_a = srv.result(jid, serve.RHO, 0)      # (256, 256)
"""
import ctypes

import numpy as np

from . import native as nat

IC_UNIFORM, IC_DROPLET, IC_STRATIFIED = 0, 1, 2
DELIVER_RHO, DELIVER_F = 1, 2
QUEUED, RUNNING, DONE, FAILED, CANCELLED, TIMED_OUT = range(6)
STATE_NAMES = ["queued", "running", "done", "failed", "cancelled", "timed_out", "suspended"]
RHO, F = nat.RHO, nat.F
SCL_ERR_BUSY, SCL_ERR_TIMEOUT = 5, 6

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Job(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("mask", _ip), ("porosity", ctypes.c_double), ("rmin", ctypes.c_int), ("rmax", ctypes.c_int),
        ("walls", ctypes.c_int), ("geom_seed", ctypes.c_uint64),
        ("f", _dp * 2), ("ic", ctypes.c_int), ("rho_hi", ctypes.c_double), ("rho_lo", ctypes.c_double),
        ("radius", ctypes.c_double), ("noise", ctypes.c_double), ("ic_seed", ctypes.c_uint64),
        ("tau", ctypes.c_double * 2), ("g_fluid", ctypes.c_double * 2), ("g_solid", ctypes.c_double * 2),
        ("force", _dp * 2), ("steps", ctypes.c_longlong), ("release_ms", ctypes.c_double),
        ("timeout_ms", ctypes.c_double), ("deliver", ctypes.c_int), ("rho0", ctypes.c_double * 2),
    ]


class JobStatus(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int), ("steps_done", ctypes.c_longlong),
        ("release_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
        ("mass", ctypes.c_double * 2), ("mass0", ctypes.c_double * 2), ("bytes", ctypes.c_size_t),
        ("error", ctypes.c_char * 160),
    ]

    def as_dict(self):
        return dict(state=STATE_NAMES[self.state], steps_done=self.steps_done, release_ms=self.release_ms,
                    admit_ms=self.admit_ms, done_ms=self.done_ms, mass=list(self.mass), mass0=list(self.mass0),
                    bytes=self.bytes, error=self.error.decode(errors="replace"))


class Config(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int), ("budget_bytes", ctypes.c_size_t), ("max_jobs", ctypes.c_int),
        ("max_pending", ctypes.c_int), ("round_steps", ctypes.c_int), ("tile_cols", ctypes.c_int),
        ("tile_rows", ctypes.c_int), ("lookahead", ctypes.c_int), ("mass_tol", ctypes.c_double),
        ("max_bypass", ctypes.c_int), ("steps_per_launch", ctypes.c_int), ("band_rows", ctypes.c_int),
        ("use_graphs", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_longlong), ("admitted", ctypes.c_longlong), ("done", ctypes.c_longlong),
        ("failed", ctypes.c_longlong), ("cancelled", ctypes.c_longlong), ("timed_out", ctypes.c_longlong),
        ("rejected_busy", ctypes.c_longlong), ("rejected", ctypes.c_longlong), ("rounds", ctypes.c_longlong),
        ("job_steps", ctypes.c_longlong), ("lattice_updates", ctypes.c_longlong), ("busy_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double), ("setup_ms", ctypes.c_double), ("finish_ms", ctypes.c_double),
        ("clock_ms", ctypes.c_double), ("resident", ctypes.c_int), ("queued", ctypes.c_int),
        ("max_resident", ctypes.c_int), ("arena_bytes", ctypes.c_size_t), ("arena_used", ctypes.c_size_t),
        ("arena_peak", ctypes.c_size_t), ("tile_cols", ctypes.c_int), ("tile_rows", ctypes.c_int),
        ("lookahead", ctypes.c_int), ("round_steps", ctypes.c_int), ("workgroups", ctypes.c_int),
        ("steps_per_launch", ctypes.c_int), ("fluid_updates", ctypes.c_longlong),
        ("pair_updates", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def bind(lib, create="scl_server_create"):
    """Declare the server entry points of `lib` (libsclbm.so, or a library
    that links the same scheduler with another backend and names its
    constructor `create`)."""
    vp = ctypes.c_void_p
    sig = {
        "scl_server_default_config": [ctypes.POINTER(Config)],
        create: [ctypes.POINTER(Config), ctypes.POINTER(vp)],
        "scl_server_submit": [vp, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_longlong)],
        "scl_server_status": [vp, ctypes.c_longlong, ctypes.POINTER(JobStatus)],
        "scl_server_wait": [vp, ctypes.c_longlong, ctypes.c_double, ctypes.POINTER(JobStatus)],
        "scl_server_cancel": [vp, ctypes.c_longlong],
        "scl_server_suspend": [vp, ctypes.c_longlong],
        "scl_server_resume": [vp, ctypes.c_longlong],
        "scl_server_result": [vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_size_t],
        "scl_server_release": [vp, ctypes.c_longlong],
        "scl_server_drain": [vp, ctypes.c_double],
        "scl_server_get_stats": [vp, ctypes.POINTER(Stats)],
        "scl_server_destroy": [vp],
        "scl_job_mask": [ctypes.POINTER(Job), _ip, ctypes.c_size_t],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.scl_last_error.restype = ctypes.c_char_p
    return lib


class ServerError(nat.SclError):
    pass


class Busy(ServerError):
    """Back-pressure: the server's queue is full."""


class WaitTimeout(ServerError):
    pass


def _check(lib, rc):
    if rc == 0:
        return
    msg = lib.scl_last_error().decode(errors="replace")
    if rc == SCL_ERR_BUSY:
        raise Busy(rc, msg)
    if rc == SCL_ERR_TIMEOUT:
        raise WaitTimeout(rc, msg)
    if rc == nat.SCL_ERR_RANGE:
        raise nat.RangeError(rc, msg)
    raise ServerError(rc, msg)


def make_job(nx, ny, steps, mask=None, porosity=1.0, rmin=8, rmax=20, walls=False, geom_seed=0, f=None,
             ic=IC_UNIFORM, rho_hi=1.0, rho_lo=1.0, radius=0.0, noise=0.0, ic_seed=0, tau=(1.0, 1.0),
             g_fluid=(0.0, 0.0), g_solid=(0.0, 0.0), force=(None, None), release_ms=0.0, timeout_ms=0.0,
             deliver=0, rho0=(0.0, 0.0)):
    """An scl_job plus the arrays it points to (kept alive in job._keep)."""
    j = Job()
    keep = []
    j.nx, j.ny = int(nx), int(ny)
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.int32).reshape(-1)
        assert m.size == nx * ny
        keep.append(m)
        j.mask = m.ctypes.data_as(_ip)
    j.porosity, j.rmin, j.rmax, j.walls, j.geom_seed = float(porosity), int(rmin), int(rmax), int(bool(walls)), int(geom_seed)
    if f is not None:
        for c in range(2):
            if f[c] is None:
                continue
            a = np.ascontiguousarray(f[c], dtype=np.float64).reshape(-1)
            assert a.size == 9 * nx * ny
            keep.append(a)
            j.f[c] = a.ctypes.data_as(_dp)
    j.ic, j.rho_hi, j.rho_lo, j.radius, j.noise, j.ic_seed = int(ic), float(rho_hi), float(rho_lo), float(radius), float(noise), int(ic_seed)
    for c in range(2):
        j.tau[c], j.g_fluid[c], j.g_solid[c] = float(tau[c]), float(g_fluid[c]), float(g_solid[c])
        if force[c] is not None:
            a = np.ascontiguousarray(force[c], dtype=np.float64)
            assert a.shape == (9,)
            keep.append(a)
            j.force[c] = a.ctypes.data_as(_dp)
    j.steps, j.release_ms, j.timeout_ms, j.deliver = int(steps), float(release_ms), float(timeout_ms), int(deliver)
    j.rho0[0], j.rho0[1] = float(rho0[0]), float(rho0[1])
    j._keep = keep
    return j


class Server:
    """One serving-layer instance.  Keyword arguments override the fields of
    scl_server_default_config (precision may be given by name)."""

    def __init__(self, lib=None, create="scl_server_create", budget_gib=None, **cfg):
        if lib is None:
            lib = nat.load()
        self.lib = bind(lib, create)
        c = Config()
        _check(self.lib, self.lib.scl_server_default_config(ctypes.byref(c)))
        if "precision" in cfg and isinstance(cfg["precision"], str):
            cfg["precision"] = nat.PRECISION[cfg["precision"]]
        if budget_gib is not None:
            cfg["budget_bytes"] = int(budget_gib * (1 << 30))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.config = c
        self._dims = {}
        h = ctypes.c_void_p()
        _check(self.lib, getattr(self.lib, create)(ctypes.byref(c), ctypes.byref(h)))
        self.h = h

    def submit(self, job=None, **kw):
        j = job if job is not None else make_job(**kw)
        jid = ctypes.c_longlong()
        _check(self.lib, self.lib.scl_server_submit(self.h, ctypes.byref(j), ctypes.byref(jid)))
        self._dims[jid.value] = (j.nx, j.ny)
        return jid.value

    def status(self, jid):
        st = JobStatus()
        _check(self.lib, self.lib.scl_server_status(self.h, jid, ctypes.byref(st)))
        return st.as_dict()

    def wait(self, jid, timeout_ms=-1.0):
        st = JobStatus()
        _check(self.lib, self.lib.scl_server_wait(self.h, jid, float(timeout_ms), ctypes.byref(st)))
        return st.as_dict()

    def cancel(self, jid):
        _check(self.lib, self.lib.scl_server_cancel(self.h, jid))

    def suspend(self, jid):
        """Pre-empt a job: its exact state leaves the device at the next round
        boundary (at once if queued) until resume()."""
        _check(self.lib, self.lib.scl_server_suspend(self.h, jid))

    def resume(self, jid):
        _check(self.lib, self.lib.scl_server_resume(self.h, jid))

    def result(self, jid, which, comp):
        """A deliverable of a finished job: RHO (ny x nx) or F (9 x ny x nx)."""
        nx, ny = self._dims[jid]
        shape = (9, ny, nx) if which == F else (ny, nx)
        out = np.empty(shape, dtype=np.float64)
        _check(self.lib, self.lib.scl_server_result(self.h, jid, which, comp, out.ctypes.data_as(_dp), out.size))
        return out

    def release(self, jid):
        _check(self.lib, self.lib.scl_server_release(self.h, jid))
        self._dims.pop(jid, None)

    def drain(self, timeout_ms=-1.0):
        _check(self.lib, self.lib.scl_server_drain(self.h, float(timeout_ms)))

    def stats(self):
        s = Stats()
        _check(self.lib, self.lib.scl_server_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.scl_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def job_mask(job=None, lib=None, **kw):
    """The mask the server builds for a job (host computation, no device)."""
    lib = bind(lib if lib is not None else nat.load())
    j = job if job is not None else make_job(**kw)
    out = np.empty(j.nx * j.ny, dtype=np.int32)
    _check(lib, lib.scl_job_mask(ctypes.byref(j), out.ctypes.data_as(_ip), out.size))
    return out.reshape(j.ny, j.nx)


def _bind_geometry(lib):
    u64p = ctypes.POINTER(ctypes.c_uint64)
    llp = ctypes.POINTER(ctypes.c_longlong)
    lib.scl_mask_raster.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, llp, ctypes.c_int,
                                    u64p, ctypes.c_size_t, llp, _dp]
    lib.scl_job_mask_bits.argtypes = [ctypes.POINTER(Job), u64p, ctypes.c_size_t, llp, _dp]
    for fn in (lib.scl_mask_raster, lib.scl_job_mask_bits):
        fn.restype = ctypes.c_int
    lib.scl_last_error.restype = ctypes.c_char_p
    return lib


def unpack_bits(bits, nx):
    """A bit mask (ny rows of ceil(nx/64) uint64 words, bit x & 63 of word
    x >> 6 = site x) as ny x nx ints."""
    b = np.ascontiguousarray(bits, dtype=np.uint64)
    by = b.view(np.uint8).reshape(b.shape[0], -1)
    return np.unpackbits(by, axis=1, bitorder="little")[:, :nx].astype(np.int32)


def job_mask_bits(job=None, **kw):
    """A job's mask through the device pipeline (host sampler, device raster,
    pack and count): (bits ny x ceil(nx/64) uint64, fluid sites, (sampler
    ms, device ms))."""
    lib = _bind_geometry(nat.load())
    j = job if job is not None else make_job(**kw)
    out = np.zeros((j.ny, (j.nx + 63) // 64), dtype=np.uint64)
    nf = ctypes.c_longlong(0)
    ms = (ctypes.c_double * 2)()
    _check(lib, lib.scl_job_mask_bits(ctypes.byref(j), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                      out.size, ctypes.byref(nf), ms))
    return out, nf.value, (ms[0], ms[1])


def mask_raster(nx, ny, circles=(), wall_thickness=0, walls_along=0):
    """Geometry::add_walls + add_circle calls rasterised on the device:
    (bits, fluid sites, device ms).  walls_along 0 none, 1 "x", 2 "y";
    circles (D, xc, yc) triples."""
    lib = _bind_geometry(nat.load())
    c = np.ascontiguousarray(np.asarray(circles, dtype=np.int64).reshape(-1))
    out = np.zeros((ny, (nx + 63) // 64), dtype=np.uint64)
    nf = ctypes.c_longlong(0)
    ms = ctypes.c_double(0.0)
    _check(lib, lib.scl_mask_raster(int(nx), int(ny), int(wall_thickness), int(walls_along),
                                    c.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(c) // 3,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), out.size,
                                    ctypes.byref(nf), ctypes.byref(ms)))
    return out, nf.value, ms.value
