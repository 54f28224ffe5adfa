"""ctypes binding of lib/libch4g.so (include/ch4g.h).

There is no fallback: if the HIP library is missing or no GPU is visible the
calls raise.  Arrays are numpy float64, shaped [NZ, NY, NX, ...] like the
reference's interior nodes (x fastest).
"""
import ctypes as C
import os

import numpy as np

from .abi import NPOP, Config, Job, Output, Particle, ParticleConfig, ServeParams, output_dtype, particle_dtype

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libch4g.so")
INCLUDE_DIR = os.path.join(os.path.dirname(PKG_ROOT), "include")

_lib = None

# every function include/ch4g.h declares (checked by tests/test_abi.py)
EXPORTS = [
    "ch4g_last_error", "ch4g_device_count", "ch4g_create", "ch4g_destroy", "ch4g_initial_conditions",
    "ch4g_set_populations", "ch4g_get_populations", "ch4g_step", "ch4g_hydro_fields", "ch4g_dump_averages",
    "ch4g_get_hit_phases", "ch4g_set_hit_phases", "ch4g_dump_rulers", "ch4g_step_timed", "ch4g_steps_per_launch", "ch4g_set_graphs", "ch4g_set_fusion", "ch4g_synchronize", "ch4g_itime", "ch4g_device_bytes",
    "ch4g_footprint_bytes", "ch4g_particles_init", "ch4g_particles_count", "ch4g_get_particles",
    "ch4g_set_particles", "ch4g_step_batch", "ch4g_serve", "ch4g_dump_averages_noalloc", "ch4g_particle_size",
    "ch4g_particle_types",
]


class Ch4gError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Ch4gError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.ch4g_last_error.restype = C.c_char_p
    L.ch4g_device_count.restype = C.c_int
    L.ch4g_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.ch4g_destroy.argtypes = [P]
    L.ch4g_destroy.restype = None
    L.ch4g_initial_conditions.argtypes = [P]
    L.ch4g_set_populations.argtypes = [P, dp, dp]
    L.ch4g_get_populations.argtypes = [P, dp, dp]
    L.ch4g_step.argtypes = [P, C.c_int]
    L.ch4g_hydro_fields.argtypes = [P, dp, dp, dp]
    L.ch4g_dump_averages.argtypes = [P, C.POINTER(Output)]
    L.ch4g_dump_averages_noalloc.argtypes = [P, P, P, P, C.POINTER(Output)]
    L.ch4g_dump_rulers.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
    L.ch4g_get_hit_phases.argtypes = [P, dp]
    L.ch4g_set_hit_phases.argtypes = [P, dp]
    L.ch4g_step_timed.argtypes = [P, C.c_int, C.POINTER(C.c_double)]
    L.ch4g_set_graphs.argtypes = [P, C.c_int]
    L.ch4g_set_fusion.argtypes = [P, C.c_int]
    L.ch4g_steps_per_launch.argtypes = [P, C.POINTER(C.c_int)]
    L.ch4g_synchronize.argtypes = [P]
    L.ch4g_itime.argtypes = [P]
    L.ch4g_itime.restype = C.c_int64
    L.ch4g_device_bytes.argtypes = [P]
    L.ch4g_device_bytes.restype = C.c_int64
    L.ch4g_footprint_bytes.argtypes = [C.POINTER(Config)]
    L.ch4g_footprint_bytes.restype = C.c_int64
    L.ch4g_particles_init.argtypes = [P, C.POINTER(ParticleConfig)]
    L.ch4g_particles_count.argtypes = [P, C.POINTER(C.c_int64)]
    L.ch4g_get_particles.argtypes = [P, P, C.c_int64, C.POINTER(C.c_int64)]
    L.ch4g_particle_types.argtypes = [P, P, C.c_int, C.POINTER(C.c_int)]
    L.ch4g_set_particles.argtypes = [P, P, C.c_int64]
    L.ch4g_step_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int]
    L.ch4g_serve.argtypes = [C.POINTER(ServeParams), C.POINTER(Job), C.c_int32]
    L.ch4g_particle_size.restype = C.c_int64
    if L.ch4g_particle_size() != C.sizeof(Particle):
        raise Ch4gError(f"ch4g_particle is {L.ch4g_particle_size()} bytes in {LIB_PATH}, "
                        f"{C.sizeof(Particle)} in ch4g/abi.py")
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def check(rc):
    if rc != 0:
        raise Ch4gError(f"ch4g error {rc}: {lib().ch4g_last_error().decode()}")


def device_count():
    return lib().ch4g_device_count()


def footprint_bytes(cfg):
    return lib().ch4g_footprint_bytes(C.byref(cfg))


def step_batch(lattices, nsteps):
    """Advance several lattices together (one launch per step for all)."""
    arr = (C.c_void_p * len(lattices))(*[l.handle for l in lattices])
    check(lib().ch4g_step_batch(arr, len(lattices), int(nsteps)))


def serve(jobs, mode=0, hbm_budget=0, chunk_steps=0, max_active=0, use_graphs=True, schedule=0):
    """Run a request stream (list of abi.Job, modified in place with the
    admission/finish times, averages and status); returns the list."""
    arr = (Job * len(jobs))(*jobs)
    sp = ServeParams()
    sp.hbm_budget, sp.mode, sp.chunk_steps, sp.max_active = int(hbm_budget), int(mode), int(chunk_steps), int(max_active)
    sp.use_graphs = 1 if use_graphs else 0
    sp.schedule = int(schedule)
    check(lib().ch4g_serve(C.byref(sp), arr, len(jobs)))
    return list(arr)


class Lattice:
    """A device-resident lattice (one ch4g_lattice)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._L = lib()
        h = C.c_void_p()
        check(self._L.ch4g_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.shape = (cfg.NZ, cfg.NY, cfg.NX)

    def close(self):
        if getattr(self, "_h", None):
            self._L.ch4g_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def initial_conditions(self):
        check(self._L.ch4g_initial_conditions(self._h))

    def set_populations(self, p, g=None):
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        g = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        check(self._L.ch4g_set_populations(self._h, _ptr(p), _ptr(g)))

    def get_populations(self):
        n = self.cfg.nodes()
        p = np.zeros((n, NPOP))
        g = np.zeros((n, NPOP))
        check(self._L.ch4g_get_populations(self._h, _ptr(p), _ptr(g)))
        return p.reshape(self.shape + (NPOP,)), g.reshape(self.shape + (NPOP,))

    def step(self, n=1):
        check(self._L.ch4g_step(self._h, int(n)))

    def step_timed(self, n):
        """Run n steps; return their device time in ms (HIP events on the
        lattice stream)."""
        ms = C.c_double()
        check(self._L.ch4g_step_timed(self._h, int(n), C.byref(ms)))
        return ms.value

    def steps_per_launch(self):
        """Steps one steady-state launch advances (2: two-step LDS kernel)."""
        k = C.c_int()
        check(self._L.ch4g_steps_per_launch(self._h, C.byref(k)))
        return k.value

    def set_graphs(self, enable):
        check(self._L.ch4g_set_graphs(self._h, 1 if enable else 0))

    def set_fusion(self, enable):
        """Two steps per launch on or off (ch4g_set_fusion); results are identical."""
        check(self._L.ch4g_set_fusion(self._h, 1 if enable else 0))

    def synchronize(self):
        check(self._L.ch4g_synchronize(self._h))

    @property
    def itime(self):
        return self._L.ch4g_itime(self._h)

    @property
    def device_bytes(self):
        return self._L.ch4g_device_bytes(self._h)

    def hydro_fields(self):
        n = self.cfg.nodes()
        d = np.zeros(n)
        u = np.zeros((n, 3))
        t = np.zeros(n)
        check(self._L.ch4g_hydro_fields(self._h, _ptr(d), _ptr(u), _ptr(t)))
        s = self.shape
        return {"dens": d.reshape(s), "u": u.reshape(s + (3,)), "t": t.reshape(s)}

    def dump_averages(self):
        out = Output()
        check(self._L.ch4g_dump_averages(self._h, C.byref(out)))
        return out.as_dict()

    def dump_averages_noalloc(self, dens_ptr, u_ptr, t_ptr):
        """ch4g_dump_averages_noalloc: the caller's device buffers (integer
        device addresses, or None) receive the FP64 fields dens[N], u[3N],
        t[N]; the library allocates nothing."""
        out = Output()
        check(self._L.ch4g_dump_averages_noalloc(self._h, dens_ptr, u_ptr, t_ptr, C.byref(out)))
        return out.as_dict()

    def dump_rulers(self):
        """x, y, z rulers (plane averages) as numpy arrays of output records."""
        c = self.cfg
        r = [np.zeros(n, dtype=output_dtype()) for n in (c.NX, c.NY, c.NZ)]
        check(self._L.ch4g_dump_rulers(self._h, *[x.ctypes.data_as(C.c_void_p) for x in r]))
        return tuple(r)

    def hit_phases(self):
        a = np.zeros((6, 3))
        check(self._L.ch4g_get_hit_phases(self._h, _ptr(a)))
        return a

    def set_hit_phases(self, phi):
        a = np.ascontiguousarray(phi, dtype=np.float64).reshape(6, 3)
        check(self._L.ch4g_set_hit_phases(self._h, _ptr(a)))

    # ---- Lagrangian particles
    def particles_init(self, pcfg):
        check(self._L.ch4g_particles_init(self._h, C.byref(pcfg)))

    def particles_count(self):
        n = C.c_int64()
        check(self._L.ch4g_particles_count(self._h, C.byref(n)))
        return n.value

    def particles(self):
        """Particle records in array order (numpy structured array with the
        fields of struct ch4g_particle)."""
        n = self.particles_count()
        a = np.zeros(n, dtype=particle_dtype())
        got = C.c_int64()
        check(self._L.ch4g_get_particles(self._h, a.ctypes.data if n else None, n, C.byref(got)))
        return a[:got.value]

    def particle_types(self):
        """Per-type values of the last particles_init (particle_properties.dat)."""
        from .abi import ParticleType
        n = C.c_int()
        check(self._L.ch4g_particle_types(self._h, None, 0, C.byref(n)))
        out = (ParticleType * max(n.value, 1))()
        check(self._L.ch4g_particle_types(self._h, out, n.value, C.byref(n)))
        return [{f: getattr(out[i], f) for f, _ in ParticleType._fields_} for i in range(n.value)]

    def set_particles(self, a):
        a = np.ascontiguousarray(a, dtype=particle_dtype())
        check(self._L.ch4g_set_particles(self._h, a.ctypes.data if len(a) else None, len(a)))
