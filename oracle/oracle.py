"""ctypes wrapper of the FP64 CPU oracle (oracle/ch4_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libch4_oracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [P]
        L.oracle_destroy.argtypes = [P]
        L.oracle_initial_conditions.argtypes = [P]
        L.oracle_set_populations.argtypes = [P, dp, dp]
        L.oracle_get_populations.argtypes = [P, dp, dp]
        L.oracle_step.argtypes = [P, C.c_int]
        L.oracle_itime.argtypes = [P]
        L.oracle_itime.restype = C.c_long
        L.oracle_hydro_fields.argtypes = [P, dp, dp, dp, dp]
        L.oracle_get_hit_phases.argtypes = [P, dp]
        L.oracle_dump_averages.argtypes = [P, P]
        L.oracle_dump_rulers.argtypes = [P, P, P, P, P]
        L.oracle_srand48.argtypes = [P, C.c_long]
        L.oracle_drand48.argtypes = [P]
        L.oracle_drand48.restype = C.c_double
        L.oracle_random_gauss.argtypes = [P, C.c_double, C.c_double]
        L.oracle_random_gauss.restype = C.c_double
        L.oracle_particles_init.argtypes = [P, P]
        L.oracle_particles_count.argtypes = [P]
        L.oracle_particles_count.restype = C.c_long
        L.oracle_get_particles.argtypes = [P, P, C.c_long]
        L.oracle_get_particles.restype = C.c_long
        L.oracle_set_particles.argtypes = [P, P, C.c_long]
        L.oracle_particles_probe.argtypes = [P, dp, dp]
        L.oracle_particles_move.argtypes = [P, dp, dp, C.c_int]
        L.oracle_set_plant.argtypes = [C.c_int]
        L.oracle_set_plant.restype = None
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One reference-structured FP64 lattice on the CPU."""

    def __init__(self, cfg):
        from ch4g.abi import LB_TEMPERATURE  # noqa: F401  (config layout lives in the product header)
        self.cfg = cfg
        self._L = lib()
        self._h = self._L.oracle_create(C.byref(cfg))
        if not self._h:
            raise MemoryError("oracle_create failed")
        self.shape = (cfg.NZ, cfg.NY, cfg.NX)

    def close(self):
        if self._h:
            self._L.oracle_destroy(self._h)
            self._h = None

    __del__ = close

    def initial_conditions(self):
        self._L.oracle_initial_conditions(self._h)

    def set_populations(self, p, g=None):
        p = np.ascontiguousarray(p, dtype=np.float64)
        g = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        self._L.oracle_set_populations(self._h, _ptr(p), _ptr(g))

    def get_populations(self):
        n = self.cfg.nodes()
        p = np.empty((n, 19))
        g = np.empty((n, 19))
        self._L.oracle_get_populations(self._h, _ptr(p), _ptr(g))
        return p.reshape(self.shape + (19,)), g.reshape(self.shape + (19,))

    def step(self, n=1):
        self._L.oracle_step(self._h, int(n))

    @property
    def itime(self):
        return self._L.oracle_itime(self._h)

    def hydro_fields(self):
        n = self.cfg.nodes()
        d = np.empty(n)
        u = np.empty((n, 3))
        t = np.empty(n)
        f = np.empty((n, 3))
        self._L.oracle_hydro_fields(self._h, _ptr(d), _ptr(u), _ptr(t), _ptr(f))
        s = self.shape
        return {"dens": d.reshape(s), "u": u.reshape(s + (3,)), "t": t.reshape(s),
                "force": f.reshape(s + (3,))}

    def hit_phases(self):
        a = np.empty((6, 3))
        self._L.oracle_get_hit_phases(self._h, _ptr(a))
        return a

    def particles_init(self, pcfg):
        if self._L.oracle_particles_init(self._h, C.byref(pcfg)) != 0:
            raise MemoryError("oracle_particles_init failed")

    def particles(self):
        """Particle records in array order, numpy structured array."""
        from ch4g.abi import particle_dtype
        n = self._L.oracle_particles_count(self._h)
        a = np.zeros(n, dtype=particle_dtype())
        self._L.oracle_get_particles(self._h, a.ctypes.data, n)
        return a

    def set_particles(self, a):
        from ch4g.abi import particle_dtype
        a = np.ascontiguousarray(a, dtype=particle_dtype())
        self._L.oracle_set_particles(self._h, a.ctypes.data, len(a))

    def probe(self, u=None, t=None):
        """Interpolate given interior fields u [NZ,NY,NX,3], t [NZ,NY,NX] at the particles."""
        u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        t = None if t is None else np.ascontiguousarray(t, dtype=np.float64)
        self._L.oracle_particles_probe(self._h, _ptr(u), _ptr(t))
        return self.particles()

    def move(self, u=None, t=None, first=False):
        """One particle sub-step on given interior fields held fixed (no fluid step)."""
        u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        t = None if t is None else np.ascontiguousarray(t, dtype=np.float64)
        self._L.oracle_particles_move(self._h, _ptr(u), _ptr(t), int(first))
        return self.particles()

    def dump_averages(self):
        from ch4g.abi import Output
        out = Output()
        self._L.oracle_dump_averages(self._h, C.byref(out))
        return out.as_dict()

    def dump_rulers(self):
        """(global averages dict, x ruler [NX], y ruler [NY], z ruler [NZ]);
        rulers are numpy structured arrays with the output members."""
        from ch4g.abi import Output, output_dtype
        out = Output()
        r = [np.zeros(n, dtype=output_dtype()) for n in (self.cfg.NX, self.cfg.NY, self.cfg.NZ)]
        self._L.oracle_dump_rulers(self._h, C.byref(out), *[x.ctypes.data_as(C.c_void_p) for x in r])
        return (out.as_dict(),) + tuple(r)


# mistakes that oracle_set_plant() can switch in (ch4_oracle.c, PLANT_*)
PLANT = {"hit_ax_in_fx": 1, "hit_no_reversion": 2, "hit_exponent": 3, "buoy_t_bot": 4, "buoy_sign": 5,
         "channel_lx": 6, "hit_noise": 7}


class planted:
    """Context manager: the oracle with one planted mistake (process-wide)."""

    def __init__(self, name):
        self.code = PLANT[name]

    def __enter__(self):
        lib().oracle_set_plant(self.code)
        return self

    def __exit__(self, *a):
        lib().oracle_set_plant(0)


class RNG:
    """The oracle's drand48 / random_gauss restatement, for pinning tests."""

    def __init__(self, seed):
        from ch4g.abi import make_config
        self._o = Oracle(make_config(1, 1, 1, 0, seed=seed))
        self._L = self._o._L

    def drand48(self):
        return self._L.oracle_drand48(self._o._h)

    def random_gauss(self, mu=0.0, sigma=1.0):
        return self._L.oracle_random_gauss(self._o._h, mu, sigma)
