"""ch4g — MI355X lattice-Boltzmann time step (host-side Python binding).

The compute path is the HIP library lib/libch4g.so behind include/ch4g.h;
this package only binds it with ctypes (see ch4g.lattice).
"""
from .abi import *  # noqa: F401,F403
from .lattice import Ch4gError, Lattice, device_count, footprint_bytes, lib, serve, step_batch  # noqa: F401,E402
