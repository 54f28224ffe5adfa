"""Synthetic workloads of the reference's run types (y is the wall-normal /
vertical axis, as in the reference) and the served-ensemble request stream.

Nothing here computes: these are configurations handed to the C-ABI.
"""
import math

import numpy as np

from . import abi


def rb2d(NX=512, NY=256, NZ=1, Ra=1e8, Pr=1.0, Uff=0.05, precision=abi.FP64, seed=12345, noise=1e-3):
    """Rayleigh-Benard: fluid + temperature, Boussinesq buoyancy, no-slip and
    fixed-T walls at y (T_bot = 1, T_top = 0), linear T plus seeded uniform
    noise.  beta g = Uff^2 / (dT H), nu = Uff H sqrt(Pr/Ra), kappa = nu / Pr."""
    H = float(NY)
    bg = Uff * Uff / H
    nu = Uff * H * (Pr / Ra) ** 0.5
    kappa = nu / Pr
    flags = (abi.LB_FLUID | abi.LB_TEMPERATURE | abi.TEMPERATURE_BUOYANCY | abi.FLUID_BC_Y |
             abi.TEMPERATURE_BC_Y | abi.TEMPERATURE_INITIAL_LINEAR | abi.TEMPERATURE_INITIAL_ADD_NOISE)
    return abi.make_config(NX, NY, NZ, flags, precision=precision, seed=seed, tau_u=3 * nu + 0.5,
                           tau_t=3 * kappa + 0.5, beta_t=bg, gravity_y=1.0, T_bot=1.0, T_top=0.0, T_ref=0.0,
                           noise_t=noise)


def channel3d(NX=64, NY=64, NZ=64, tau=0.6, g=1e-6, precision=abi.FP32_DEV):
    """Force-driven channel: fluid only, no-slip y walls, F_x = g
    (Poiseuille forcing F_x = 2 Amp_x 4 nu / LY^2, forcing.c:369)."""
    nu = (tau - 0.5) / 3.0
    return abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.FLUID_FORCING_POISEUILLE | abi.FLUID_BC_Y,
                           precision=precision, tau_u=tau, Amp_x=g * NY * NY / (8.0 * nu))


# Smallest viscosity a served convection job is given (tau >= 0.5015).  At
# free-fall velocity Uff, nu = Uff H sqrt(Pr/Ra), so a grid of height H
# resolves Ra <= (Uff H / NU_MIN)^2: 1e9 needs H >= 316 at Uff = 0.05.  The
# ensemble draws Ra log-uniform in 1e6..1e9 and caps it there: a job whose
# grid has fewer rows runs at the largest Ra the grid resolves (H = 100:
# 1e8, H = 64: 4.1e7).
NU_MIN = 5e-4


def ensemble(njobs, seed=2026, rate=20.0, steps=(1e3, 1e5), precision=abi.FP32_DEV, Uff=0.05):
    """Served-ensemble request stream: Poisson arrivals at `rate` jobs/s;
    half 2-D Rayleigh-Benard jobs with grids log-uniform from 64x64 to
    1024x512 (NX = 64*16^s, NY = 64*8^s) and Ra log-uniform in 1e6..1e9,
    capped where the grid cannot resolve it with tau >= 0.5015; half 3-D
    channel jobs with cubic grids log-uniform in 64..192 and body force
    1e-6; steps log-uniform in `steps`.  Returns a list of abi.Job."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(njobs):
        t += rng.exponential(1.0 / rate)
        nsteps = int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))))
        if i % 2 == 0:
            s = rng.uniform(0.0, 1.0)
            NX, NY = int(round(64 * 16 ** s)), int(round(64 * 8 ** s))
            Ra = math.exp(rng.uniform(math.log(1e6), math.log(1e9)))
            Ra = min(Ra, (Uff * NY / NU_MIN) ** 2)
            cfg = rb2d(NX, NY, 1, Ra=Ra, precision=precision, seed=int(rng.integers(1, 2 ** 31 - 1)), Uff=Uff)
        else:
            L = int(round(math.exp(rng.uniform(math.log(64), math.log(192)))))
            cfg = channel3d(L, L, L, precision=precision)
        j = abi.Job()
        j.cfg = cfg
        j.nsteps = nsteps
        j.arrival_s = t
        jobs.append(j)
    return jobs


def sweep(njobs, seed=7, rate=200.0, steps=(2e3, 2e4), precision=abi.FP32_DEV):
    """Parameter-sweep request stream: many small 2-D Rayleigh-Benard jobs
    (grids 64x32, 96x48, 128x64, 192x96 or 256x128, Ra log-uniform in
    1e5..1e7, steps log-uniform in `steps`) arriving as a burst (Poisson at
    `rate` jobs/s).  Each job alone is too small to fill the GPU.  Returns a
    list of abi.Job."""
    rng = np.random.default_rng(seed)
    grids = [(64, 32), (96, 48), (128, 64), (192, 96), (256, 128)]
    t = 0.0
    jobs = []
    for _ in range(njobs):
        t += rng.exponential(1.0 / rate)
        NX, NY = grids[int(rng.integers(0, len(grids)))]
        Ra = math.exp(rng.uniform(math.log(1e5), math.log(1e7)))
        nsteps = int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))))
        j = abi.Job()
        j.cfg = rb2d(NX, NY, 1, Ra=Ra, precision=precision, seed=int(rng.integers(1, 2 ** 31 - 1)))
        j.nsteps = nsteps
        j.arrival_s = t
        jobs.append(j)
    return jobs


def job_work(j):
    """Node updates of a job."""
    return int(j.cfg.NX) * int(j.cfg.NY) * int(j.cfg.NZ) * int(j.nsteps)
