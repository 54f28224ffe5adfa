"""Named workloads of the reference's streaming step (sizes in lattice units).

y is the wall-normal / vertical axis (the reference's convention).
"""
from ch4g import abi


def poiseuille(NX=4, NY=64, NZ=4, tau=1.0, g=1e-6, precision=abi.FP64):
    """Force-driven channel, halfway bounce-back walls at y, F_x = g.

    Poiseuille forcing builds F_x = 2 Amp_x 4 nu / LY^2 (forcing.c:369), so
    Amp_x = g LY^2 / (8 nu) gives F_x = g."""
    nu = (tau - 0.5) / 3.0
    return abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.FLUID_FORCING_POISEUILLE | abi.FLUID_BC_Y,
                           precision=precision, tau_u=tau, Amp_x=g * NY * NY / (8.0 * nu))


def rb2d(NX=512, NY=256, NZ=1, Ra=1e8, Pr=1.0, Uff=0.05, precision=abi.FP64, seed=12345, noise=1e-3):
    """Rayleigh-Benard: fluid + temperature, buoyancy, no-slip and fixed-T
    walls at y (T_bot = 1, T_top = 0), linear T plus seeded uniform noise.

    beta g = Uff^2 / (dT H), nu = Uff H sqrt(Pr/Ra), kappa = nu / Pr."""
    H = float(NY)
    dT = 1.0
    bg = Uff * Uff / (dT * H)
    nu = Uff * H * (Pr / Ra) ** 0.5
    kappa = nu / Pr
    flags = (abi.LB_FLUID | abi.LB_TEMPERATURE | abi.TEMPERATURE_BUOYANCY | abi.FLUID_BC_Y |
             abi.TEMPERATURE_BC_Y | abi.TEMPERATURE_INITIAL_LINEAR | abi.TEMPERATURE_INITIAL_ADD_NOISE)
    return abi.make_config(NX, NY, NZ, flags, precision=precision, seed=seed, tau_u=3 * nu + 0.5,
                           tau_t=3 * kappa + 0.5, beta_t=bg, gravity_y=1.0, T_bot=1.0, T_top=0.0, T_ref=0.0,
                           noise_t=noise)


def hit(N=32, tau=0.5015, amp=1e-5, les_cs=0.1, precision=abi.FP64, seed=4321, les=True):
    """Forced homogeneous isotropic turbulence, periodic, 6 modes |k|^2 <= 2,
    optional Smagorinsky closure: les=True or "local" for the local-strain
    form, "fd" for the finite-difference form (the reference's default)."""
    flags = abi.LB_FLUID | abi.FLUID_FORCING_HIT
    if les:
        flags |= abi.FLUID_LES
        if les != "fd":
            flags |= abi.FLUID_LES_LOCAL_STRAIN
    return abi.make_config(N, N, N, flags, precision=precision, seed=seed, tau_u=tau, Amp_x=amp, Amp_y=amp,
                           Amp_z=amp, les_cs=les_cs)


def channel3d(NX=64, NY=64, NZ=64, tau=0.6, g=1e-6, precision=abi.FP32_DEV):
    """3-D channel: fluid only, no-slip y walls, constant body force g."""
    return poiseuille(NX, NY, NZ, tau=tau, g=g, precision=precision)
