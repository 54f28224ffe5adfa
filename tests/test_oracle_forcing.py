"""Pins of the oracle's build_forcing branches (CPU only).

Each pin checks a consequence that the mathematics or physics fixes, not the
formula itself, and each is shown to fail on a plausible mistake planted in
the oracle (oracle_set_plant, test-only):

* HIT force (forcing.c:576-590, modes initial_conditions.c:842-849):
  F_x depends on (y, z) only, F_y on (x, z), F_z on (x, y) (divergence free);
  its 2-D Fourier transform holds only the modes |k|^2 <= 2 of the forced
  shells, with amplitude |k|^(-2/3) Amp and phase 2 pi (phi + k/(2 L)) (cell
  centres at i + 1/2).  Planted: sin(k_x x) in F_x; pow(k^2, -2/3).
* HIT phase walk (forcing.c:47-69): an Ornstein-Uhlenbeck process,
  phi_{n+1} = (1 - a) phi_n + sqrt(2 D) xi with a = 0.1 / LX (dt = 1) and
  D = 0.001 / LX: regression slope 1 - a, residual variance 2 D, stationary
  variance 2 D / (2 a - a^2).  Planted: no mean reversion; noise without
  its factor 2.
* Buoyancy (forcing.c:620-667, theta = t - (T_bot + T_top)/2 by default):
  the conduction state relaxes to rest with the hydrostatic density
  rho(y) ~ exp(beta g / cs^2 * int (T - T0) dy), the barometric profile of
  initial_conditions.c:72-84 (peaked at mid-height for T0 = the mean wall
  temperature); the conduction state of a no-slip, fixed-T layer is
  linearly stable at Ra = 1500 and unstable at Ra = 2000 (Ra_c = 1707.8).
  Planted: theta = t - T_bot; a flipped sign.
* Channel forcing (forcing.c:374-379): the steady flow between no-slip
  walls is the exact lattice Poiseuille solution for g = 2 Amp_x^2 / LY;
  in a periodic box each forcing component accelerates the mean flow at
  its own rate, <u>(n) = (n + 1/2) F.  Planted: LX in place of LY.
"""
import math

import numpy as np
import pytest

from ch4g import abi
from oracle import Oracle, planted
import configs


def hit_cfg(NX, NY, NZ, amp=(1e-3, 2e-3, 3e-3), seed=4321):
    return abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.FLUID_FORCING_HIT, seed=seed, tau_u=0.8, Amp_x=amp[0],
                           Amp_y=amp[1], Amp_z=amp[2])


MODES = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1]])


def hit_force_step1(cfg):
    o = Oracle(cfg)
    o.initial_conditions()
    o.step(1)
    return o.hydro_fields()["force"], o.hit_phases()


def hit_structure_and_spectrum_ok(cfg):
    """True when the step-1 HIT force has the structure and spectrum of the
    forced shells (see the module docstring)."""
    F, phi = hit_force_step1(cfg)
    NZ, NY, NX = F.shape[:3]
    amp = (cfg.Amp_x, cfg.Amp_y, cfg.Amp_z)
    fac = (MODES ** 2).sum(1) ** (-1.0 / 3.0)  # |k|^(-2/3)
    L = (NX, NY, NZ)
    # component c does not depend on coordinate c (array axis 2 - c)
    for c in range(3):
        Fc = F[..., c]
        free = 2 - c
        if np.ptp(Fc, axis=free).max() != 0.0:
            return False
        plane = np.take(Fc, 0, axis=free)  # 2-D array over the other two array axes, in order
        rest = [ax for ax in range(3) if ax != free]
        spec = np.fft.fft2(plane) / plane.size
        # expected coefficients: for each dependent coordinate d, the modes
        # with k_d = 1 contribute sin(2 pi ((i + 1/2)/L_d + phi_d)), the
        # others the constant sin(2 pi phi_d)
        exp = np.zeros_like(spec)
        for pos, ax in enumerate(rest):
            d = 2 - ax  # array axis -> coordinate (x = 0, y = 1, z = 2)
            for ii in range(6):
                s = fac[ii] * amp[c]
                if MODES[ii, d] == 1:
                    ph = 2 * math.pi * (phi[ii, d] + 0.5 / L[d])
                    idx = [0, 0]
                    idx[pos] = 1
                    exp[tuple(idx)] += s * np.exp(1j * ph) / 2j
                    idx[pos] = -1
                    exp[tuple(idx)] -= s * np.exp(-1j * ph) / 2j
                else:
                    exp[0, 0] += s * math.sin(2 * math.pi * phi[ii, d])
        if np.abs(spec - exp).max() > 1e-12 * max(amp):
            return False
    return True


def test_hit_force_structure_and_spectrum():
    cfg = hit_cfg(12, 16, 20)
    assert hit_structure_and_spectrum_ok(cfg)
    for mistake in ("hit_ax_in_fx", "hit_exponent"):
        with planted(mistake):
            assert not hit_structure_and_spectrum_ok(cfg), mistake


def ou_statistics(LX=4, steps=20000, burn=400):
    """Regression of the HIT phases phi_{n+1} on phi_n over a run."""
    o = Oracle(hit_cfg(LX, 4, 4, seed=99))
    o.initial_conditions()
    ph = np.empty((steps, 18))
    for n in range(steps):
        o.step(1)
        ph[n] = o.hit_phases().ravel()
    x, y = ph[burn:-1].ravel(), ph[burn + 1:].ravel()
    slope = np.polyfit(x, y, 1)[0]
    resid = y - slope * x
    return 1.0 - slope, resid.var(), ph[burn:].var()


def ou_ok(LX, stats):
    a, rvar, svar = stats
    a0, D = 0.1 / LX, 0.001 / LX
    return (abs(a / a0 - 1) < 0.15 and abs(rvar / (2 * D) - 1) < 0.03
            and abs(svar / (2 * D / (2 * a0 - a0 * a0)) - 1) < 0.25)


@pytest.mark.slow
def test_hit_phase_walk_is_ornstein_uhlenbeck():
    stats = ou_statistics()
    print("  1-a, residual variance, stationary variance:", stats)
    assert ou_ok(4, stats)
    for mistake in ("hit_no_reversion", "hit_noise"):
        with planted(mistake):
            assert not ou_ok(4, ou_statistics(steps=6000)), mistake


def hydrostatic_error(H=32, bg=1e-4, steps=30000):
    """Conduction layer at rest: relative deviation of rho(y) from the
    barometric profile, and max |u|."""
    nu = 0.1
    cfg = configs.rb2d(NX=4, NY=H, NZ=1, Ra=300, noise=0.0)
    cfg = cfg.copy(beta_t=bg, gravity_y=1.0, tau_u=3 * nu + 0.5, tau_t=3 * nu + 0.5)
    cfg.flags &= ~abi.TEMPERATURE_INITIAL_ADD_NOISE
    o = Oracle(cfg)
    o.initial_conditions()
    o.step(steps)
    h = o.hydro_fields()
    y = np.arange(H) + 0.5
    T0 = 0.5 * (cfg.T_bot + cfg.T_top)
    # T(y) = 1 - y/H  =>  int_0^y (T - T0) dy' = (1 - T0) y - y^2 / (2 H)
    rc = np.exp(bg / (1.0 / 3.0) * ((1 - T0) * y - y * y / (2 * H)))
    rc /= rc.mean()
    r = h["dens"][0, :, 0] / h["dens"][0, :, 0].mean()
    return np.abs(r - rc).max() / np.ptp(rc), np.abs(h["u"]).max()


def test_buoyancy_hydrostatic_barometric_profile():
    err, umax = hydrostatic_error()
    print(f"  barometric profile: max error / variation {err:.2e}, max|u| {umax:.2e}")
    assert err < 1e-3 and umax < 1e-5
    for mistake in ("buoy_t_bot", "buoy_sign"):
        with planted(mistake):
            err, _ = hydrostatic_error()
            assert err > 0.5, mistake


def onset_energy(Ra, steps=20000):
    """Kinetic energy <u^2> of a 32 x 16 no-slip convection layer (one
    critical wavelength, 2.0 H) seeded with 1e-2 temperature noise."""
    cfg = configs.rb2d(NX=32, NY=16, NZ=1, Ra=Ra, Uff=0.2, seed=99, noise=1e-2)
    o = Oracle(cfg)
    o.initial_conditions()
    o.step(steps)
    a = o.dump_averages()
    return a["ux2"] + a["uy2"] + a["uz2"]


@pytest.mark.slow
def test_buoyancy_onset_of_convection():
    e_sub, e_sup = onset_energy(1500), onset_energy(2000)
    print(f"  <u^2> after 20000 steps: Ra 1500 {e_sub:.3e}, Ra 2000 {e_sup:.3e}")
    assert e_sub < 1e-7 and e_sup > 1e-5
    with planted("buoy_sign"):
        assert onset_energy(2000) < 1e-7


def channel_profile_error(tau=1.0, H=16, g=1e-6):
    nu = (tau - 0.5) / 3
    amp = math.sqrt(g * H / 2.0)  # F_x = 2 Amp_x^2 / LY = g
    cfg = abi.make_config(4, H, 1, abi.LB_FLUID | abi.FLUID_FORCING_CHANNEL | abi.FLUID_BC_Y, tau_u=tau, Amp_x=amp)
    o = Oracle(cfg)
    o.initial_conditions()
    o.step(int(40 / (math.pi ** 2 * nu / H ** 2)))
    u = o.hydro_fields()["u"][0, :, 0, 0]
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y) + (g / nu) * (16 * tau ** 2 - 16 * tau + 1) / 24
    return np.abs(u - ua).max() / ua.max()


@pytest.mark.parametrize("tau", [0.7, 1.0])
def test_channel_forcing_steady_poiseuille(tau):
    assert channel_profile_error(tau) < 1e-9
    with planted("channel_lx"):
        assert channel_profile_error(tau) > 0.5


def test_forcing_components_accelerate_a_periodic_box():
    """Periodic 6 x 10 x 14 box, Poiseuille and channel terms on all three
    axes: <u>(n) = (n + 1/2) F with F = 2 Amp_a (4 nu / L^2) + 2 Amp_a^2 / L,
    L = LY for x and z and LX for y (forcing.c:366-379)."""
    tau = 0.9
    nu = (tau - 0.5) / 3
    A = (3e-3, 4e-3, 5e-3)
    LX, LY, LZ = 6, 10, 14
    cfg = abi.make_config(LX, LY, LZ, abi.LB_FLUID | abi.FLUID_FORCING_POISEUILLE | abi.FLUID_FORCING_CHANNEL,
                          tau_u=tau, Amp_x=A[0], Amp_y=A[1], Amp_z=A[2])
    o = Oracle(cfg)
    o.initial_conditions()
    n = 25
    o.step(n)
    u = o.hydro_fields()["u"].reshape(-1, 3).mean(0)
    Ls = (LY, LX, LY)
    F = np.array([2 * A[c] * 4 * nu / Ls[c] ** 2 + 2 * A[c] ** 2 / Ls[c] for c in range(3)])
    np.testing.assert_allclose(u, (n + 0.5) * F, rtol=1e-10)
    # z uses LY, not LZ: the alternative differs by far more than the tolerance
    assert abs(2 * A[2] * 4 * nu / LZ ** 2 + 2 * A[2] ** 2 / LZ - F[2]) > 1e-3 * F[2]
