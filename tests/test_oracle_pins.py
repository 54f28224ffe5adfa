"""Pins of the FP64 CPU oracle (CPU only).

The reference has no tests or golden files (src/tests/CMakeLists.txt is
empty) and cannot be built here (it needs MPI, parallel HDF5 and a generated
define.dat), so each part of the oracle is pinned by something the
mathematics fixes, chosen so that a plausible mistake fails it:

* drand48 / random_gauss        <- glibc drand48 itself; the polar method's
                                   distribution and rejection rate
* D3Q19 equilibrium + collision <- the textbook D3Q19 equilibrium is a fixed
                                   point of the periodic dynamics
* streaming + periodic halo     <- exact mass/momentum conservation
* bounce-back + Guo + u shift   <- exact discrete Poiseuille solution with the
                                   BGK halfway-bounce-back slip
                                   u_s = (g/nu)(16 tau^2 - 16 tau + 1)/24
* viscosity nu = (tau-1/2)/3    <- shear-wave decay rate
* fixed-T walls (same-direction
  population rule)              <- linear conduction profile is exact only
                                   with that rule when tau_t != 1; Nu = 1
* force lag (u carries F(T_{n-1}))
                                <- periodic momentum budget
                                   dJ = sum rho_{n-1}[F_n + (F_{n-1}-F_n)/(2 tau)]
* NOZEROMODE mean removal       <- a uniform force on a periodic fluid at
                                   rest then does nothing; added to HIT
                                   forcing it leaves the mean velocity at ~0
"""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

from ch4g import abi
from oracle import RNG, Oracle
import configs

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CV = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [1, -1, 0],
               [-1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1], [0, 1, 1], [0, 1, -1],
               [0, -1, 1], [0, -1, -1]], dtype=float)


def textbook_feq(rho, u):
    """D3Q19 second-order equilibrium with cs^2 = 1/3 (Qian, d'Humieres,
    Lallemand 1992)."""
    cu = u @ CV.T
    u2 = (u * u).sum(-1)[..., None]
    return rho[..., None] * W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u2)


# --------------------------------------------------------------------------
def test_drand48_matches_glibc():
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = ctypes.c_double
    for seed in (0, 1, 12345, 2 ** 31 - 1, 1796224233):
        libc.srand48(ctypes.c_long(seed))
        r = RNG(seed)
        a = [libc.drand48() for _ in range(2000)]
        b = [r.drand48() for _ in range(2000)]
        assert a == b


def test_random_gauss_is_standard_normal():
    r = RNG(777)
    x = np.array([r.random_gauss(0.0, 1.0) for _ in range(200000)])
    assert abs(x.mean()) < 0.01
    assert abs(x.var() - 1.0) < 0.01
    # Kolmogorov-Smirnov distance against the normal CDF
    xs = np.sort(x)
    cdf = 0.5 * (1 + np.vectorize(math.erf)(xs / math.sqrt(2)))
    ks = np.max(np.abs(cdf - np.arange(1, len(xs) + 1) / len(xs)))
    assert ks < 0.005


def test_random_gauss_consumes_polar_pairs():
    """The polar method accepts a pair with probability pi/4 and uses two
    uniforms per trial; mean consumption = 8/pi uniforms per normal."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = ctypes.c_double
    r = RNG(99)
    ngauss = 20000
    for _ in range(ngauss):
        r.random_gauss()
    nxt = r.drand48()
    libc.srand48(ctypes.c_long(99))
    used = 0
    while libc.drand48() != nxt:
        used += 1
    assert used % 2 == 0
    assert abs(used / ngauss - 8 / math.pi) < 0.03


# --------------------------------------------------------------------------
def test_equilibrium_is_fixed_point():
    """A uniform textbook equilibrium in a periodic box without forcing is a
    fixed point of collide+stream: the oracle's equilibrium, weights and
    velocity set must match the textbook D3Q19 ones."""
    cfg = abi.make_config(3, 4, 5, abi.LB_FLUID, tau_u=0.8)
    o = Oracle(cfg)
    rho = np.full((5, 4, 3), 1.02)
    u = np.broadcast_to(np.array([0.03, -0.02, 0.015]), (5, 4, 3, 3))
    p = textbook_feq(rho, u)
    o.set_populations(p.reshape(-1, 19))
    o.step(10)
    p2, _ = o.get_populations()
    np.testing.assert_allclose(p2, p, rtol=0, atol=2e-15)


def test_periodic_conservation():
    """Random populations, periodic box, no force: total mass and momentum
    are conserved to round-off; the state relaxes (not a no-op)."""
    rng = np.random.default_rng(1)
    cfg = abi.make_config(6, 5, 4, abi.LB_FLUID, tau_u=0.7)
    o = Oracle(cfg)
    p = W * (1 + 0.05 * rng.standard_normal((4, 5, 6, 19)))
    o.set_populations(p.reshape(-1, 19))
    m0 = p.sum()
    j0 = (p @ CV).sum(axis=(0, 1, 2))
    o.step(50)
    p2, _ = o.get_populations()
    assert abs(p2.sum() - m0) < 1e-13 * m0
    np.testing.assert_allclose((p2 @ CV).sum(axis=(0, 1, 2)), j0, atol=1e-13)
    assert np.abs(p2 - p).max() > 1e-3


@pytest.mark.parametrize("tau", [0.6, 0.8, 1.0, 1.3])
def test_poiseuille_steady_closed_form(tau):
    """Steady force-driven channel: the BGK/Guo/halfway-bounce-back lattice
    solution is the parabola plus a tau-dependent slip,
    u(y) = g y (H-y) / (2 nu) + (g/nu)(16 tau^2 - 16 tau + 1)/24, y = j + 1/2,
    exact in exact arithmetic (derivation in this file's module docstring).
    The slip is ~1e-4 of u_max; the tolerance is 1e-9."""
    H, g = 16, 1e-6
    nu = (tau - 0.5) / 3
    o = Oracle(configs.poiseuille(2, H, 1, tau=tau, g=g))
    o.initial_conditions()
    o.step(int(40 / (math.pi ** 2 * nu / H ** 2)))
    h = o.hydro_fields()
    assert h["force"][0, 0, 0, 0] == g
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y) + (g / nu) * (16 * tau ** 2 - 16 * tau + 1) / 24
    u = h["u"][0, :, 0, 0]
    assert np.abs(u - ua).max() / ua.max() < 1e-9
    # bounce-back symmetry and no transverse flow
    np.testing.assert_allclose(u, u[::-1], rtol=1e-10)
    assert np.abs(h["u"][..., 1:]).max() < 1e-10 * ua.max()


@pytest.mark.slow
def test_poiseuille_startup_series():
    """Configuration 4x64x4 (walls at y), tau = 1, g = 1e-6, 10000 steps:
    the start-up Fourier series of the continuum problem,
    u = g/(2nu) y(H-y) - sum_{n odd} 4 g H^2/(nu pi^3 n^3) sin(n pi y/H) e^{-n^2 pi^2 nu t/H^2},
    agrees with the lattice to the lattice's O(1/H^2) discretisation error
    (the n = 1 transient is still 1.8 % of u_max at t = 1e4)."""
    H, g, tau, T = 64, 1e-6, 1.0, 10000
    nu = (tau - 0.5) / 3
    o = Oracle(configs.poiseuille(4, H, 4, tau=tau, g=g))
    o.initial_conditions()
    o.step(T)
    u = o.hydro_fields()["u"][:, :, :, 0]
    assert np.abs(u - u[0:1, :, 0:1]).max() < 1e-18  # homogeneous in x and z
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y)
    for n in range(1, 200, 2):
        ua -= 4 * g * H * H / (nu * math.pi ** 3 * n ** 3) * np.sin(n * math.pi * y / H) * math.exp(
            -n * n * math.pi ** 2 * nu * T / H ** 2)
    umax = g * H * H / (8 * nu)
    err = np.abs(u[0, :, 0] - ua).max() / umax
    assert err < 2e-4, err


def _shear_decay_rate(L, tau, steps):
    k = 2 * math.pi / L
    cfg = abi.make_config(1, L, 1, abi.LB_FLUID, tau_u=tau)
    o = Oracle(cfg)
    y = np.arange(L) + 0.5
    u = np.zeros((1, L, 1, 3))
    u[0, :, 0, 0] = 1e-4 * np.sin(k * y)
    o.set_populations(textbook_feq(np.ones((1, L, 1)), u).reshape(-1, 19))
    amps = []
    for _ in range(2):
        o.step(steps)
        ux = o.hydro_fields()["u"][0, :, 0, 0]
        amps.append((ux * np.sin(k * y)).sum() * 2 / L)
    return math.log(amps[0] / amps[1]) / steps / (k * k)


def test_shear_wave_viscosity():
    """Decay of u_x = U sin(2 pi y / L) in a periodic box: rate nu_eff k^2
    with nu_eff = nu (1 + a k^2 + ...).  Richardson extrapolation over
    L = 32, 64 removes the O(k^2) lattice term and must give
    nu = (tau - 1/2)/3 (the reference's measure_viscosity.sh check)."""
    tau = 0.8
    nu = (tau - 0.5) / 3
    r1 = _shear_decay_rate(32, tau, 400) / nu
    r2 = _shear_decay_rate(64, tau, 1600) / nu
    assert abs(r1 - 1) < 5e-3
    r0 = (4 * r2 - r1) / 3
    assert abs(r0 - 1) < 2e-5, (r1, r2, r0)


@pytest.mark.parametrize("tau_t", [0.8, 1.4])
def test_conduction_linear_profile_and_nusselt(tau_t):
    """Pure conduction between fixed-T walls (no buoyancy): the linear
    profile T = T_bot - dT y/H (y = j + 1/2) is an exact steady state of the
    reference's wall rule g_q = g*_q + 2 w_q (T_w - T), and Nu = 1.  With an
    anti-bounce-back rule the wall would sit at a tau_t-dependent position
    and this profile would not be steady for tau_t != 1."""
    H = 12
    flags = abi.LB_FLUID | abi.LB_TEMPERATURE | abi.FLUID_BC_Y | abi.TEMPERATURE_BC_Y
    cfg = abi.make_config(2, H, 1, flags, tau_u=1.0, tau_t=tau_t, T_bot=1.0, T_top=0.0)
    o = Oracle(cfg)
    y = np.arange(H) + 0.5
    T = 1.0 - y / H
    p = np.broadcast_to(W, (1, H, 2, 19))
    g = (T[None, :, None, None] * W) * np.ones((1, H, 2, 1))
    o.set_populations(p.reshape(-1, 19), g.reshape(-1, 19))
    o.step(300)
    t = o.hydro_fields()["t"][0, :, 0]
    np.testing.assert_allclose(t, T, rtol=0, atol=1e-14)
    av = o.dump_averages()
    assert abs(av["nuy"] - 1.0) < 1e-12


def test_buoyancy_momentum_budget_pins_force_lag():
    """Periodic box with buoyancy and a diffusing temperature field: per
    step the total momentum changes by sum rho_{n-1}[F_n + (F_{n-1} - F_n)/(2 tau)]
    (collision relaxes j towards rho u = j + rho F_{n-1}/2, Guo adds
    (1 - 1/(2 tau)) rho F_n).  The quadratic buoyancy term makes the total
    force change while the temperature variance diffuses; without the
    one-step lag the correction term vanishes and the budget fails."""
    rng = np.random.default_rng(3)
    NX, NY, NZ = 8, 6, 4
    tau = 0.9
    flags = abi.LB_FLUID | abi.LB_TEMPERATURE | abi.TEMPERATURE_BUOYANCY
    cfg = abi.make_config(NX, NY, NZ, flags, tau_u=tau, tau_t=0.7, beta_t=0.01, beta2_t=0.2, gravity_y=1.0, gravity_x=0.3,
                          T_bot=1.0, T_top=0.0)
    o = Oracle(cfg)
    T = 0.5 + 0.3 * rng.standard_normal((NZ, NY, NX))
    g = T[..., None] * W
    p = np.broadcast_to(W, (NZ, NY, NX, 19))
    o.set_populations(p.reshape(-1, 19), g.reshape(-1, 19))
    prev = None
    checked = 0
    for n in range(7):
        h = o.hydro_fields()
        pp, _ = o.get_populations()
        J = (pp @ CV).sum(axis=(0, 1, 2))
        if prev is not None:
            rho_p, F_p, J_p = prev
            F_n = h["force"]
            expect = (rho_p[..., None] * (F_n + (F_p - F_n) / (2 * tau))).sum(axis=(0, 1, 2))
            np.testing.assert_allclose(J - J_p, expect, rtol=1e-10, atol=1e-12 * np.abs(expect).max())
            if n >= 2:  # steps 1 and 0 share F(T_0); the lag shows from step 2 on
                lagless = (rho_p[..., None] * F_n).sum(axis=(0, 1, 2))
                assert np.abs(lagless - expect).max() > 1e-3 * np.abs(expect).max()
            checked += 1
        prev = (h["dens"], h["force"], J)
        o.step(1)
    assert checked == 6


def test_nozeromode_removes_the_mean_force():
    flags = abi.LB_FLUID | abi.FLUID_FORCING_POISEUILLE | abi.FLUID_FORCING_NOZEROMODE
    o = Oracle(abi.make_config(6, 6, 6, flags, tau_u=0.8, Amp_x=1e-3))
    o.initial_conditions()
    o.step(40)
    h = o.hydro_fields()
    assert np.abs(h["u"]).max() < 1e-16 and np.abs(h["force"]).max() < 1e-18
    # HIT modes + a uniform force: the modes drive the flow; their k-free
    # terms (sin 2 pi phi of the modes without a y or z wave number) and the
    # uniform part have a non-zero mean, which the switch removes, so the
    # mean velocity stays near zero (the momentum change is sum rho F, and
    # rho - 1 = O(u^2) leaves ~1e-9) while without it it grows to ~0.07
    base = abi.LB_FLUID | abi.FLUID_FORCING_HIT | abi.FLUID_FORCING_POISEUILLE
    res = {}
    for nz in (0, abi.FLUID_FORCING_NOZEROMODE):
        o = Oracle(abi.make_config(8, 8, 8, base | nz, tau_u=0.6, Amp_x=2e-3, Amp_y=1e-4, Amp_z=1e-4, seed=5))
        o.initial_conditions()
        o.step(30)
        res[nz] = o.hydro_fields()["u"]
    m0 = np.abs(res[0].reshape(-1, 3).mean(0))
    u = res[abi.FLUID_FORCING_NOZEROMODE]
    m1 = np.abs(u.reshape(-1, 3).mean(0))
    assert m0[0] > 1e-2 and np.all(m1 < 1e-7 * m0[0])
    assert np.abs(u).max() > 1e-5
