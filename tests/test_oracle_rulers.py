"""Pins of the oracle's dump_averages rulers (CPU only): the plane averages
along x, y and z written to velocity_averages_{x,y,z}.dat and
temperature_averages_{x,y,z}.dat (output.c:212-491, normalisation :520-690).

* every ruler averages back to the global average of the same member (the
  global averages are pinned in test_oracle_pins.py), except the temperature
  members of the z ruler, which the reference scales by 1/(SY SZ) instead of
  1/(SX SY): their mean is the global value times SX/SZ;
* a conduction state (u = 0, T linear in y) has T(y) = T_bot - dT y/H on
  every y plane and a Nusselt number of exactly 1 on every plane, the wall
  planes included (one-sided differences are exact on a linear profile);
* a shear wave u_x = U sin(k y) has plane means U sin(k y), U^2 sin^2(k y)
  and dissipation nu (U cos(k y) sin k)^2 on each y plane (centred
  differences, trig identity), and zero mean velocity on every x plane.
"""
import numpy as np

from ch4g import abi
from oracle import Oracle

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CV = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [1, -1, 0],
               [-1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1], [0, 1, 1], [0, 1, -1],
               [0, -1, 1], [0, -1, -1]], dtype=float)

FLUID_MEMBERS = ["ux", "uy", "uz", "ux2", "uy2", "uz2", "rho", "ene", "eps", "rho2"]
TEMP_MEMBERS = ["t", "t2", "epst", "dxt", "dyt", "dzt", "uxt", "uyt", "uzt", "nux", "nuy", "nuz"]


def feq(rho, u):
    cu = u @ CV.T
    u2 = (u * u).sum(-1)[..., None]
    return rho[..., None] * W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u2)


def close(a, b, scale):
    return abs(a - b) <= 1e-12 * scale + 1e-15


def test_rulers_average_to_the_global_values():
    rng = np.random.default_rng(3)
    NX, NY, NZ = 5, 7, 3
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.LB_TEMPERATURE | abi.FLUID_BC_Y | abi.TEMPERATURE_BC_Y,
                          tau_u=0.8, tau_t=0.7, T_bot=1.0, T_top=0.0)
    o = Oracle(cfg)
    p = W * (1 + 0.05 * rng.standard_normal((NZ, NY, NX, 19)))
    g = W * (0.5 + 0.05 * rng.standard_normal((NZ, NY, NX, 19)))
    o.set_populations(p.reshape(-1, 19), g.reshape(-1, 19))
    glob, rx, ry, rz = o.dump_rulers()
    assert glob == o.dump_averages()
    for r in (rx, ry, rz):
        for m in FLUID_MEMBERS:
            assert close(r[m].mean(), glob[m], np.abs(r[m]).max()), m
    for r in (rx, ry):
        for m in TEMP_MEMBERS:
            assert close(r[m].mean(), glob[m], np.abs(r[m]).max()), m
    for m in TEMP_MEMBERS:
        assert close(rz[m].mean(), glob[m] * NX / NZ, np.abs(rz[m]).max()), m
    np.testing.assert_array_equal(rx["x"], np.arange(NX) + 0.5)
    np.testing.assert_array_equal(ry["y"], np.arange(NY) + 0.5)
    np.testing.assert_array_equal(rz["z"], np.arange(NZ) + 0.5)
    assert not rx["y"].any() and not ry["x"].any() and not rz["x"].any()


def test_conduction_profile_and_unit_nusselt_on_every_plane():
    NX, NY, NZ, tau_t = 4, 12, 2, 0.8
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.LB_TEMPERATURE | abi.FLUID_BC_Y | abi.TEMPERATURE_BC_Y,
                          tau_u=0.9, tau_t=tau_t, T_bot=1.0, T_top=0.0)
    y = np.arange(NY) + 0.5
    T = 1.0 - y / NY
    shape = (NZ, NY, NX)
    p = feq(np.ones(shape), np.zeros(shape + (3,)))
    g = feq(np.broadcast_to(T[None, :, None], shape).copy(), np.zeros(shape + (3,)))
    o = Oracle(cfg)
    o.set_populations(p.reshape(-1, 19), g.reshape(-1, 19))
    _, rx, ry, rz = o.dump_rulers()
    np.testing.assert_allclose(ry["t"], T, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ry["nuy"], 1.0, rtol=1e-12)
    np.testing.assert_allclose(rx["nuy"], 1.0, rtol=1e-12)
    np.testing.assert_allclose(ry["dyt"], -1.0 / NY, rtol=1e-12)
    kappa = (tau_t - 0.5) / 3
    np.testing.assert_allclose(ry["epst"], kappa / NY ** 2, rtol=1e-12)


def test_shear_wave_plane_means():
    NX, NY, NZ, U, tau = 3, 16, 2, 0.02, 0.7
    nu = (tau - 0.5) / 3
    k = 2 * np.pi / NY
    j = np.arange(NY)
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID, tau_u=tau)
    shape = (NZ, NY, NX)
    u = np.zeros(shape + (3,))
    u[..., 0] = (U * np.sin(k * j))[None, :, None]
    o = Oracle(cfg)
    o.set_populations(feq(np.ones(shape), u).reshape(-1, 19))
    _, rx, ry, _ = o.dump_rulers()
    np.testing.assert_allclose(ry["ux"], U * np.sin(k * j), rtol=0, atol=1e-15)
    np.testing.assert_allclose(ry["ux2"], (U * np.sin(k * j)) ** 2, rtol=0, atol=1e-16)
    np.testing.assert_allclose(ry["eps"], nu * (U * np.cos(k * j) * np.sin(k)) ** 2, rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(rx["ux"], 0.0, atol=1e-15)
    np.testing.assert_allclose(ry["rho"], 1.0, rtol=1e-14)
