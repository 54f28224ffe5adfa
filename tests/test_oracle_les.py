"""Pins of the oracle's Smagorinsky closures (CPU only).

The relaxation time the collision used at each node is read back from one
step.  With populations f = f_eq(rho, u) + delta, where delta carries no mass
and no momentum (delta_q = eps w_q (c_qz^2 - c_qy^2)), the post-collision
population of q = 5 (c = (0,0,1), which streams within a y row and never
meets a y wall) is f_eq + (1 - 1/tau) delta, so after the pull
tau(x) = 1 / (1 - (f_5(x + e_z) - f_eq,5(x)) / delta_5(x)).

The expected tau comes from closed forms (tau = 3 (nu + Cs^2 |S|) + 1/2,
les.c:160-180, :253), each chosen so that a plausible mistake fails:

* finite-difference |S| (les.c:31-43) of u_x = U sin(k y): the centred
  difference is exactly U cos(k y) sin(k) (trig identity), |S| = |du_x/dy|;
  a missing symmetrisation or 1/2 changes |S| by sqrt(2) or 2;
* at no-slip y walls the one-sided second-order difference (gradients.c,
  SECOND_ORDER_GRAD) is exact on a parabola; centred differences with the
  periodic images would not be;
* local strain (les.c:20-30): |S| = sqrt(2 Q:Q) / (2 cs^2 rho tau_u) with
  Q = sum_q c c delta; for this delta Q = diag(0, -2eps/9, 2eps/9), and the
  reference's S.zz = S.xz assignment (gradients.c:65) drops the zz term;
* tau_t_les with Pr_t = 1 (les.c:273-297): tau_t - 1/2 - 3 kappa equals
  tau_u_les - 1/2 - 3 nu = 3 nu_t.
"""
import numpy as np
import pytest

from ch4g import abi
from oracle import Oracle

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CV = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [1, -1, 0],
               [-1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1], [0, 1, 1], [0, 1, -1],
               [0, -1, 1], [0, -1, -1]], dtype=float)
EPS = 1e-4
DELTA = EPS * W * (CV[:, 2] ** 2 - CV[:, 1] ** 2)


def feq(rho, u):
    cu = u @ CV.T
    u2 = (u * u).sum(-1)[..., None]
    return rho[..., None] * W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u2)


def taus_after_one_step(cfg, u, with_t=False):
    """Relaxation times of the first collision, per node, read back from q=5
    of the fluid (and of the temperature when with_t)."""
    shape = (cfg.NZ, cfg.NY, cfg.NX)
    rho = np.ones(shape)
    p = feq(rho, u) + DELTA
    g = feq(np.ones(shape), u) + DELTA if with_t else None
    o = Oracle(cfg)
    o.set_populations(p.reshape(-1, 19), None if g is None else g.reshape(-1, 19))
    o.step(1)
    p2, g2 = o.get_populations()
    o.close()

    def back(pop_new, pop_eq):
        post = np.roll(pop_new[..., 5], -1, axis=0)  # post-collision value at the source node
        return 1.0 / (1.0 - (post - pop_eq[..., 5]) / DELTA[5])

    tau_u = back(p2, feq(rho, u))
    tau_t = back(g2, feq(np.ones(shape), u)) if with_t else None
    return tau_u, tau_t


def sine_shear(NX, NY, NZ, U):
    k = 2 * np.pi / NY
    j = np.arange(NY)
    u = np.zeros((NZ, NY, NX, 3))
    u[..., 0] = (U * np.sin(k * j))[None, :, None]
    S = np.abs(U * np.cos(k * j) * np.sin(k))[None, :, None] * np.ones((NZ, NY, NX))
    return u, S


def test_fd_closure_sine_shear():
    NX, NY, NZ, U, cs, tau = 4, 32, 3, 0.02, 0.18, 0.6
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.FLUID_LES, tau_u=tau, les_cs=cs)
    u, S = sine_shear(NX, NY, NZ, U)
    got, _ = taus_after_one_step(cfg, u)
    want = 3 * ((tau - 0.5) / 3 + cs * cs * S) + 0.5
    np.testing.assert_allclose(got, want, rtol=1e-9)
    assert np.ptp(want) > 1e5 * 1e-9 * want.max()  # the closure is resolvable


def test_fd_closure_one_sided_at_walls():
    NX, NY, NZ, A, cs, tau = 3, 12, 2, 4e-4, 0.18, 0.55
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.FLUID_BC_Y | abi.FLUID_LES, tau_u=tau, les_cs=cs)
    y = np.arange(NY) + 0.5
    u = np.zeros((NZ, NY, NX, 3))
    u[..., 0] = (A * y * (NY - y))[None, :, None]
    S = np.abs(A * (NY - 2 * y))[None, :, None] * np.ones((NZ, NY, NX))
    got, _ = taus_after_one_step(cfg, u)
    want = 3 * ((tau - 0.5) / 3 + cs * cs * S) + 0.5
    np.testing.assert_allclose(got, want, rtol=1e-9)


def test_local_strain_closure_with_stress_quirk():
    NX, NY, NZ, cs, tau = 3, 4, 5, 0.18, 0.7
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.FLUID_LES | abi.FLUID_LES_LOCAL_STRAIN, tau_u=tau,
                          les_cs=cs)
    u = np.zeros((NZ, NY, NX, 3))
    got, _ = taus_after_one_step(cfg, u)
    qyy = -2 * EPS / 9  # Q_zz = +2 eps / 9 is replaced by Q_xz = 0
    S = np.sqrt(2 * qyy * qyy) / (2 * (1 / 3) * 1.0 * tau)
    want = 3 * ((tau - 0.5) / 3 + cs * cs * S) + 0.5
    np.testing.assert_allclose(got, np.full(got.shape, want), rtol=1e-9)
    S_sym = np.sqrt(2 * 2 * qyy * qyy) / (2 * (1 / 3) * tau)
    assert abs(3 * cs * cs * (S_sym - S)) > 1e3 * 1e-9 * want  # the quirk is resolvable


def test_temperature_les_follows_fluid_les():
    NX, NY, NZ, U, cs, tau, tau_t = 4, 16, 2, 0.03, 0.18, 0.58, 0.66
    flags = abi.LB_FLUID | abi.LB_TEMPERATURE | abi.FLUID_LES | abi.TEMPERATURE_LES
    cfg = abi.make_config(NX, NY, NZ, flags, tau_u=tau, tau_t=tau_t, les_cs=cs)
    u, S = sine_shear(NX, NY, NZ, U)
    got_u, got_t = taus_after_one_step(cfg, u, with_t=True)
    nu_t = cs * cs * S
    np.testing.assert_allclose(got_u, 3 * ((tau - 0.5) / 3 + nu_t) + 0.5, rtol=1e-9)
    np.testing.assert_allclose(got_t, 3 * ((tau_t - 0.5) / 3 + nu_t) + 0.5, rtol=1e-9)


def test_temperature_les_needs_fluid_les():
    """Without LB_FLUID_LES the temperature keeps its molecular tau_t
    (fluid.c:1083 requires both switches)."""
    NX, NY, NZ, U, tau, tau_t = 4, 16, 2, 0.03, 0.58, 0.66
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID | abi.LB_TEMPERATURE | abi.TEMPERATURE_LES, tau_u=tau,
                          tau_t=tau_t)
    u, _ = sine_shear(NX, NY, NZ, U)
    got_u, got_t = taus_after_one_step(cfg, u, with_t=True)
    np.testing.assert_allclose(got_u, tau, rtol=1e-9)
    np.testing.assert_allclose(got_t, tau_t, rtol=1e-9)
