"""Pins of the oracle's Lagrangian particles (CPU only).

The reference ships no particle tests or fixtures, so each part of the
restatement (oracle/ch4_oracle_lagrange.c) is pinned by what the mathematics
or a library fixes, each chosen so that a plausible slip (a swapped weight,
a wrong ghost sign, a missing factor, an off-by-one cell) fails it:

* type ladders, names, positions  <- the parameter arithmetic of
                                     param.in and glibc drand48 itself
* trilinear interpolation         <- exact for linear fields; the no-slip
                                     ghost keeps a wall-vanishing linear
                                     field linear; periodic halo = image
* velocity gradient               <- centred and one-sided second-order
                                     differences are exact for quadratic
                                     fields; the interpolation fix gives the
                                     first-cell gradient in the ghost half-cell
* tracer advance                  <- uniform flow: x_n = x_0 + n u dt
* inertial advance                <- Stokes relaxation v -> u (1 - e^{-t/tau})
                                     (second-order AB2 error); terminal
                                     velocity -g tau (1 - beta)
* walls and re-injection          <- mirror / stop rules; glibc drand48 draws
* two-way feedback                <- deposits sum to V_p f_p (weights sum to
                                     one) with f_p in closed form
* orientation                     <- ladders and random_vector draws from the
                                     glibc stream; Jeffery orbits in simple
                                     shear (closed form in the shear plane,
                                     orbit constant in 3-D); the gyrotaxis
                                     term by hand; swimming in a straight line
* particle temperature            <- cp ladder after the orientation ones;
                                     decay rate of a sphere at Nusselt
                                     number 2 with radius and density from
                                     tau_drag and beta; the scheme's fixed
                                     point; re-injection wall temperatures;
                                     the conduction profile at the start
"""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

from ch4g import abi
from oracle import Oracle

LIBC = ctypes.CDLL(ctypes.util.find_library("c"))
LIBC.drand48.restype = ctypes.c_double


def glibc_stream(seed, n):
    LIBC.srand48(ctypes.c_long(seed))
    return np.array([LIBC.drand48() for _ in range(n)])


def quiescent(NX, NY, NZ, flags=0, tau=0.8, seed=7, **kw):
    cfg = abi.make_config(NX, NY, NZ, abi.LB_FLUID | flags, tau_u=tau, seed=seed, **kw)
    o = Oracle(cfg)
    o.initial_conditions()
    return o, cfg


def uniform_flow(o, cfg, u0):
    W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    CV = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                   [1, -1, 0], [-1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1], [0, 1, 1],
                   [0, 1, -1], [0, -1, 1], [0, -1, -1]], dtype=float)
    cu = CV @ np.asarray(u0)
    feq = W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * float(np.dot(u0, u0)))
    p = np.broadcast_to(feq, (cfg.NZ, cfg.NY, cfg.NX, 19)).copy()
    o.set_populations(p)


def records(n):
    return np.zeros(n, dtype=abi.particle_dtype())


# --------------------------------------------------------------------------
def test_type_ladders_and_names():
    o, _ = quiescent(8, 8, 8)
    flags = abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS
    o.particles_init(abi.make_particle_config(flags, 30, tau_drag=(3, 0.5, 1.5), beta_coeff=(2, 0.5, 1.0)))
    a = o.particles()
    assert len(a) == 30
    assert np.array_equal(a["name"], np.arange(30))
    ntypes = 3 * 2 + 1
    for r in a:
        t = int(r["name"]) % ntypes
        if t == 0:
            assert r["tau_drag"] == 0.0 and r["beta_coeff"] == 1.0
        else:
            h = t - 1
            assert r["tau_drag"] == 0.5 + (h % 3) * 0.5
            assert r["beta_coeff"] == 0.5 + ((h // 3) % 2) * 0.5
    # geometric ladder: min * step^i with step = (max/min)^(1/(types-1))
    o2, _ = quiescent(8, 8, 8)
    o2.particles_init(abi.make_particle_config(abi.LAGRANGE_TAUDRAG_INCREMENT_LOG, 8, tau_drag=(4, 0.1, 10.0)))
    t = o2.particles()["tau_drag"][:4]
    step = (10.0 / 0.1) ** (1 / 3)
    assert np.allclose(t, 0.1 * step ** np.arange(4), rtol=1e-15)
    assert t[-1] == pytest.approx(10.0, rel=1e-14)


def test_positions_follow_the_drand48_stream():
    # fluid only, no noise: the first draws after srand48 are the positions
    o, cfg = quiescent(16, 8, 4, seed=99)
    o.particles_init(abi.make_particle_config(0, 20, tau_drag=(1, 2.0, 2.0)))
    a = o.particles()
    r = glibc_stream(99, 60).reshape(20, 3)
    assert np.array_equal(a["x"], r[:, 0] * 16)
    assert np.array_equal(a["y"], r[:, 1] * 8)
    assert np.array_equal(a["z"], r[:, 2] * 4)
    # with temperature noise the k,j,i noise draws come first
    cfg2 = abi.make_config(4, 4, 2, abi.LB_FLUID | abi.LB_TEMPERATURE | abi.TEMPERATURE_INITIAL_ADD_NOISE,
                           seed=5, noise_t=1e-3)
    o2 = Oracle(cfg2)
    o2.initial_conditions()
    o2.particles_init(abi.make_particle_config(0, 3, tau_drag=(1, 1.0, 1.0)))
    r2 = glibc_stream(5, 32 + 9)[32:].reshape(3, 3)
    assert np.array_equal(o2.particles()["x"], r2[:, 0] * 4)


def _probe(o, pos, u=None, t=None):
    a = records(len(pos))
    a["x"], a["y"], a["z"] = pos[:, 0], pos[:, 1], pos[:, 2]
    o.set_particles(a)
    return o.probe(u, t)


def _centres(cfg):
    z, y, x = np.meshgrid(np.arange(cfg.NZ) + 0.5, np.arange(cfg.NY) + 0.5, np.arange(cfg.NX) + 0.5, indexing="ij")
    return x, y, z


def test_trilinear_exact_for_linear_fields():
    o, cfg = quiescent(12, 10, 8, flags=abi.LB_TEMPERATURE)
    o.particles_init(abi.make_particle_config(abi.LAGRANGE_GRADIENT, 1, tau_drag=(1, 1.0, 1.0)))
    x, y, z = _centres(cfg)
    a0, B = np.array([0.1, -0.2, 0.3]), np.array([[0.01, 0.02, -0.03], [0.005, -0.01, 0.02], [0.03, 0.01, 0.004]])
    u = a0 + np.stack([x, y, z], -1) @ B.T
    t = 0.7 + 0.05 * x - 0.02 * y + 0.01 * z
    rng = np.random.default_rng(0)
    pos = np.column_stack([rng.uniform(0.5, 11.5, 200), rng.uniform(0.5, 9.5, 200), rng.uniform(0.5, 7.5, 200)])
    r = _probe(o, pos, u, t)
    want = a0 + pos @ B.T
    got = np.column_stack([r["ux"], r["uy"], r["uz"]])
    assert np.allclose(got, want, rtol=0, atol=1e-14)
    assert np.allclose(r["t"], 0.7 + 0.05 * pos[:, 0] - 0.02 * pos[:, 1] + 0.01 * pos[:, 2], atol=1e-14)
    # gradient of a linear field: centred differences are exact away from the periodic seams
    inner = (pos[:, 0] > 1.5) & (pos[:, 0] < 10.5) & (pos[:, 1] > 1.5) & (pos[:, 1] < 8.5) & \
            (pos[:, 2] > 1.5) & (pos[:, 2] < 6.5)
    G = np.column_stack([r[n] for n in ["dx_ux", "dy_ux", "dz_ux", "dx_uy", "dy_uy", "dz_uy",
                                        "dx_uz", "dy_uz", "dz_uz"]])[inner]
    assert np.allclose(G, B.reshape(1, 9), atol=1e-14)


def test_periodic_halo_is_the_image():
    o, cfg = quiescent(8, 6, 4)
    o.particles_init(abi.make_particle_config(0, 1, tau_drag=(1, 1.0, 1.0)))
    x, y, z = _centres(cfg)
    f = np.sin(2 * np.pi * x / 8) + 0.3 * np.cos(2 * np.pi * y / 6) + 0.1 * z
    u = np.stack([f, 2 * f, -f], -1)
    # across the x seam, y and z on cell centres: blend of u(NX-1) and u(0)
    pos = np.array([[7.8, 2.5, 1.5], [0.2, 2.5, 1.5], [8.3, 2.5, 1.5], [-0.2, 2.5, 1.5]])
    r = _probe(o, pos, u)
    fl, fr = f[1, 2, 7], f[1, 2, 0]
    for p, got in zip(pos[:, 0], r["ux"]):
        xw = p % 8
        w = (xw - 7.5) if xw >= 7.5 else (xw + 0.5)
        assert got == pytest.approx((1 - w) * fl + w * fr, abs=1e-15)


def test_noslip_ghost_and_gradient_near_wall():
    o, cfg = quiescent(6, 8, 6, flags=abi.FLUID_BC_Y)
    o.particles_init(abi.make_particle_config(abi.LAGRANGE_GRADIENT, 1, tau_drag=(1, 1.0, 1.0)))
    x, y, z = _centres(cfg)
    # u_x = a y vanishes at the bottom wall: the ghost u_g = -u keeps it linear
    a = 0.01
    u = np.stack([a * y, 0 * y, 0 * y], -1)
    pos = np.array([[2.7, 0.1, 3.2], [1.1, 0.45, 2.9], [4.0, 0.3, 0.7]])
    r = _probe(o, pos, u)
    assert np.allclose(r["ux"], a * pos[:, 1], atol=1e-16)
    # quadratic field: centred and one-sided second-order differences are exact
    q = lambda X, Y, Z: 0.01 * X * X - 0.02 * X * Y + 0.003 * Y * Y + 0.004 * Z * Z + 0.05 * Y  # noqa: E731
    dqdy = lambda X, Y, Z: -0.02 * X + 0.006 * Y + 0.05  # noqa: E731
    u = np.stack([q(x, y, z), 0 * y, 0 * y], -1)
    pos = np.array([[2.2, 0.7, 2.5], [3.1, 7.2, 2.6], [2.9, 3.3, 3.1]])
    r = _probe(o, pos, u)
    assert np.allclose(r["dy_ux"], dqdy(pos[:, 0], pos[:, 1], pos[:, 2]), atol=1e-14)
    # ghost half-cell (y < 1/2): the interpolation fix uses the first-row gradient
    pos = np.array([[2.5, 0.2, 2.5]])
    r = _probe(o, pos, u)
    assert r["dy_ux"][0] == pytest.approx(dqdy(2.5, 0.5, 2.5), abs=1e-14)


def test_temperature_gradient_of_the_conduction_profile():
    """gradient_scalar at the particles (gradients.c:505-827 with
    LB_TEMPERATURE_BC_Y): for the linear conduction profile the fixed-T
    ghost t_g = 2 T_w - t continues the line, so centred differences, the
    one-sided wall rows and the ghost-row fix all give dT/dy = -1/H exactly
    (to rounding), and dT/dx = dT/dz = 0 exactly."""
    o, cfg = quiescent(6, 8, 5, flags=abi.LB_TEMPERATURE | abi.TEMPERATURE_BC_Y, T_bot=1.0, T_top=0.0, T_ref=0.0)
    o.particles_init(abi.make_particle_config(abi.LAGRANGE_GRADIENT, 1, tau_drag=(1, 1.0, 1.0)))
    x, y, z = _centres(cfg)
    t = 1.0 - y / 8.0
    rng = np.random.default_rng(4)
    pos = np.column_stack([rng.uniform(0, 6, 300), rng.uniform(0.01, 7.99, 300), rng.uniform(0, 5, 300)])
    r = _probe(o, pos, np.zeros(x.shape + (3,)), t)
    assert np.allclose(r["t"], 1.0 - pos[:, 1] / 8.0, rtol=0, atol=1e-15)
    assert np.allclose(r["dy_t"], -1.0 / 8.0, rtol=0, atol=1e-15)
    assert not r["dx_t"].any() and not r["dz_t"].any()


def test_gradient_at_a_periodic_ghost_is_one_sided():
    """Without walls the interpolation fix takes the second-order one-sided
    difference at the ghost node itself, over its periodic image
    (gradients.c:413-488 for u, :762-825 for T), not a centred difference of
    images.  For f = b x_c on the interior centres of an N-periodic axis the
    ghost at -1 holds f(N-1) = b (N - 1/2), so
      ghost:        -3/2 f(-1) + 2 f(0) - 1/2 f(1) = b (1 - 3N/2)
      first node:   (f(1) - f(-1)) / 2              = b (2 - N) / 2
    and a particle at x = 1/4 (weights 1/4 and 3/4) reads their blend,
    -5 b for N = 8 (a centred difference at the ghost would give -3 b);
    the mirror case x = N - 1/4 uses the ghost at N the same way."""
    o, cfg = quiescent(8, 6, 4, flags=abi.LB_TEMPERATURE)
    o.particles_init(abi.make_particle_config(abi.LAGRANGE_GRADIENT, 1, tau_drag=(1, 1.0, 1.0)))
    x, y, z = _centres(cfg)
    b, N = 0.01, 8
    u = np.stack([b * x, 0 * x, 0 * x], -1)
    pos = np.array([[0.25, 2.5, 1.5], [N - 0.25, 2.5, 1.5]])
    r = _probe(o, pos, u, b * x)
    want = 0.25 * b * (1 - 1.5 * N) + 0.75 * b * (2 - N) / 2
    assert want == pytest.approx(-5 * b)
    assert np.allclose(r["dx_ux"], want, rtol=1e-14, atol=0)
    assert np.allclose(r["dx_t"], want, rtol=1e-14, atol=0)


def test_tracers_move_with_uniform_flow():
    o, cfg = quiescent(16, 8, 8, tau=0.9)
    u0 = np.array([0.031, -0.017, 0.009])
    uniform_flow(o, cfg, u0)
    o.particles_init(abi.make_particle_config(abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_WRAP, 10,
                                              tau_drag=(1, 0.0, 0.0)))
    a0 = o.particles()
    n = 40
    o.step(n)
    a = o.particles()
    for c, S, u in zip("xyz", (16, 8, 8), u0):
        want = np.mod(a0[c] + n * u, S)
        d = np.abs(a[c] - want)
        assert np.all(np.minimum(d, S - d) < 1e-12)
        assert np.all((a[c] >= 0) & (a[c] < S))
    assert np.allclose(a["vx"], u0[0], atol=1e-15)


def test_stokes_relaxation_and_terminal_velocity():
    tau_p = 20.0
    o, cfg = quiescent(8, 8, 8, tau=0.8)
    u0 = np.array([0.02, 0.0, -0.01])
    uniform_flow(o, cfg, u0)
    o.particles_init(abi.make_particle_config(0, 4, tau_drag=(1, tau_p, tau_p)))
    n = 60
    o.step(n)
    v = np.column_stack([o.particles()[c] for c in ("vx", "vy", "vz")])
    exact = u0 * (1 - math.exp(-n / tau_p))
    # AB2 after an Euler start: O(dt^2/tau^2) relative error of the transient
    assert np.allclose(v, exact, rtol=2e-3, atol=1e-6)
    assert not np.allclose(v, u0 * (1 - math.exp(-n / (2 * tau_p))), rtol=2e-3)
    # settling in a quiescent fluid: -g tau (1 - beta), Du/Dt = 0
    g, tau_p, beta = 1e-3, 5.0, 0.6
    o2, _ = quiescent(6, 6, 6, gravity_y=g)
    flags = abi.LAGRANGE_GRAVITY | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS
    o2.particles_init(abi.make_particle_config(flags, 3, tau_drag=(1, tau_p, tau_p), beta_coeff=(1, beta, beta)))
    o2.step(400)
    a = o2.particles()
    assert np.allclose(a["vy"], -g * tau_p * (1 - beta), rtol=1e-12)
    assert np.allclose(a["vx"], 0.0, atol=1e-20)


def _one_particle(o, **fields):
    a = records(1)
    for k, v in fields.items():
        a[k] = v
    o.set_particles(a)


def test_walls_mirror_stop_and_reinject():
    g = 1e-3
    # elastic: mirror position, reverse the normal velocity
    o, cfg = quiescent(6, 8, 6, flags=abi.FLUID_BC_Y, gravity_y=g)
    o.particles_init(abi.make_particle_config(0, 1, tau_drag=(1, 1e9, 1e9)))
    _one_particle(o, x=2.0, y=0.2, z=3.0, vy=-0.3, vx=0.01, tau_drag=1e9)
    o.step(1)
    p = o.particles()[0]
    assert p["y"] == pytest.approx(0.1, abs=1e-12) and p["vy"] == pytest.approx(0.3, rel=1e-8)
    assert p["vx"] == pytest.approx(0.01, rel=1e-8)
    _one_particle(o, x=2.0, y=7.9, z=3.0, vy=0.3, tau_drag=1e9, vy_old=0.3)
    o.step(1)
    p = o.particles()[0]
    assert p["y"] == pytest.approx(7.8, abs=1e-12) and p["vy"] < 0
    # inelastic: stop on the wall
    o, _ = quiescent(6, 8, 6, flags=abi.FLUID_BC_Y)
    o.particles_init(abi.make_particle_config(abi.LAGRANGE_BC_INELASTIC, 1, tau_drag=(1, 1e9, 1e9)))
    _one_particle(o, x=2.0, y=0.2, z=3.0, vy=-0.3, vx=0.01, tau_drag=1e9)
    o.step(1)
    p = o.particles()[0]
    assert p["y"] == 0.0 and p["vx"] == 0.0 and p["vy"] == 0.0
    # re-injection: x = SX r, y + SY, z = SZ r with the next two drand48 draws;
    # Stokes velocity -g (1-beta) tau, halved
    seed = 31
    o, _ = quiescent(6, 8, 4, flags=abi.FLUID_BC_Y, seed=seed, gravity_y=g)
    flags = (abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT | abi.LAGRANGE_GRAVITY |
             abi.LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF | abi.LAGRANGE_WRAP)
    o.particles_init(abi.make_particle_config(flags, 2, tau_drag=(1, 4.0, 4.0)))
    # consumed: 6 draws of positions; then overwrite with a falling particle
    a = records(2)
    a["x"], a["y"], a["z"] = [1.0, 3.0], [0.05, 4.0], [1.0, 2.0]
    a["vy"], a["tau_drag"] = [-0.2, 0.0], [4.0, 4.0]
    o.set_particles(a)
    o.step(1)
    r = glibc_stream(seed, 8)[6:]
    p = o.particles()
    assert p["x"][0] == 6 * r[0] and p["z"][0] == 4 * r[1]
    assert p["y"][0] == pytest.approx(0.05 - 0.2 + 8, abs=1e-12)
    assert p["vy"][0] == -g * (1 - 0.0) * 4.0 / 2.0
    assert p["y"][1] == 4.0 and p["x"][1] == 3.0


def test_twoway_feedback_deposit():
    g, tau_p, beta = 1e-4, 3.0, 0.6
    o, cfg = quiescent(8, 8, 8, tau=0.8, gravity_y=g)
    flags = (abi.LAGRANGE_GRAVITY | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS |
             abi.LAGRANGE_TWOWAY_MOMENTUM)
    o.particles_init(abi.make_particle_config(flags, 1, tau_drag=(1, tau_p, tau_p), beta_coeff=(1, beta, beta)))
    _one_particle(o, x=3.3, y=4.6, z=2.8, tau_drag=tau_p, beta_coeff=beta)
    o.step(1)  # first move: a = -g (1 - beta) in y, Du/Dt = 0
    p = o.particles()[0]
    assert p["ay"] == pytest.approx(-g * (1 - beta), rel=1e-12)
    o.step(1)  # second step: deposit on the force field
    F = o.hydro_fields()["force"]
    nu = (0.8 - 0.5) / 3
    vol = 4 / 3 * 3.14159265359 * (3 * nu * beta * tau_p) ** 1.5
    dr = (3 - beta) / (2 * beta)
    fpy = 0.0 - (-g) + dr * (-g - p["ay"])
    assert F[..., 1].sum() == pytest.approx(vol * fpy, rel=1e-12)
    assert np.count_nonzero(F[..., 1]) == 8
    # trilinear weights: the node at (3,4,2) (centre 3.5,4.5,2.5) gets (1-0.2)(1-0.1)(1-0.3)
    assert F[2, 4, 3, 1] == pytest.approx(vol * fpy * 0.8 * 0.9 * 0.7, rel=1e-12)
    assert F[..., 0].sum() == pytest.approx(0.0, abs=1e-30)


# --------------------------------------------------------------------------
# orientation (lagrange.c:3470-3752)
ORIENT = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY


def _shear(cfg, gamma):
    """u_x = gamma (y - NY/2): uniform velocity gradient d u_x / d y = gamma."""
    x, y, z = _centres(cfg)
    u = np.zeros(x.shape + (3,))
    u[..., 0] = gamma * (y - cfg.NY / 2)
    return u


def test_orientation_ladders_and_initial_draws():
    seed = 41
    o, cfg = quiescent(8, 8, 4, seed=seed)
    flags = (abi.LAGRANGE_ADD_TRACER | ORIENT | abi.LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS |
             abi.LAGRANGE_ORIENTATION_ACTIVE)
    pc = abi.make_particle_config(flags, 40, tau_drag=(2, 0.0, 1.0), aspect_ratio=(3, 1.0, 5.0),
                                  gyrotaxis_velocity=(2, 0.5, 2.0), swim_velocity=(2, 0.0, 0.01))
    o.particles_init(pc)
    a = o.particles()
    nt = 2 * 3 * 2 * 2 + 1
    for r in a:
        t = int(r["name"]) % nt
        if t == 0:
            assert (r["tau_drag"], r["aspect_ratio"], r["gyrotaxis_velocity"], r["swim_velocity"]) == (0, 1, 0, 0)
            continue
        h = t - 1  # tau varies fastest, then aspect ratio, gyrotaxis, swimming
        assert r["tau_drag"] == [0.0, 1.0][h % 2]
        assert r["aspect_ratio"] == [1.0, 3.0, 5.0][(h // 2) % 3]
        assert r["gyrotaxis_velocity"] == pytest.approx([0.5, 2.0][(h // 6) % 2], rel=1e-15)
        assert r["swim_velocity"] == [0.0, 0.01][(h // 12) % 2]
    # per particle: x, y, z, then random_vector's theta and phi
    d = glibc_stream(seed, 5 * 40).reshape(40, 5)
    assert np.array_equal(a["x"], d[:, 0] * 8)
    theta, phi = np.arccos(2.0 * d[:, 3] - 1.0), 2 * 3.14159265359 * d[:, 4]
    assert np.allclose(a["px"], np.sin(theta) * np.sin(phi), rtol=0, atol=1e-15)  # numpy vs glibc sin: 1 ulp
    assert np.allclose(a["pz"], np.cos(theta), rtol=0, atol=1e-15)
    assert np.allclose(a["px"] ** 2 + a["py"] ** 2 + a["pz"] ** 2, 1.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("lam", [1.0, 3.0, 0.25])
def test_jeffery_orbit_in_plane_matches_the_closed_form(lam):
    """In u_x = gamma y, p = (cos th, sin th, 0) obeys
    dth/dt = -gamma (cos^2 th + lam^2 sin^2 th) / (lam^2 + 1), solved by
    p ~ (lam cos psi, sin psi), psi = psi0 - gamma lam t / (lam^2 + 1)
    (lam = 1: rigid rotation with the vorticity, -gamma/2)."""
    gamma, n = 0.01, 600
    o, cfg = quiescent(8, 16, 4)
    o.particles_init(abi.make_particle_config(ORIENT, 1, tau_drag=(1, 0.0, 0.0), aspect_ratio=(1, lam, lam)))
    th0 = np.linspace(0.1, 3.0, 5)
    a = records(len(th0))
    a["name"] = np.arange(len(th0))
    a["x"], a["y"], a["z"] = 4.0, 8.0, 2.0
    a["px"], a["py"] = np.cos(th0), np.sin(th0)
    a["aspect_ratio"] = lam
    o.set_particles(a)
    u = _shear(cfg, gamma)
    for s in range(n):
        r = o.move(u, first=(s == 0))
    assert np.allclose(r["dy_ux"], gamma, rtol=1e-12)
    w = gamma * lam / (lam * lam + 1)
    psi = np.arctan2(lam * np.sin(th0), np.cos(th0)) - w * n
    e = np.stack([lam * np.cos(psi), np.sin(psi)], -1)
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    got = np.stack([r["px"], r["py"]], -1)
    # AB2 after an Euler step: observed <= 6e-5 over these 600 steps; a
    # dropped -(p.S'p)p term gives 4e-3, a transposed gradient 0.3-2
    assert np.abs(got - e).max() < 1e-4
    assert np.all(r["pz"] == 0.0)
    # the last dp/dt is the right-hand side at the previous p: its length is
    # |dth/dt| there (one step, an angle of < 0.01, away)
    rate = gamma * (r["px"] ** 2 + lam * lam * r["py"] ** 2) / (lam * lam + 1)
    assert np.allclose(np.hypot(r["dt_px"], r["dt_py"]), rate, rtol=2e-2)


def test_jeffery_orbit_constant_in_three_dimensions():
    """(px^2 + lam^2 py^2) / pz^2 is a constant of Jeffery's equation in
    u_x = gamma y (the orbit constant)."""
    gamma, lam, n = 0.02, 4.0, 400
    o, cfg = quiescent(8, 16, 4, seed=3)
    o.particles_init(abi.make_particle_config(ORIENT, 6, tau_drag=(1, 0.0, 0.0), aspect_ratio=(1, lam, lam)))
    a = o.particles()
    a["x"], a["y"], a["z"] = 4.0, 8.0, 2.0
    o.set_particles(a)
    c0 = (a["px"] ** 2 + lam ** 2 * a["py"] ** 2) / a["pz"] ** 2
    u = _shear(cfg, gamma)
    for s in range(n):
        r = o.move(u, first=(s == 0))
    c1 = (r["px"] ** 2 + lam ** 2 * r["py"] ** 2) / r["pz"] ** 2
    assert np.allclose(c1, c0, rtol=2e-4)
    assert np.allclose(r["px"] ** 2 + r["py"] ** 2 + r["pz"] ** 2, 1.0, atol=1e-15)
    assert not np.allclose(np.stack([r["px"], r["py"]]), np.stack([a["px"], a["py"]]), atol=1e-2)


def test_gyrotaxis_first_column_and_swimming():
    # no flow, no gravity switch: a = -1/(2 v0) (0, 0, 1), added to the first
    # column of S only (lagrange.c:3553-3562).  By hand for p = (0.6, 0, 0.8),
    # v0 = 2: S[2][0] = -0.15, S p = (0, 0, -0.09), p.S p = -0.072,
    # F = S p + 0.072 p = (0.0432, 0, -0.0324); Euler: p + F, normalised.
    o, cfg = quiescent(6, 6, 6)
    flags = ORIENT | abi.LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS
    o.particles_init(abi.make_particle_config(flags, 1, tau_drag=(1, 0.0, 0.0), gyrotaxis_velocity=(1, 2.0, 2.0)))
    a = records(1)
    a["x"], a["y"], a["z"] = 3.0, 3.0, 3.0
    a["px"], a["pz"], a["aspect_ratio"], a["gyrotaxis_velocity"] = 0.6, 0.8, 3.0, 2.0
    o.set_particles(a)
    r = o.move(np.zeros((6, 6, 6, 3)), first=True)[0]
    q = np.array([0.6 + 0.0432, 0.0, 0.8 - 0.0324])
    q /= np.linalg.norm(q)
    assert np.allclose([r["px"], r["py"], r["pz"]], q, rtol=1e-14, atol=1e-16)
    assert np.allclose([r["dt_px"], r["dt_py"], r["dt_pz"]], [0.0432, 0.0, -0.0324], rtol=1e-13)
    # active tracer in still fluid: straight line at swim_velocity along p
    o, cfg = quiescent(16, 8, 8)
    flags = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_ACTIVE | abi.LAGRANGE_WRAP
    o.particles_init(abi.make_particle_config(flags, 4, tau_drag=(1, 0.0, 0.0), swim_velocity=(1, 0.01, 0.01)))
    a0 = o.particles()
    o.step(30)
    a1 = o.particles()
    for c, pc, S in (("x", "px", 16), ("y", "py", 8), ("z", "pz", 8)):
        d = np.abs(a1[c] - np.mod(a0[c] + 30 * 0.01 * a0[pc], S))
        assert np.all(np.minimum(d, S - d) < 1e-12)
    assert np.array_equal(a1["px"], a0["px"]) or np.allclose(a1["px"], a0["px"], rtol=0, atol=2e-16)


def test_orientation_turned_away_from_the_walls():
    """LAGRANGE_ORIENTATION_BC: a particle that crosses a y wall leaves with
    p_y pointing into the fluid (lagrange.c:3866-3870, :3942-3946)."""
    for bc in (0, abi.LAGRANGE_ORIENTATION_BC):
        o, cfg = quiescent(6, 8, 6, flags=abi.FLUID_BC_Y)
        flags = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ORIENTATION | bc
        o.particles_init(abi.make_particle_config(flags, 2, tau_drag=(1, 1e9, 1e9)))
        a = records(2)
        a["x"], a["y"], a["z"] = [2.0, 3.0], [0.2, 7.9], [3.0, 2.0]
        a["vy"], a["vy_old"], a["tau_drag"] = [-0.3, 0.3], [-0.3, 0.3], 1e9
        a["px"], a["py"] = [0.6, 0.6], [-0.8, 0.8]
        o.set_particles(a)
        r = o.move(np.zeros((6, 8, 6, 3)), first=True)
        assert r["y"][0] == pytest.approx(0.1, abs=1e-12) and r["y"][1] == pytest.approx(7.8, abs=1e-12)
        want = [0.8, -0.8] if bc else [-0.8, 0.8]
        assert np.allclose(r["py"], want, rtol=1e-15) and np.allclose(r["px"], 0.6, rtol=1e-15)


def test_second_orientation_stays_normal_to_p():
    """n starts as p x v / |p x v| (one more random_vector per particle) and
    dn/dt = W n - (p.S'n) p keeps p.n = 0: d(p.n)/dt = (Wp).n + p.(Wn) +
    (S'p).n - p.(S'n) = 0 for W antisymmetric and S' symmetric."""
    seed = 17
    o, cfg = quiescent(8, 16, 4, seed=seed)
    flags = ORIENT | abi.LAGRANGE_ORIENTATION_SECONDORIENTATION
    o.particles_init(abi.make_particle_config(flags, 8, tau_drag=(1, 0.0, 0.0), aspect_ratio=(1, 4.0, 4.0)))
    a = o.particles()
    d = glibc_stream(seed, 7 * 8).reshape(8, 7)
    t2, f2 = np.arccos(2.0 * d[:, 5] - 1.0), 2 * 3.14159265359 * d[:, 6]
    v = np.stack([np.sin(t2) * np.sin(f2), np.sin(t2) * np.cos(f2), np.cos(t2)], -1)
    p0 = np.stack([a["px"], a["py"], a["pz"]], -1)
    n0 = np.cross(p0, v)
    n0 /= np.linalg.norm(n0, axis=1, keepdims=True)
    assert np.allclose(np.stack([a["nx"], a["ny"], a["nz"]], -1), n0, rtol=0, atol=1e-14)
    a["x"], a["y"], a["z"] = 4.0, 8.0, 2.0
    o.set_particles(a)
    u = _shear(cfg, 0.02)
    for s in range(400):
        r = o.move(u, first=(s == 0))
    p = np.stack([r["px"], r["py"], r["pz"]], -1)
    n = np.stack([r["nx"], r["ny"], r["nz"]], -1)
    assert np.abs((p * n).sum(1)).max() < 1e-4
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-15)
    assert np.abs(n - n0).max() > 0.05  # n did turn


# --------------------------------------------------------------------------
# Particle temperature (LAGRANGE_TEMPERATURE, lagrange.c:579-617, :862-864,
# :1004-1015, :3839-3842, :3915-3918, :4007-4020).
TPART = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS | abi.LAGRANGE_TEMPERATURE


def _thermal(NX=6, NY=8, NZ=4, flags=0, tau_u=0.8, **kw):
    kw.setdefault("tau_t", 0.6)
    return quiescent(NX, NY, NZ, flags=abi.LB_TEMPERATURE | flags, tau=tau_u, **kw)


def test_particle_temperature_ladder_and_start():
    """cp_coeff steps vary slowest, after the orientation ladders (the
    tracer gets 0); every particle starts at T_bot; the log switch gives
    min * step^i, step = (max/min)^(1/(types-1))."""
    o, _ = _thermal(T_bot=1.25, T_top=0.25)
    flags = (TPART | abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_ACTIVE)
    o.particles_init(abi.make_particle_config(flags, 60, tau_drag=(2, 1.0, 3.0), beta_coeff=(1, 0.5, 0.5),
                                              swim_velocity=(2, 0.0, 0.01), cp_coeff=(3, 0.5, 2.0)))
    a = o.particles()
    ntypes = 2 * 2 * 3 + 1
    for r in a:
        t = int(r["name"]) % ntypes
        if t == 0:
            assert r["cp_p"] == 0.0 and r["tau_drag"] == 0.0
            continue
        h = t - 1
        assert r["tau_drag"] == 1.0 + (h % 2) * 2.0
        assert r["swim_velocity"] == ((h // 2) % 2) * 0.01
        assert r["cp_p"] == 0.5 + ((h // 4) % 3) * 0.75
    assert np.all(a["t_p"] == 1.25)
    o2, _ = _thermal()
    o2.particles_init(abi.make_particle_config(TPART | abi.LAGRANGE_TEMPERATURE_INCREMENT_LOG, 9,
                                               tau_drag=(1, 2.0, 2.0), beta_coeff=(1, 0.5, 0.5),
                                               cp_coeff=(3, 0.5, 2.0)))
    assert np.array_equal(o2.particles()["cp_p"][:3], [0.5, 1.0, 2.0])


def _heat_rate(nu, kappa, tau, beta, cp):
    """1 / tau_heat of a small sphere from its heat transfer at Nusselt
    number 2 (no convection): m c_p dT_p/dt = 4 pi a^2 (k / a) (T - T_p),
    i.e. 1/tau_h = 3 kappa rho_f c_f / (a^2 rho_p c_p) with the fluid's
    c_f = 1; the radius from the Stokes response time with added mass,
    a^2 = 3 nu beta tau (lagrange_feedback.c:87), and the density ratio
    rho_p / rho_f = (3 - beta) / (2 beta) (lagrange_feedback.c:84)."""
    a2 = 3.0 * nu * beta * tau
    rho_ratio = (3.0 - beta) / (2.0 * beta)
    return 3.0 * kappa / (a2 * rho_ratio * cp)


@pytest.mark.parametrize("factor", [1.0, 4.0])
def test_particle_temperature_relaxation_rate(factor):
    """In a fluid at t = 0 the particle's temperature decays by exp(-dt /
    tau_h) per sub-step, tau_h from the sphere's heat transfer (not from
    the reference's own expression); in a fluid at constant T0 a particle
    starting at 0 takes f T0 e^{-f} in one sub-step (the scheme's explicit
    source with the exponential factor), and the map's fixed point is
    f T0 / (e^f - 1)."""
    tau_u, tau_t, tau, beta, cp = 0.8, 0.6, 2.5, 0.7, 1.6
    nu, kappa = (tau_u - 0.5) / 3.0, (tau_t - 0.5) / 3.0
    o, cfg = _thermal(tau_u=tau_u, tau_t=tau_t)
    pc = abi.make_particle_config(TPART, 3, tau_drag=(1, tau, tau), beta_coeff=(1, beta, beta),
                                  cp_coeff=(1, cp, cp), time_dt_lagr_factor=factor)
    o.particles_init(pc)
    a = o.particles()
    a["t_p"] = [1.0, -0.5, 3.0]
    o.set_particles(a)
    u = np.zeros((cfg.NZ, cfg.NY, cfg.NX, 3))
    t0 = np.zeros((cfg.NZ, cfg.NY, cfg.NX))
    f_want = _heat_rate(nu, kappa, tau, beta, cp) / factor
    b = o.move(u, t0, first=True)
    f = -np.log(b["t_p"] / np.array([1.0, -0.5, 3.0]))
    assert np.allclose(f, f_want, rtol=1e-12, atol=0)
    # constant fluid temperature T0, particle from 0
    T0 = 0.75
    a = o.particles()
    a["t_p"] = 0.0
    o.set_particles(a)
    b = o.move(u, t0 + T0)
    assert np.allclose(b["t"], T0, rtol=0, atol=1e-15)
    assert np.allclose(b["t_p"], f_want * T0 * math.exp(-f_want), rtol=1e-12)
    assert not np.allclose(b["t_p"], T0 * (1 - math.exp(-f_want)), rtol=1e-6)
    fixed = f_want * T0 / math.expm1(f_want)
    for _ in range(int(35 / f_want)):  # the start is forgotten by e^{-35}
        b = o.move(u, t0 + T0)
    assert np.allclose(b["t_p"], fixed, rtol=1e-9)


def test_tracer_temperature_is_not_a_number():
    """Tracers have tau_drag = 0 and cp_p = 0: the rate is +inf and
    (t_p + inf t) exp(-inf) is NaN, as the reference computes it."""
    o, cfg = _thermal()
    o.particles_init(abi.make_particle_config(TPART | abi.LAGRANGE_ADD_TRACER, 4, tau_drag=(1, 2.0, 2.0),
                                              beta_coeff=(1, 0.5, 0.5), cp_coeff=(1, 1.0, 1.0)))
    o.step(1)
    a = o.particles()
    tracer = a["name"] % 2 == 0  # two types: the tracer and one inertial
    assert np.all(np.isnan(a["t_p"][tracer])) and np.all(np.isfinite(a["t_p"][~tracer]))


def test_reinjected_particle_takes_the_wall_temperature():
    """Crossing the bottom wall re-injects at the top with t_p = T_top,
    crossing the top re-injects at the bottom with t_p = T_bot; the
    relaxation of the same sub-step then acts on that value."""
    tau_u, tau_t, tau, beta, cp = 0.8, 0.6, 4.0, 0.5, 1.0
    nu, kappa = (tau_u - 0.5) / 3.0, (tau_t - 0.5) / 3.0
    o, cfg = _thermal(6, 8, 4, flags=abi.FLUID_BC_Y | abi.TEMPERATURE_BC_Y, T_bot=1.0, T_top=0.0, T_ref=0.0,
                      tau_u=tau_u, tau_t=tau_t)
    flags = (TPART | abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT |
             abi.LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM | abi.LAGRANGE_WRAP)
    o.particles_init(abi.make_particle_config(flags, 3, tau_drag=(1, tau, tau), beta_coeff=(1, beta, beta),
                                              cp_coeff=(1, cp, cp)))
    a = o.particles()
    a["x"], a["y"], a["z"] = [1.0, 2.0, 3.0], [0.05, 7.95, 4.0], [1.0, 2.0, 3.0]
    a["vy"], a["vy_old"] = [-0.2, 0.2, 0.0], [-0.2, 0.2, 0.0]
    a["t_p"] = 0.5
    o.set_particles(a)
    x, y, z = _centres(cfg)
    b = o.move(np.zeros(x.shape + (3,)), 1.0 - y / 8.0, first=True)
    f = _heat_rate(nu, kappa, tau, beta, cp)
    assert b["y"][0] > 7.0 and b["y"][1] < 1.0 and b["y"][2] == 4.0
    assert b["t_p"][0] == pytest.approx((0.0 + f * b["t"][0]) * math.exp(-f), rel=1e-13)
    assert b["t_p"][1] == pytest.approx((1.0 + f * b["t"][1]) * math.exp(-f), rel=1e-13)
    assert b["t_p"][2] == pytest.approx((0.5 + f * b["t"][2]) * math.exp(-f), rel=1e-13)


def test_initial_temperature_from_the_fluid():
    """LAGRANGE_INITIAL_TEMPERATURE_FLUID: t_p (and t, dx_t..dz_t) start as
    the fluid's temperature at the particle; for the linear conduction
    profile that is T_bot + (T_top - T_bot) y / H at every height (the
    fixed-T ghost continues the line)."""
    o, cfg = _thermal(6, 8, 5, flags=abi.FLUID_BC_Y | abi.TEMPERATURE_BC_Y, T_bot=1.0, T_top=0.0, T_ref=0.0)
    W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    x, y, z = _centres(cfg)
    t = 1.0 - y / 8.0
    p = np.broadcast_to(W, (cfg.NZ, cfg.NY, cfg.NX, 19)).copy()
    g = t[..., None] * W
    o.set_populations(p, g)
    o.particles_init(abi.make_particle_config(TPART | abi.LAGRANGE_INITIAL_TEMPERATURE_FLUID, 200,
                                              tau_drag=(1, 2.0, 2.0), beta_coeff=(1, 0.5, 0.5),
                                              cp_coeff=(1, 1.0, 1.0)))
    a = o.particles()
    assert np.allclose(a["t_p"], 1.0 - a["y"] / 8.0, rtol=0, atol=1e-15)
    assert np.array_equal(a["t"], a["t_p"])
    assert np.allclose(a["dy_t"], -1.0 / 8.0, rtol=0, atol=1e-15)
    assert not a["ux"].any() and not a["dx_ux"].any()
    # without the switch: T_bot
    o2, _ = _thermal(6, 8, 5, flags=abi.FLUID_BC_Y | abi.TEMPERATURE_BC_Y, T_bot=1.0, T_top=0.0, T_ref=0.0)
    o2.set_populations(p, g)
    o2.particles_init(abi.make_particle_config(TPART, 5, tau_drag=(1, 2.0, 2.0), beta_coeff=(1, 0.5, 0.5),
                                               cp_coeff=(1, 1.0, 1.0)))
    assert np.all(o2.particles()["t_p"] == 1.0) and not o2.particles()["t"].any()


# --------------------------------------------------------------------------
# Initial particle velocity (lagrange.c:968-991).
def test_initial_velocity_from_the_fluid():
    """LAGRANGE_INITIAL_VELOCITY_FLUID: v starts as the fluid velocity at the
    particle (u and grad u set with it); in a uniform flow that is u0 with a
    zero gradient, for tracers and inertial particles alike."""
    o, cfg = quiescent(8, 6, 4, tau=0.9)
    u0 = np.array([0.021, -0.013, 0.007])
    uniform_flow(o, cfg, u0)
    flags = (abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS |
             abi.LAGRANGE_INITIAL_VELOCITY_FLUID)
    o.particles_init(abi.make_particle_config(flags, 40, tau_drag=(2, 1.0, 5.0), beta_coeff=(1, 0.5, 0.5)))
    a = o.particles()
    for v, u, c in (("vx", "ux", 0), ("vy", "uy", 1), ("vz", "uz", 2)):
        assert np.array_equal(a[v], a[u])
        assert np.allclose(a[v], u0[c], rtol=1e-14, atol=0)
    for g in ("dx_ux", "dy_uy", "dz_uz", "dy_ux"):
        assert np.allclose(a[g], 0.0, atol=1e-15)
    # without the switch the particles start at rest and u is not interpolated
    o2, _ = quiescent(8, 6, 4, tau=0.9)
    uniform_flow(o2, cfg, u0)
    o2.particles_init(abi.make_particle_config(flags & ~abi.LAGRANGE_INITIAL_VELOCITY_FLUID, 40,
                                               tau_drag=(2, 1.0, 5.0), beta_coeff=(1, 0.5, 0.5)))
    b = o2.particles()
    assert not b["vx"].any() and not b["ux"].any()


def test_initial_stokes_velocity_is_the_terminal_velocity():
    """LAGRANGE_INITIAL_VELOCITY_STOKES starts v at -g (1 - beta) tau, the
    settling velocity the particle dynamics reach in a fluid at rest: the
    velocity then stays put, while a particle started at rest is still far
    from it after the same time; with the fluid-velocity start as well, the
    Stokes value wins (it is set second) and u keeps the interpolated fluid
    velocity."""
    g, tau_p, beta = 1e-3, 5.0, 0.6
    flags = abi.LAGRANGE_GRAVITY | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS
    pc = abi.make_particle_config(flags | abi.LAGRANGE_INITIAL_VELOCITY_STOKES, 3, tau_drag=(1, tau_p, tau_p),
                                  beta_coeff=(1, beta, beta))
    o, _ = quiescent(6, 6, 6, gravity_y=g)
    o.particles_init(pc)
    vt = -g * tau_p * (1 - beta)
    assert np.all(o.particles()["vy"] == vt) and not o.particles()["vx"].any()
    o.step(8)
    assert np.allclose(o.particles()["vy"], vt, rtol=1e-12)
    o2, _ = quiescent(6, 6, 6, gravity_y=g)
    o2.particles_init(abi.make_particle_config(flags, 3, tau_drag=(1, tau_p, tau_p), beta_coeff=(1, beta, beta)))
    o2.step(8)
    assert np.all(np.abs(o2.particles()["vy"] - vt) > 0.1 * abs(vt))
    o3, cfg = quiescent(6, 6, 6, gravity_y=g)
    uniform_flow(o3, cfg, np.array([0.01, 0.0, 0.0]))
    o3.particles_init(abi.make_particle_config(flags | abi.LAGRANGE_INITIAL_VELOCITY_STOKES |
                                               abi.LAGRANGE_INITIAL_VELOCITY_FLUID, 3, tau_drag=(1, tau_p, tau_p),
                                               beta_coeff=(1, beta, beta)))
    c = o3.particles()
    assert np.all(c["vy"] == vt) and not c["vx"].any() and np.allclose(c["ux"], 0.01, rtol=1e-14)


# --------------------------------------------------------------------------
# Variable gravity (LAGRANGE_GRAVITY_VARIABLE, lagrange.c:347-384,
# :3314-3335, :3824-3827, :3905-3908).
GVAR = abi.LAGRANGE_GRAVITY | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS | abi.LAGRANGE_GRAVITY_VARIABLE


def test_gravity_coeff_ladder_order():
    """gravity_coeff steps come after beta and before the orientation and
    cp ladders; the tracer gets 0; the log switch gives min * step^i."""
    o, _ = _thermal()
    flags = GVAR | abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_TEMPERATURE
    o.particles_init(abi.make_particle_config(flags, 80, tau_drag=(2, 1.0, 3.0), beta_coeff=(1, 0.5, 0.5),
                                              gravity_coeff=(3, 0.5, 2.0), cp_coeff=(2, 1.0, 2.0)))
    ntypes = 2 * 3 * 2 + 1
    for r in o.particles():
        t = int(r["name"]) % ntypes
        if t == 0:
            assert r["gravity_coeff"] == 0.0
            continue
        h = t - 1
        assert r["tau_drag"] == 1.0 + (h % 2) * 2.0
        assert r["gravity_coeff"] == 0.5 + ((h // 2) % 3) * 0.75
        assert r["cp_p"] == 1.0 + ((h // 6) % 2) * 1.0
    o2, _ = quiescent(6, 6, 6)
    o2.particles_init(abi.make_particle_config(GVAR | abi.LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG, 3,
                                               tau_drag=(1, 1.0, 1.0), beta_coeff=(1, 0.5, 0.5),
                                               gravity_coeff=(3, 0.5, 2.0)))
    assert np.array_equal(o2.particles()["gravity_coeff"], [0.5, 1.0, 2.0])


def test_variable_gravity_settling_and_reinjection():
    """Each type settles at -c g tau (1 - beta): the factor c scales the
    weight and the added-mass buoyancy alike (with c on one of them only the
    velocity would be -g tau (c - beta) or -g tau (1 - c beta)); a
    re-injected particle takes the Stokes velocity of its own c."""
    g, tau_p, beta = 1e-3, 5.0, 0.6
    o, _ = quiescent(6, 6, 6, gravity_y=g)
    o.particles_init(abi.make_particle_config(GVAR, 6, tau_drag=(1, tau_p, tau_p), beta_coeff=(1, beta, beta),
                                              gravity_coeff=(3, 0.5, 2.0)))
    o.step(400)
    a = o.particles()
    assert np.allclose(a["vy"], -a["gravity_coeff"] * g * tau_p * (1 - beta), rtol=1e-12)
    o2, _ = quiescent(6, 8, 4, flags=abi.FLUID_BC_Y, gravity_y=g)
    flags = GVAR | abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT | \
        abi.LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM | abi.LAGRANGE_WRAP
    o2.particles_init(abi.make_particle_config(flags, 3, tau_drag=(1, tau_p, tau_p), beta_coeff=(1, beta, beta),
                                               gravity_coeff=(3, 0.5, 2.0)))
    b = o2.particles()
    b["y"], b["vy"], b["vy_old"] = 0.05, -0.2, -0.2
    o2.set_particles(b)
    o2.step(1)
    c = o2.particles()
    assert np.all(c["y"] > 7.0)
    assert np.array_equal(c["vy"], -g * c["gravity_coeff"] * (1 - beta) * tau_p)


# --------------------------------------------------------------------------
# Eulerian probes and the initial-position switches (lagrange.c:280-282,
# :774-814, :3024-3033).
def test_static_probes_stay_and_record_the_flow():
    """LAGRANGE_ADD_TRACER_STATIC: type 0 (tau_drag -1) is a probe that
    keeps its position and records the fluid velocity as its own; type 1
    is the tracer, which moves with the flow."""
    o, cfg = quiescent(8, 6, 4, tau=0.9)
    u0 = np.array([0.02, -0.01, 0.005])
    uniform_flow(o, cfg, u0)
    flags = abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_ADD_TRACER_STATIC | abi.LAGRANGE_WRAP
    o.particles_init(abi.make_particle_config(flags, 12, tau_drag=(2, 1.0, 4.0)))
    a0 = o.particles()
    ntypes = 2 + 2
    kind = a0["name"].astype(int) % ntypes
    assert np.all(a0["tau_drag"][kind == 0] == -1.0) and np.all(a0["tau_drag"][kind == 1] == 0.0)
    assert set(a0["tau_drag"][kind >= 2]) == {1.0, 4.0}
    o.step(10)
    a = o.particles()
    probe, tracer = kind == 0, kind == 1
    for c, v, u in (("x", "vx", 0), ("y", "vy", 1), ("z", "vz", 2)):
        assert np.array_equal(a[c][probe], a0[c][probe])
        assert np.allclose(a[v][probe], u0[u], rtol=1e-13)
        d = np.abs(a[c][tracer] - np.mod(a0[c][tracer] + 10 * u0[u], (8, 6, 4)[u]))
        assert np.all(d < 1e-12)


def test_initial_grid_positions():
    """LAGRANGE_INITIAL_GRID with one particle per node: every particle on
    its own cell centre, x fastest; the position draws are still consumed
    (the next draws of the stream follow them)."""
    o, cfg = quiescent(4, 3, 2, seed=11)
    o.particles_init(abi.make_particle_config(0, 24, tau_drag=(1, 2.0, 2.0), flags2=abi.LAGRANGE_INITIAL_GRID))
    a = o.particles()
    i = np.arange(24)
    assert np.array_equal(a["x"], i % 4 + 0.5)
    assert np.array_equal(a["y"], (i // 4) % 3 + 0.5)
    assert np.array_equal(a["z"], (i // 12) % 2 + 0.5)
    # a coarser grid: 8 particles on 4x3x2 nodes -> spacing 3^(1/3) (integer 24 // 8)
    o2, _ = quiescent(4, 3, 2, seed=11)
    o2.particles_init(abi.make_particle_config(0, 8, tau_drag=(1, 2.0, 2.0), flags2=abi.LAGRANGE_INITIAL_GRID))
    b = o2.particles()
    dp = 3.0 ** (1.0 / 3.0)
    nx, ny, nz = int(4 // dp), int(3 // dp), max(int(2 // dp), 1)
    j = np.arange(8)
    assert np.allclose(b["x"], dp / 2 + (j % nx) * dp, rtol=1e-15)
    assert np.allclose(b["y"], dp / 2 + ((j // nx) % ny) * dp, rtol=1e-15)
    assert np.allclose(b["z"], dp / 2 + ((j // (nx * ny)) % nz) * dp, rtol=1e-15)


def test_initial_pairs_at_unit_distance():
    """LAGRANGE_INITIAL_PAIRS: particle 2k-1 sits at distance 1 from particle
    2k (k >= 1), on the box side of it; each pair takes two extra draws after
    the position draws of its second particle."""
    seed = 5
    o, cfg = quiescent(16, 12, 8, seed=seed)
    n = 21
    o.particles_init(abi.make_particle_config(0, n, tau_drag=(1, 2.0, 2.0), flags2=abi.LAGRANGE_INITIAL_PAIRS))
    a = o.particles()
    pos = np.column_stack([a["x"], a["y"], a["z"]])
    for k in range(2, n, 2):
        assert np.linalg.norm(pos[k - 1] - pos[k]) == pytest.approx(1.0, abs=1e-14)
        assert np.all(pos[k - 1] >= 0) and np.all(pos[k - 1] < (16, 12, 8))
    # draws: 3 for particle 0, 3 for particle 1 (overwritten), 3 + 2 for particle 2, ...
    r = glibc_stream(seed, 3 * n + 2 * ((n - 1) // 2))
    assert a["x"][0] == 16 * r[0] and a["x"][2] == 16 * r[6] and a["x"][4] == 16 * r[6 + 5 + 3]


# --------------------------------------------------------------------------
# Lift and wake drag of the added-mass model (lagrange.c:3242-3264, :3365-3380).
def _accel(flags2, u_field, v, tau=4.0, beta=0.6, nu_tau=0.8):
    o, cfg = quiescent(8, 8, 8, tau=nu_tau)
    flags = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS
    o.particles_init(abi.make_particle_config(flags, 1, tau_drag=(1, tau, tau), beta_coeff=(1, beta, beta),
                                              flags2=flags2))
    a = o.particles()
    a["x"], a["y"], a["z"] = 3.3, 4.2, 3.7
    a["vx"], a["vy"], a["vz"] = v
    o.set_particles(a)
    r = o.move(u_field(cfg), None, first=True)[0]
    return np.array([r["ax"], r["ay"], r["az"]]), r, (nu_tau - 0.5) / 3.0


def test_lift_is_the_slip_cross_vorticity():
    """LAGRANGE_ADDEDMASS_LIFT adds (beta/3) (u - v) x omega: in the shear
    u = (gamma (y - 4), 0, 0) the vorticity is (0, 0, -gamma)."""
    gamma, beta = 2e-3, 0.6

    def shear(cfg):
        x, y, z = _centres(cfg)
        u = np.zeros(x.shape + (3,))
        u[..., 0] = gamma * (y - 4.0)
        return u
    v = (0.001, -0.0007, 0.0004)
    a_l, r, _ = _accel(abi.LAGRANGE_ADDEDMASS_LIFT, shear, v, beta=beta)
    a_0, _, _ = _accel(0, shear, v, beta=beta)
    slip = np.array([r["ux"], r["uy"], r["uz"]]) - np.array(v)
    omega = np.array([0.0, 0.0, -gamma])
    assert np.allclose(a_l - a_0, beta / 3.0 * np.cross(slip, omega), rtol=1e-10, atol=1e-22)
    assert abs((a_l - a_0)[1]) > 0


def test_wake_drag_follows_schiller_naumann():
    """LAGRANGE_ADDEDMASS_WAKEDRAG adds 0.15 Re_p^0.687 (u - v) / tau, with
    the particle Reynolds number of the slip velocity and the diameter of
    the Stokes particle, d = 2 sqrt(3 beta nu tau)."""
    u0 = np.array([0.03, -0.01, 0.02])

    def uniform(cfg):
        return np.broadcast_to(u0, (cfg.NZ, cfg.NY, cfg.NX, 3)).copy()
    v = (0.0, 0.004, -0.002)
    tau, beta = 4.0, 0.6
    a_w, r, nu = _accel(abi.LAGRANGE_ADDEDMASS_WAKEDRAG, uniform, v, tau=tau, beta=beta)
    a_0, _, _ = _accel(0, uniform, v, tau=tau, beta=beta)
    slip = u0 - np.array(v)
    d = 2.0 * math.sqrt(3.0 * beta * nu * tau)
    re = np.linalg.norm(slip) * d / nu
    assert np.allclose(a_w - a_0, 0.15 * re ** 0.687 * slip / tau, rtol=1e-12)


# --------------------------------------------------------------------------
# Stokes drag of a spheroid (LAGRANGE_ORIENTATION_DRAG, lagrange.c:3267-3310).
def _happel_brenner(alpha):
    """Resistance of a spheroid with semi-axes a_par = alpha a_perp (along
    the axis) and a_perp, across and along the axis, over that of the
    sphere of radius a_perp (Happel & Brenner, Low Reynolds number
    hydrodynamics, the prolate and oblate translation formulas)."""
    if alpha > 1:
        e = math.sqrt(1 - 1 / alpha ** 2)
        L = math.log((1 + e) / (1 - e))
        across = 32 * alpha * e ** 3 / (6 * (2 * e + (3 * e * e - 1) * L))
        along = 16 * alpha * e ** 3 / (6 * ((1 + e * e) * L - 2 * e))
    else:
        e = math.sqrt(1 - alpha * alpha)
        s = math.asin(e)
        across = 16 * e ** 3 / (6 * ((1 + 2 * e * e) * s - e * math.sqrt(1 - e * e)))
        along = 8 * e ** 3 / (6 * (e * math.sqrt(1 - e * e) - (1 - 2 * e * e) * s))
    return across, along


@pytest.mark.parametrize("alpha", [4.0, 1.0, 0.4])
def test_spheroid_drag_matches_happel_brenner(alpha):
    """In a uniform flow a spheroid at rest takes the acceleration
    c (u - v) / (tau alpha), c the translational resistance along or across
    its axis relative to the sphere of radius a_perp (tau_drag is that
    sphere's response time, the mass grows with alpha)."""
    tau = 4.0
    u0 = 0.02
    flags = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY
    across, along = _happel_brenner(alpha) if alpha != 1.0 else (1.0, 1.0)
    for p_axis, c in (((1.0, 0.0, 0.0), along), ((0.0, 1.0, 0.0), across), ((0.0, 0.0, 1.0), across)):
        o, cfg = quiescent(8, 8, 8, tau=0.8)
        o.particles_init(abi.make_particle_config(flags, 1, tau_drag=(1, tau, tau), aspect_ratio=(1, alpha, alpha),
                                                  flags2=abi.LAGRANGE_ORIENTATION_DRAG))
        a = o.particles()
        a["x"], a["y"], a["z"] = 3.3, 4.2, 3.7
        a["px"], a["py"], a["pz"] = p_axis
        a["vx"] = a["vy"] = a["vz"] = 0.0
        o.set_particles(a)
        u = np.zeros((8, 8, 8, 3))
        u[..., 0] = u0
        r = o.move(u, None, first=True)[0]
        assert r["ax"] == pytest.approx(c * u0 / (tau * alpha), rel=1e-13)
        assert abs(r["ay"]) < 1e-18 and abs(r["az"]) < 1e-18


# --------------------------------------------------------------------------
# Radius and density parametrisation (lagrange.c:150-260, :3775-3788).
RAD = abi.LAGRANGE_RADIUSANDDENSITY


def test_radius_density_give_the_stokes_response_time():
    """tau_drag and beta_coeff from the radius and density ladders (radius
    fastest, then density; the tracer has r = 0): without added mass the
    Stokes time 2 rho_p r^2 / (9 mu); with it, beta = 3 rho_f / (rho_f +
    2 rho_p) and the added-mass time 2 (rho_p + rho_f / 2) r^2 / (9 mu),
    rho_f = 1, mu = nu."""
    nu = (0.8 - 0.5) / 3.0
    for flags in (abi.LAGRANGE_ADD_TRACER, abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS):
        o, _ = quiescent(6, 6, 6, tau=0.8)
        o.particles_init(abi.make_particle_config(flags, 10, particle_radius=(2, 0.5, 1.5),
                                                  particle_density=(2, 1.0, 4.0), flags2=RAD))
        a = o.particles()
        for r in a:
            t = int(r["name"]) % 5
            if t == 0:
                assert r["tau_drag"] == 0.0
                continue
            rad, rho = (0.5, 1.5)[(t - 1) % 2], (1.0, 4.0)[((t - 1) // 2) % 2]
            if flags & abi.LAGRANGE_ADDEDMASS:
                assert r["beta_coeff"] == pytest.approx(3.0 / (1.0 + 2.0 * rho), rel=1e-15)
                assert r["tau_drag"] == pytest.approx(2 * (rho + 0.5) * rad ** 2 / (9 * nu), rel=1e-14)
            else:
                assert r["tau_drag"] == pytest.approx(2 * rho * rad ** 2 / (9 * nu), rel=1e-14)


def test_walls_at_the_particle_radius():
    """LAGRANGE_BC_RADIUSandDENSITY: an elastic particle of radius r that
    ends a step at height y < r is mirrored about y = r (its centre cannot
    come closer to the wall than its radius)."""
    o, cfg = quiescent(6, 8, 4, flags=abi.FLUID_BC_Y, tau=0.8)
    flags = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS
    o.particles_init(abi.make_particle_config(flags, 1, particle_radius=(1, 0.5, 0.5),
                                              particle_density=(1, 2.0, 2.0),
                                              flags2=RAD | abi.LAGRANGE_BC_RADIUSANDDENSITY))
    a = o.particles()
    a["x"], a["y"], a["z"] = 2.0, 0.55, 2.0
    a["vy"] = -0.3
    o.set_particles(a)
    u = np.zeros((4, 8, 6, 3))
    r = o.move(u, None, first=True)[0]
    assert r["y"] == pytest.approx(2 * 0.5 - (0.55 - 0.3), abs=1e-12)
    assert r["vy"] > 0


# --------------------------------------------------------------------------
# Swimming modes of LAGRANGE_ORIENTATION_ACTIVE (lagrange.c:3134-3186).
ACTIVE = (abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY |
          abi.LAGRANGE_ORIENTATION_ACTIVE | abi.LAGRANGE_ORIENTATION_SECONDORIENTATION)


def _swimmer_velocity(mode, aspect=2.0, t_fields=(None,)):
    o, cfg = quiescent(8, 8, 8, tau=0.8, flags=abi.LB_TEMPERATURE if t_fields[0] is not None else 0)
    pc = abi.make_particle_config(ACTIVE, 1, tau_drag=(1, 0.0, 0.0), aspect_ratio=(1, aspect, aspect),
                                  swim_velocity=(1, 0.01, 0.01), flags2=mode)
    o.particles_init(pc)
    a = o.particles()
    a["x"], a["y"], a["z"] = 3.3, 4.2, 3.7
    a["px"], a["py"], a["pz"] = 1.0, 0.0, 0.0
    a["nx"], a["ny"], a["nz"] = 0.0, 0.0, 1.0
    o.set_particles(a)
    u = np.zeros((8, 8, 8, 3))
    u[..., 1] = 0.002
    out = []
    for k, tf in enumerate(t_fields):
        t = None if tf is None else np.full((8, 8, 8), tf)
        r = o.move(u, t, first=(k == 0))[0]
        out.append(np.array([r["vx"], r["vy"], r["vz"]]))
    return out


def test_swimming_modes():
    """Default: u + s p; BALLISTIC: s p alone; SECONDORIENTATION: u + s n;
    SELECTIVEORIENTATION: along p for rods (aspect >= 1), along n for
    disks; TEMPERATURE: u + s p only while the particle's temperature rises
    (dt_t > 0)."""
    s, u = 0.01, np.array([0.0, 0.002, 0.0])
    p, n = np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0])
    assert np.allclose(_swimmer_velocity(0)[0], u + s * p, rtol=1e-14, atol=1e-18)
    assert np.allclose(_swimmer_velocity(abi.LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC)[0], s * p, rtol=1e-14,
                       atol=1e-18)
    assert np.allclose(_swimmer_velocity(abi.LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION)[0], u + s * n,
                       rtol=1e-14, atol=1e-18)
    sel = abi.LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION
    assert np.allclose(_swimmer_velocity(sel, aspect=2.0)[0], u + s * p, rtol=1e-14, atol=1e-18)
    assert np.allclose(_swimmer_velocity(sel, aspect=0.5)[0], u + s * n, rtol=1e-14, atol=1e-18)
    # temperature: first sub-step keeps dt_t = 0, then warming, then cooling
    v = _swimmer_velocity(abi.LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE, t_fields=(0.5, 0.6, 0.55))
    assert np.allclose(v[0], u, rtol=1e-14, atol=1e-18)
    assert np.allclose(v[1], u + s * p, rtol=1e-14, atol=1e-18)
    assert np.allclose(v[2], u, rtol=1e-14, atol=1e-18)


# Copepod jumps (LAGRANGE_ORIENTATION_ACTIVE_JUMP, lagrange.c:3048-3132).
def _jumper(flags2, gamma, csr, jt, steps, t_field=None, seed=7):
    """One swimming tracer with p along z, the vorticity axis of the shear
    u_x = gamma (y - 4), which Jeffery's equation leaves fixed; returns the
    initial records and those after each sub-step."""
    o, cfg = quiescent(8, 8, 8, tau=0.8, seed=seed, flags=abi.LB_TEMPERATURE if t_field is not None else 0)
    pc = abi.make_particle_config(ACTIVE & ~abi.LAGRANGE_ORIENTATION_SECONDORIENTATION, 1, tau_drag=(1, 0.0, 0.0),
                                  aspect_ratio=(1, 2.0, 2.0), swim_velocity=(1, 0.01, 0.01),
                                  critical_shear_rate=(1, csr, csr), jump_time=(1, jt, jt),
                                  flags2=abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP | flags2)
    o.particles_init(pc)
    a0 = o.particles()
    a = a0.copy()
    a["x"], a["y"], a["z"] = 3.3, 4.2, 3.7
    a["px"], a["py"], a["pz"] = 0.0, 0.0, 1.0
    a["time_from_jump"] = 100.0 * jt
    o.set_particles(a)
    u = np.zeros((8, 8, 8, 3))
    u[..., 0] = gamma * (np.arange(8.0)[None, :, None] - 4.0)
    t = None if t_field is None else np.full((8, 8, 8), t_field)
    return a0, [o.move(u, t, first=(k == 0))[0] for k in range(steps)]


def test_jump_timing_and_decay():
    """The first draw after the orientation sets time_from_jump = 10 r
    jump_time; the shear rate sqrt(2 S:S) of a simple shear is gamma; above
    the critical rate the tracer jumps along p with the swim speed, the
    amplitude then decays as exp(-t / jump_time) until t exceeds
    -jump_time ln 0.01, when the next jump starts; below it nothing
    happens; HIGHPASS jumps below the critical rate instead; with
    JUMP_TEMPERATURE the fluid temperature takes the shear rate's place."""
    s, jt, gamma = 0.01, 2.0, 1e-3
    a0, rs = _jumper(0, gamma, 5e-4, jt, 24, seed=11)
    assert a0["jump_time"][0] == jt and a0["critical_shear_rate"][0] == 5e-4
    assert a0["time_from_jump"][0] == 10.0 * glibc_stream(11, 6)[5] * jt
    assert a0["px_jump"][0] == a0["px"][0] and a0["pz_jump"][0] == a0["pz"][0]
    dt = rs[0]["time_from_jump"]
    assert dt == 1.0
    assert rs[0]["shear_rate"] == pytest.approx(gamma, rel=1e-12)
    vz = np.array([r["vz"] for r in rs])
    dur = -jt * np.log(0.01)
    want, tfj = [], 100.0 * jt
    for k in range(24):
        if tfj > dur:
            want.append(s)
            tfj = 0.0
        else:
            want.append(s * np.exp(-tfj / jt))
        tfj += dt
    assert np.allclose(vz, want, rtol=1e-14, atol=0)
    assert want[10] == s and want[9] == pytest.approx(s * np.exp(-4.5))  # 9 dt <= dur < 10 dt
    assert np.array_equal([r["pz_jump"] for r in rs], np.ones(24))
    # below the critical rate: no jump; HIGHPASS inverts the test
    _, quiet = _jumper(0, gamma, 2e-3, jt, 3)
    assert all(r["vz"] == 0.0 for r in quiet)
    _, hp = _jumper(abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS, gamma, 2e-3, jt, 2)
    assert hp[0]["vz"] == s and hp[1]["vz"] == pytest.approx(s * np.exp(-1.0 / jt), rel=1e-14)
    _, hp = _jumper(abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS, gamma, 5e-4, jt, 2)
    assert hp[0]["vz"] == 0.0
    # temperature as the stimulus
    T = abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE
    assert _jumper(T, gamma, 0.3, jt, 1, t_field=0.5)[1][0]["vz"] == s
    assert _jumper(T, gamma, 0.6, jt, 1, t_field=0.5)[1][0]["vz"] == 0.0


def test_sediment_counts_bottom_depositions():
    """LAGRANGE_BC_INELASTIC_SEDIMENT (no nucleation): each crossing of the
    bottom wall adds one deposition, a crossing of the top wall none."""
    o, cfg = quiescent(6, 8, 4, flags=abi.FLUID_BC_Y)
    flags = abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT | \
        abi.LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM | abi.LAGRANGE_WRAP
    o.particles_init(abi.make_particle_config(flags, 2, tau_drag=(1, 1e9, 1e9),
                                              flags2=abi.LAGRANGE_BC_INELASTIC_SEDIMENT))
    a = o.particles()
    a["x"], a["z"] = 2.0, 2.0
    a["y"], a["vy"], a["vy_old"] = [0.05, 7.95], [-0.2, 0.2], [-0.2, 0.2]
    o.set_particles(a)
    u = np.zeros((4, 8, 6, 3))
    b = o.move(u, None, first=True)
    assert b["y"][0] > 7.0 and b["y"][1] < 1.0
    assert list(b["sediment"]) == [1.0, 0.0]


# --------------------------------------------------------------------------
# Random reorientation (LAGRANGE_ORIENTATION_RANDOM, lagrange.c:3698-3713).
def test_random_reorientation_draws_and_isotropy():
    """Every sub-step each particle's p is a new random_vector from the next
    two draws of the stream, in particle order; dp/dt is the change from
    the previous p over the sub-step; the directions are uniform on the
    sphere (mean p_z^2 = 1/3, which theta uniform in [0, pi] would make
    1/2)."""
    seed = 21
    o, cfg = quiescent(8, 8, 8, seed=seed)
    flags = abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ORIENTATION
    n = 5
    o.particles_init(abi.make_particle_config(flags, n, tau_drag=(1, 2.0, 2.0),
                                              flags2=abi.LAGRANGE_ORIENTATION_RANDOM))
    a0 = o.particles()
    o.step(1)
    a = o.particles()
    r = glibc_stream(seed, 5 * n + 2 * n)[5 * n:].reshape(n, 2)
    theta, phi = np.arccos(2.0 * r[:, 0] - 1.0), 2 * 3.14159265359 * r[:, 1]
    v = np.column_stack([np.sin(theta) * np.sin(phi), np.sin(theta) * np.cos(phi), np.cos(theta)])
    p = np.column_stack([a["px"], a["py"], a["pz"]])
    assert np.allclose(p, v, rtol=0, atol=1e-15)
    dp = np.column_stack([a["dt_px"], a["dt_py"], a["dt_pz"]])
    p0 = np.column_stack([a0["px"], a0["py"], a0["pz"]])
    assert np.allclose(dp, v - p0, rtol=0, atol=1e-14)
    o2, _ = quiescent(16, 16, 16, seed=3)
    o2.particles_init(abi.make_particle_config(flags, 20000, tau_drag=(1, 2.0, 2.0),
                                               flags2=abi.LAGRANGE_ORIENTATION_RANDOM))
    o2.step(1)
    b = o2.particles()
    assert abs(np.mean(b["pz"] ** 2) - 1.0 / 3.0) < 0.01
    assert abs(np.mean(b["px"])) < 0.02 and abs(np.mean(b["py"])) < 0.02


# --------------------------------------------------------------------------
# Polymer conformation tensor (LAGRANGE_POLYMER, polymer.c:150-240).
def test_polymer_conformation_steady_shear_and_relaxation():
    """The reference's evolution dc/dt = A^T c + c A + (I - c) / tau_p with
    A[i][j] = d_j u_i: in the shear u_x = g y its fixed point is c_xx = 1,
    c_xy = g tau_p, c_yy = 1 + 2 (g tau_p)^2, c_zz = 1 (the transpose of the
    upper-convected form, whose c_xx would grow instead), reached by the
    Adams-Bashforth steps; at rest c - I decays as exp(-t / tau_p) to the
    steps' second order.  The initial diagonal is a random_vector of two
    draws after the position draws."""
    g, tau_p = 0.01, 5.0
    flags = abi.LAGRANGE_GRADIENT
    o, cfg = quiescent(8, 8, 8, seed=13)
    o.particles_init(abi.make_particle_config(flags, 2, tau_drag=(1, 0.0, 0.0), flags2=abi.LAGRANGE_POLYMER,
                                              tau_polymer=tau_p))
    a = o.particles()
    r = glibc_stream(13, 10).reshape(2, 5)[:, 3:]
    th, ph = np.arccos(2 * r[:, 0] - 1), 2 * 3.14159265359 * r[:, 1]
    assert np.allclose(a["cxx"], np.sin(th) * np.sin(ph), atol=1e-15)
    assert np.allclose(a["czz"], np.cos(th), atol=1e-15) and not a["cxy"].any()
    a["x"], a["y"], a["z"] = [2.0, 5.0], [4.3, 3.6], [4.0, 4.0]
    o.set_particles(a)
    x, y, z = _centres(cfg)
    u = np.zeros(x.shape + (3,))
    u[..., 0] = g * (y - 4.0)
    for k in range(400):
        b = o.move(u, None, first=(k == 0))
    gt = g * tau_p
    assert np.allclose(b["cxx"], 1.0, rtol=1e-9)
    assert np.allclose(b["cxy"], gt, rtol=1e-9)
    assert np.allclose(b["cyy"], 1.0 + 2 * gt * gt, rtol=1e-9)
    assert np.allclose(b["czz"], 1.0, rtol=1e-9) and np.allclose(b["cxz"], 0.0, atol=1e-12)
    # at rest, tau_p = 50 lattice steps: c - I decays as exp(-t / tau_p) (the
    # first step's Adams-Bashforth start is off by 1/2 of dt / tau_p)
    tau_p = 50.0
    o2, cfg2 = quiescent(8, 8, 8, seed=13)
    o2.particles_init(abi.make_particle_config(flags, 2, tau_drag=(1, 0.0, 0.0), flags2=abi.LAGRANGE_POLYMER,
                                               tau_polymer=tau_p))
    c0 = o2.particles()
    u0 = np.zeros(x.shape + (3,))
    n = 100
    for k in range(n):
        c = o2.move(u0, None, first=(k == 0))
    for f in ("cxx", "cyy", "czz"):
        assert np.allclose(c[f] - 1.0, (c0[f] - 1.0) * math.exp(-n / tau_p), rtol=2e-2)
