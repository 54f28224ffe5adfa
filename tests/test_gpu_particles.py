"""GPU parity of the Lagrangian particles: the HIP kernels through the C-ABI
against the oracle's restatement of lagrange.c / lagrange_feedback.c.

FP64 lattices: every particle field, and with two-way coupling the fluid
populations, must be bit-identical to the oracle (both evaluate the
reference's expressions in its order without contraction; the feedback is
summed per node in particle order, as the reference's sequential loop does).
FP32-deviation lattice: particle positions within 1e-6 (absolute, lattice
units) of the FP64 oracle after 100 steps, velocities within 1e-5 of the
largest particle velocity component.
"""
import numpy as np
import pytest

from ch4g import abi, Lattice
from oracle import Oracle
import configs

pytestmark = pytest.mark.gpu

FIELDS = abi.PARTICLE_FIELDS


def rb_small(NX=64, NY=32, precision=abi.FP64, g=1e-3, seed=2024):
    """Rayleigh-Benard at Ra 1e5 with a realistic particle gravity: beta_t g
    is kept, g itself set to `g`."""
    c = configs.rb2d(NX, NY, 1, Ra=1e5, precision=precision, seed=seed)
    return c.copy(gravity_y=g, beta_t=c.beta_t * c.gravity_y / g)


def pair(cfg, pcfg, steps, records=None):
    o = Oracle(cfg)
    o.initial_conditions()
    o.particles_init(pcfg)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.particles_init(pcfg)
    if records is not None:
        o.set_particles(records)
        lat.set_particles(records)
    o.step(steps)
    lat.step(steps)
    return o, lat


def compare_particles(o, lat, exact=True, pos_atol=0.0, rel=0.0):
    a, b = o.particles(), lat.particles()
    assert len(a) == len(b), (len(a), len(b))
    bad = []
    for f in FIELDS:
        x, y = a[f], b[f]
        if exact:  # NaN where the reference's arithmetic gives NaN (tracers' t_p)
            if not np.array_equal(x, y, equal_nan=True):
                i = int(np.argmax(x != y))
                bad.append(f"{f}[{i}] oracle {x[i]!r} gpu {y[i]!r}")
    assert not bad, bad[:8]
    if not exact:
        for f in ("x", "y", "z"):
            d = np.abs(a[f] - b[f])
            S = {"x": lat.cfg.NX, "y": lat.cfg.NY, "z": lat.cfg.NZ}[f]
            d = np.minimum(d, S - d)
            assert d.max() <= pos_atol, (f, d.max())
        # velocity error relative to the largest particle speed component
        s = max(np.abs(a[f]).max() for f in ("vx", "vy", "vz"))
        for f in ("vx", "vy", "vz"):
            err = np.abs(a[f] - b[f]).max() / (s if s > 0 else 1.0)
            print(f"  {f}: max error / max|v| = {err:.3e}")
            assert err <= rel, f


def compare_fluid(o, lat):
    po, go = o.get_populations()
    pg, gg = lat.get_populations()
    assert np.array_equal(po, pg)
    if lat.cfg.has(abi.LB_TEMPERATURE):
        assert np.array_equal(go, gg)


HEAVY = abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS | \
    abi.LAGRANGE_GRAVITY


def test_rb_tracers_and_inertial_particles_fp64():
    pc = abi.make_particle_config(HEAVY, 400, tau_drag=(3, 2.0, 8.0), beta_coeff=(2, 0.6, 1.2))
    o, lat = pair(rb_small(), pc, 200)
    print("particles alive", lat.particles_count())
    compare_particles(o, lat)
    compare_fluid(o, lat)


def test_hit_les_particles_fp64():
    cfg = configs.hit(16, tau=0.51, amp=1e-4)
    flags = abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS
    pc = abi.make_particle_config(flags, 500, tau_drag=(4, 2.0, 10.0), beta_coeff=(1, 0.0, 0.0))
    o, lat = pair(cfg, pc, 100)
    compare_particles(o, lat)


@pytest.mark.parametrize("extra", [
    0,
    abi.LAGRANGE_BC_INELASTIC,
    abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT,
    abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT |
    abi.LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF,
    abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT | abi.LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM,
    abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT |
    abi.LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO,
    abi.LAGRANGE_SMALLTAUD_BETA,
    abi.LAGRANGE_SMALLTAUD_FLOATER,
    (abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT, abi.LAGRANGE_BC_INELASTIC_SEDIMENT),
], ids=["elastic", "inelastic", "reinject", "reinject_half", "reinject_norandom", "reinject_vzero",
        "smalltaud_beta", "smalltaud_floater", "reinject_sediment"])
def test_wall_rules_and_small_tau_fp64(extra):
    # strong settling so that particles reach the walls within the run
    extra, flags2 = extra if isinstance(extra, tuple) else (extra, 0)
    pc = abi.make_particle_config(HEAVY | extra, 256, tau_drag=(2, 2.0, 20.0), beta_coeff=(2, 0.2, 0.5),
                                  flags2=flags2)
    o, lat = pair(rb_small(32, 16, g=5e-3), pc, 120)
    compare_particles(o, lat)
    if flags2:
        assert lat.particles()["sediment"].max() >= 1


def test_twoway_momentum_fp64():
    pc = abi.make_particle_config(HEAVY | abi.LAGRANGE_TWOWAY_MOMENTUM, 300, tau_drag=(2, 1.0, 4.0),
                                  beta_coeff=(2, 0.5, 0.8))
    o, lat = pair(rb_small(48, 24, g=2e-3), pc, 80)
    compare_particles(o, lat)
    compare_fluid(o, lat)


def test_twoway_smalltaud_feedback_fp64():
    flags = HEAVY | abi.LAGRANGE_TWOWAY_MOMENTUM | abi.LAGRANGE_TWOWAY_SMALLTAUD | abi.LAGRANGE_SMALLTAUD_BETA
    pc = abi.make_particle_config(flags, 200, tau_drag=(2, 1.0, 4.0), beta_coeff=(1, 0.5, 0.5))
    o, lat = pair(rb_small(32, 16, g=2e-3), pc, 50)
    compare_particles(o, lat)
    compare_fluid(o, lat)


ORIENT = (abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ORIENTATION |
          abi.LAGRANGE_ORIENTATION_JEFFREY | abi.LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG |
          abi.LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS | abi.LAGRANGE_ORIENTATION_ACTIVE |
          abi.LAGRANGE_ORIENTATION_SECONDORIENTATION)


@pytest.mark.parametrize("case", ["hit", "rb_gravity"])
def test_orientation_fp64(case):
    """Jeffery orientation with aspect-ratio, gyrotaxis and swimming ladders:
    p, dp/dt and the particles they move bit-identical to the oracle."""
    if case == "hit":
        cfg, flags = configs.hit(16, tau=0.51, amp=1e-4), ORIENT
    else:
        cfg, flags = rb_small(), ORIENT | abi.LAGRANGE_GRAVITY | abi.LAGRANGE_ADDEDMASS | abi.LAGRANGE_ORIENTATION_BC
    pc = abi.make_particle_config(flags, 300, tau_drag=(2, 0.0, 4.0), beta_coeff=(1, 0.8, 0.8),
                                  aspect_ratio=(3, 0.5, 8.0), gyrotaxis_velocity=(2, 0.5, 2.0),
                                  swim_velocity=(2, 0.0, 0.002), time_dt_lagr_factor=2)
    o, lat = pair(cfg, pc, 100)
    a = lat.particles()
    assert np.unique(a["aspect_ratio"]).size == 4 and np.abs(a["dt_px"]).max() > 0
    compare_particles(o, lat)


@pytest.mark.parametrize("mode", [abi.LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC, abi.LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE,
                                  abi.LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION,
                                  abi.LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION],
                         ids=["ballistic", "temperature", "second", "selective"])
def test_swimming_modes_fp64(mode):
    """The swimming modes of LAGRANGE_ORIENTATION_ACTIVE in Rayleigh-Benard
    with rods and disks (aspect ratios 0.5 .. 8), bit-identical to the
    oracle over 100 steps."""
    flags = ORIENT | abi.LAGRANGE_GRAVITY | abi.LAGRANGE_ADDEDMASS
    pc = abi.make_particle_config(flags, 300, tau_drag=(2, 0.0, 4.0), beta_coeff=(1, 0.8, 0.8),
                                  aspect_ratio=(3, 0.5, 8.0), gyrotaxis_velocity=(2, 0.5, 2.0),
                                  swim_velocity=(2, 0.0, 0.002), time_dt_lagr_factor=2, flags2=mode)
    o, lat = pair(rb_small(), pc, 100)
    compare_particles(o, lat)


@pytest.mark.parametrize("mode", ["shear", "highpass", "temperature"])
def test_copepod_jumps_fp64(mode):
    """LAGRANGE_ORIENTATION_ACTIVE_JUMP in Rayleigh-Benard: swimming tracers
    jump when the shear rate (or, with JUMP_TEMPERATURE, the temperature)
    crosses their critical value, then decay as exp(-t / jump_time); the
    decay uses the device's exp, so the motion is held within 1e-9 of the
    oracle after 100 steps, and the set of tracers that jumped is the
    oracle's."""
    f2 = abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP
    csr = (2, 2e-7, 6e-7)
    if mode == "highpass":
        f2 |= abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS
    elif mode == "temperature":
        f2 |= abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE
        csr = (2, 0.35, 0.65)
    flags = ORIENT | abi.LAGRANGE_GRAVITY | abi.LAGRANGE_ADDEDMASS
    pc = abi.make_particle_config(flags, 300, tau_drag=(2, 0.0, 4.0), beta_coeff=(1, 0.8, 0.8),
                                  aspect_ratio=(3, 0.5, 8.0), gyrotaxis_velocity=(2, 0.5, 2.0),
                                  swim_velocity=(2, 0.0, 0.002), critical_shear_rate=csr, jump_time=(2, 2.0, 5.0),
                                  time_dt_lagr_factor=2, flags2=f2)
    o = Oracle(rb_small())
    o.initial_conditions()
    o.particles_init(pc)
    t0 = o.particles()["time_from_jump"].copy()
    o, lat = pair(rb_small(), pc, 100)
    a, b = o.particles(), lat.particles()
    assert len(a) == len(b) and np.array_equal(a["name"], b["name"])
    a_s, b_s = np.sort(a, order="name"), np.sort(b, order="name")
    tracer = a_s["tau_drag"] == 0.0
    jumped_a = a_s["time_from_jump"] < t0[a_s["name"].astype(int)] + 100 * 0.5 - 1e-9
    jumped_b = b_s["time_from_jump"] < t0[b_s["name"].astype(int)] + 100 * 0.5 - 1e-9
    swim = tracer & (a_s["swim_velocity"] > 0)
    print(f"  {mode}: {int(jumped_a[swim].sum())} of {int(swim.sum())} swimmers jumped")
    assert np.array_equal(jumped_a[tracer], jumped_b[tracer])
    assert jumped_a[swim].sum() > 0
    for f in ("time_from_jump", "px_jump", "py_jump", "pz_jump", "shear_rate"):
        err = np.abs(a[f] - b[f]).max()
        print(f"  {f}: max |gpu - oracle| = {err:.3e}")
        assert err <= 1e-9, f
    compare_particles(o, lat, exact=False, pos_atol=1e-9, rel=1e-9)


def test_random_reorientation_fp64():
    """LAGRANGE_ORIENTATION_RANDOM with random re-injection and the wall turn
    of p_y: each particle's draws in order (two for p, two per re-injection)
    from the device's copy of the stream; p from the device's acos / sincos,
    so p, dp/dt and the motion within 1e-9 of the oracle after 60 steps, and
    the re-injection positions exact."""
    flags = (HEAVY | abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY | abi.LAGRANGE_ORIENTATION_BC |
             abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT)
    pc = abi.make_particle_config(flags, 256, tau_drag=(2, 2.0, 20.0), beta_coeff=(2, 0.2, 0.5),
                                  aspect_ratio=(2, 0.5, 4.0), flags2=abi.LAGRANGE_ORIENTATION_RANDOM)
    o, lat = pair(rb_small(32, 16, g=5e-3), pc, 60)
    a, b = o.particles(), lat.particles()
    assert len(a) == len(b) and np.array_equal(a["name"], b["name"])
    for f in ("px", "py", "pz", "dt_px", "dt_py", "dt_pz"):
        err = np.abs(a[f] - b[f]).max()
        print(f"  {f}: max |gpu - oracle| = {err:.3e}")
        assert err <= 1e-9, f
    compare_particles(o, lat, exact=False, pos_atol=1e-9, rel=1e-9)


@pytest.mark.parametrize("case", ["heavy", "orientation"])
def test_polymer_conformation_fp64(case):
    """LAGRANGE_POLYMER: the conformation tensor along each trajectory (its
    initial diagonal from two draws after the orientation draws) bit-identical
    to the oracle over 100 steps."""
    flags = HEAVY if case == "heavy" else ORIENT | abi.LAGRANGE_GRAVITY | abi.LAGRANGE_ADDEDMASS
    pc = abi.make_particle_config(flags, 300, tau_drag=(2, 0.0, 4.0), beta_coeff=(1, 0.8, 0.8),
                                  aspect_ratio=(2, 0.5, 4.0), gyrotaxis_velocity=(1, 1.0, 1.0),
                                  swim_velocity=(1, 0.001, 0.001), flags2=abi.LAGRANGE_POLYMER, tau_polymer=20.0)
    o, lat = pair(rb_small(), pc, 100)
    assert np.abs(lat.particles()["dt_cxy"]).max() > 0
    compare_particles(o, lat)


def test_orientation_needs_the_gradient():
    lat = Lattice(rb_small())
    lat.initial_conditions()
    pc = abi.make_particle_config(abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY, 4)
    with pytest.raises(Exception, match="LAGRANGE_GRADIENT"):
        lat.particles_init(pc)


TPART = HEAVY | abi.LAGRANGE_TEMPERATURE


@pytest.mark.parametrize("extra", [
    abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT,
    abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT | abi.LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM |
    abi.LAGRANGE_TEMPERATURE_INCREMENT_LOG | abi.LAGRANGE_INITIAL_TEMPERATURE_FLUID,
], ids=["reinject", "norandom_log_fluid_start"])
def test_particle_temperature_fp64(extra):
    """LAGRANGE_TEMPERATURE: the cp ladder, t_p relaxing to the fluid's t
    every sub-step (lagrange.c:4007-4020), the wall temperature of a
    re-injected particle, and the start at T_bot or at the fluid's t:
    every record bit-identical to the oracle (tracers' t_p NaN in both)."""
    pc = abi.make_particle_config(TPART | extra, 256, tau_drag=(2, 2.0, 20.0), beta_coeff=(2, 0.2, 0.5),
                                  cp_coeff=(3, 0.5, 4.0), time_dt_lagr_factor=2)
    o, lat = pair(rb_small(32, 16, g=5e-3), pc, 120)
    a = lat.particles()
    assert np.unique(a["cp_p"]).size == 4
    fin = np.isfinite(a["t_p"])
    assert fin.sum() > 0 and np.ptp(a["t_p"][fin]) > 0
    compare_particles(o, lat)


def test_particle_temperature_fp32_dev_close_to_fp64():
    """Float-storage lattice: t_p within 1e-5 of the wall temperature
    difference of the FP64 oracle's after 100 steps."""
    pc = abi.make_particle_config(TPART, 300, tau_drag=(2, 2.0, 8.0), beta_coeff=(1, 0.6, 0.6),
                                  cp_coeff=(2, 0.5, 2.0))
    o = Oracle(rb_small())
    o.initial_conditions()
    o.particles_init(pc)
    o.step(100)
    cfg = rb_small(precision=abi.FP32_DEV)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.particles_init(pc)
    lat.step(100)
    a, b = o.particles(), lat.particles()
    assert len(a) == len(b) and np.array_equal(np.isnan(a["t_p"]), np.isnan(b["t_p"]))
    fin = np.isfinite(a["t_p"])
    err = np.abs(a["t_p"][fin] - b["t_p"][fin]).max() / abs(cfg.T_bot - cfg.T_top)
    print(f"  t_p: max error / |T_bot - T_top| = {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("extra", [
    abi.LAGRANGE_INITIAL_VELOCITY_FLUID,
    abi.LAGRANGE_INITIAL_VELOCITY_STOKES,
    abi.LAGRANGE_INITIAL_VELOCITY_FLUID | abi.LAGRANGE_INITIAL_VELOCITY_STOKES | abi.LAGRANGE_TEMPERATURE |
    abi.LAGRANGE_INITIAL_TEMPERATURE_FLUID,
], ids=["fluid", "stokes", "all_starts"])
def test_initial_velocity_fp64(extra):
    """The starts of initial_conditions_particles (lagrange.c:968-991,
    :1004-1015): the records right after ch4g_particles_init and after 60
    steps bit-identical to the oracle."""
    pc = abi.make_particle_config(HEAVY | extra, 256, tau_drag=(2, 2.0, 20.0), beta_coeff=(2, 0.2, 0.5),
                                  cp_coeff=(2, 0.5, 2.0))
    cfg = rb_small(32, 16, g=5e-3)
    o = Oracle(cfg)
    o.initial_conditions()
    o.particles_init(pc)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.particles_init(pc)
    a = lat.particles()
    assert np.abs(a["vy"]).max() > 0
    compare_particles(o, lat)
    o.step(60)
    lat.step(60)
    compare_particles(o, lat)


def test_variable_gravity_fp64():
    """LAGRANGE_GRAVITY_VARIABLE with its log ladder: the type's factor on
    the weight, the added-mass buoyancy and the Stokes velocity of a
    re-injected particle, bit-identical to the oracle."""
    flags = (HEAVY | abi.LAGRANGE_GRAVITY_VARIABLE | abi.LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG |
             abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT)
    pc = abi.make_particle_config(flags, 256, tau_drag=(2, 2.0, 20.0), beta_coeff=(2, 0.2, 0.5),
                                  gravity_coeff=(3, 0.5, 2.0))
    o, lat = pair(rb_small(32, 16, g=5e-3), pc, 120)
    assert np.unique(lat.particles()["gravity_coeff"]).size == 4
    compare_particles(o, lat)


@pytest.mark.parametrize("flags,flags2", [
    (abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_ADD_TRACER_STATIC | abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT |
     abi.LAGRANGE_ADDEDMASS | abi.LAGRANGE_GRAVITY, abi.LAGRANGE_INITIAL_GRID),
    (abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_ADD_TRACER_STATIC | abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT |
     abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY, abi.LAGRANGE_INITIAL_PAIRS),
], ids=["static_grid", "static_pairs_orientation"])
def test_probes_and_initial_positions_fp64(flags, flags2):
    """Eulerian probes (LAGRANGE_ADD_TRACER_STATIC) and the grid / pair
    starts: records bit-identical to the oracle after ch4g_particles_init
    (positions, the draws of the pairs before the orientation draws) and
    after 60 steps (probes in place, v = u)."""
    pc = abi.make_particle_config(flags, 256, tau_drag=(2, 2.0, 8.0), beta_coeff=(1, 0.5, 0.5),
                                  aspect_ratio=(2, 0.5, 4.0), flags2=flags2)
    c = configs.rb2d(24, 16, 8, Ra=1e5, seed=2024)  # 3-D: the grid lies inside the box
    cfg = c.copy(gravity_y=1e-3, beta_t=c.beta_t * c.gravity_y / 1e-3)
    o = Oracle(cfg)
    o.initial_conditions()
    o.particles_init(pc)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.particles_init(pc)
    compare_particles(o, lat)
    a0 = lat.particles()
    o.step(60)
    lat.step(60)
    compare_particles(o, lat)
    a = lat.particles()
    probe = a0["tau_drag"] == -1.0
    assert probe.sum() > 0 and np.array_equal(a["x"][probe], a0["x"][probe])
    assert np.array_equal(a["vx"][probe], a["ux"][probe])


@pytest.mark.parametrize("flags2", [abi.LAGRANGE_ADDEDMASS_LIFT,
                                    abi.LAGRANGE_ADDEDMASS_LIFT | abi.LAGRANGE_ADDEDMASS_WAKEDRAG],
                         ids=["lift", "lift_wakedrag"])
def test_lift_and_wake_drag_fp64(flags2):
    """The lift of the added-mass model bit-identical to the oracle; with the
    wake drag, whose pow(Re_p, 0.687) is the device's own, positions within
    1e-9 and velocities within 1e-9 of the largest velocity component after
    100 steps."""
    pc = abi.make_particle_config(HEAVY, 300, tau_drag=(2, 2.0, 8.0), beta_coeff=(2, 0.4, 0.8), flags2=flags2)
    o, lat = pair(rb_small(), pc, 100)
    if flags2 & abi.LAGRANGE_ADDEDMASS_WAKEDRAG:
        compare_particles(o, lat, exact=False, pos_atol=1e-9, rel=1e-9)
    else:
        compare_particles(o, lat)


def test_spheroid_drag_fp64():
    """LAGRANGE_ORIENTATION_DRAG with prolate and oblate types (aspect
    ratios 0.5 .. 6; response times long enough for the explicit AB2 steps): the drag tensor along p, with added mass and gravity,
    bit-identical to the oracle over 100 steps (c_perp, c_par by glibc on
    both sides)."""
    flags = (HEAVY | abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY |
             abi.LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG)
    pc = abi.make_particle_config(flags, 300, tau_drag=(2, 8.0, 20.0), beta_coeff=(1, 0.5, 0.5),
                                  aspect_ratio=(4, 0.5, 6.0), flags2=abi.LAGRANGE_ORIENTATION_DRAG,
                                  time_dt_lagr_factor=2)
    o, lat = pair(rb_small(), pc, 100)
    assert np.unique(lat.particles()["aspect_ratio"]).size == 5
    compare_particles(o, lat)


@pytest.mark.parametrize("extra2", [0, abi.LAGRANGE_BC_RADIUSANDDENSITY |
                                    abi.LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS],
                         ids=["ladders", "wall_radius_log"])
def test_radius_and_density_fp64(extra2):
    """LAGRANGE_RADIUSandDENSITY: tau_drag and beta_coeff from the radius and
    density ladders, and with LAGRANGE_BC_RADIUSandDENSITY elastic walls at
    the particle's radius; bit-identical to the oracle over 120 steps."""
    pc = abi.make_particle_config(HEAVY, 256, particle_radius=(2, 0.4, 1.0), particle_density=(2, 1.5, 3.0),
                                  flags2=abi.LAGRANGE_RADIUSANDDENSITY | extra2)
    o, lat = pair(rb_small(32, 16, g=1e-3), pc, 120)
    assert np.unique(lat.particles()["tau_drag"]).size == 5
    compare_particles(o, lat)


def test_substeps_fp64():
    pc = abi.make_particle_config(HEAVY, 128, tau_drag=(2, 1.0, 3.0), beta_coeff=(1, 0.7, 0.7),
                                  time_dt_lagr_factor=3)
    o, lat = pair(configs.poiseuille(16, 32, 8, tau=0.8, g=1e-5), pc, 60)
    compare_particles(o, lat)


def test_particles_fp32_dev_close_to_fp64():
    pc = abi.make_particle_config(HEAVY, 400, tau_drag=(3, 2.0, 8.0), beta_coeff=(2, 0.6, 1.2))
    cfg64 = rb_small()
    o = Oracle(cfg64)
    o.initial_conditions()
    o.particles_init(pc)
    o.step(100)
    lat = Lattice(rb_small(precision=abi.FP32_DEV))
    lat.initial_conditions()
    lat.particles_init(pc)
    lat.step(100)
    compare_particles(o, lat, exact=False, pos_atol=1e-6, rel=1e-5)


def test_particle_restart_roundtrip():
    pc = abi.make_particle_config(HEAVY, 64, tau_drag=(2, 1.0, 3.0), beta_coeff=(1, 0.7, 0.7))
    cfg = rb_small(32, 16)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.particles_init(pc)
    lat.step(5)
    a = lat.particles()
    lat.set_particles(a)
    assert np.array_equal(lat.particles(), a)
    # a larger set than the initial capacity
    b = np.concatenate([a, a])
    b["name"] = np.arange(len(b))
    lat.set_particles(b)
    assert lat.particles_count() == len(b)
    lat.step(3)
    assert np.isfinite(lat.particles()["x"]).all()


@pytest.mark.parametrize("case", ["hit", "rb"])
def test_particles_on_fused_fields_match_field_pass(case):
    """Float-storage lattices with one-way particles (no device random draws):
    within one ch4g_step call the step kernel writes the float u (and T) of
    the step it gathers and the particles of that step move after it; one
    step per call moves them on the separate field pass instead.  The two
    field sources differ only by float rounding (the kernel's float moments
    against the pass's double moments rounded to float), which the added-mass
    term (u - u_old)/dt amplifies: the bar is the FP32-deviation one against
    FP64 (positions within 1e-6 of the lattice spacing, velocities within
    1e-5 of the largest particle velocity component) after 60 steps, with
    particle order and count identical and fluid populations bit-identical."""
    if case == "hit":
        cfg = configs.hit(N=24, tau=0.505, amp=1e-4, les_cs=0.1, precision=abi.FP32_DEV)
        pc = abi.make_particle_config(abi.LAGRANGE_WRAP | abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_GRADIENT |
                                      abi.LAGRANGE_ADDEDMASS, 300, tau_drag=(2, 1.0, 4.0), beta_coeff=(1, 0.8, 0.8))
    else:
        cfg = rb_small(precision=abi.FP32_DEV)
        pc = abi.make_particle_config(HEAVY, 300, tau_drag=(2, 2.0, 6.0), beta_coeff=(1, 0.6, 0.6))
    fused, single = Lattice(cfg), Lattice(cfg)
    for lat in (fused, single):
        lat.initial_conditions()
        lat.particles_init(pc)
    fused.step(60)
    for _ in range(60):
        single.step(1)
    a, b = single.particles(), fused.particles()
    assert len(a) == len(b) and np.array_equal(a["name"], b["name"])
    for f in ("x", "y", "z"):
        d = np.abs(a[f] - b[f]).max()
        print(f"  {f}: max |fused - field pass| {d:.3e}")
        assert d <= 1e-6, f
    s = max(np.abs(a[f]).max() for f in ("vx", "vy", "vz"))
    for f in ("vx", "vy", "vz"):
        d = np.abs(a[f] - b[f]).max() / s
        print(f"  {f}: max |fused - field pass| / max|v| {d:.3e}")
        assert d <= 1e-5, f
    # the fields themselves: the fused path leaves the lattice where the
    # single steps leave it
    pa, ga = single.get_populations()
    pb, gb = fused.get_populations()
    assert np.array_equal(pa, pb) and np.array_equal(ga, gb)
