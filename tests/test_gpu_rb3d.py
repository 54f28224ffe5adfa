"""GPU parity of the 3-D Rayleigh-Benard configuration (fluid + temperature,
buoyancy, no-slip and fixed-T walls at y, periodic x and z) and of the 3-D
channel forcing, through the C-ABI against the FP64 oracle.

* FP64: populations, density, velocity and temperature bit-identical to the
  oracle (kernels_fp64.hip follows the reference's evaluation order without
  contraction).  NX = 45 fills 70 % of the 64-lane waves of a row, so the
  steady steps run on the dense one-node-per-lane grid; NX = 64 runs the row
  grid.  Both exercise the 3-D thermal z-neighbour gathers of every step.
* FP32-deviation (and FP32 storage with double arithmetic): relative L2 error
  per field against the FP64 oracle <= 1e-5.
* Two-way coupled heavy particles on a 32^3 lattice: particle records and
  fluid populations bit-identical; with float storage (the same per-node
  sums rounded to float planes) two runs bit-identical and within the FP32
  tolerance of the oracle.
* 512^3 (the full 3-D configuration, Ra = 1e9): initial averages against
  their closed forms, FP32-deviation against the FP64 device path after 100
  steps, mass conservation and bounded fields after 400 steps.
"""
import numpy as np
import pytest

from ch4g import abi, Lattice
from oracle import Oracle
import configs

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


def rel_l2(a, b):
    s = np.sqrt((b * b).sum())
    return np.sqrt(((a - b) ** 2).sum()) / (s if s > 0 else 1.0)


def run(cfg, steps, oracle=False):
    x = Oracle(cfg) if oracle else Lattice(cfg)
    x.initial_conditions()
    x.step(steps)
    return x


def rb3d(NX, NY, NZ, precision=abi.FP64, Ra=2e6):
    return configs.rb2d(NX=NX, NY=NY, NZ=NZ, Ra=Ra, precision=precision, seed=777)


@pytest.mark.parametrize("shape,steps", [((45, 32, 24), 300), ((64, 24, 16), 200)],
                         ids=["dense_45x32x24", "rows_64x24x16"])
def test_rb3d_fp64_bit_identical(shape, steps):
    cfg = rb3d(*shape)
    print(f"  rb3d FP64 {shape} {steps} steps", flush=True)
    o = run(cfg, steps, oracle=True)
    lat = run(cfg, steps)
    po, go = o.get_populations()
    pg, gg = lat.get_populations()
    ho, hg = o.hydro_fields(), lat.hydro_fields()
    for name, a, b in (("p", pg, po), ("g", gg, go), ("dens", hg["dens"], ho["dens"]), ("u", hg["u"], ho["u"]),
                       ("t", hg["t"], ho["t"])):
        print(f"  {name}: max rel err {rel_max(a, b):.3e} bit-identical={np.array_equal(a, b)}")
        assert np.array_equal(a, b), name
    # the flow has left the initial state: the comparison is not of zeros
    assert np.abs(ho["u"]).max() > 1e-6
    ao, ag = o.dump_averages(), lat.dump_averages()
    for k in ("nuy", "t", "t2", "ene", "eps", "epst"):
        assert abs(ag[k] - ao[k]) <= 1e-11 * max(abs(ao[k]), 1e-300), (k, ag[k], ao[k])


@pytest.mark.parametrize("precision", [abi.FP32_DEV, abi.FP32_STORE], ids=["fp32dev", "fp32store"])
def test_rb3d_fp32_against_fp64_oracle(precision):
    shape, steps = (45, 32, 24), 300
    o = run(rb3d(*shape), steps, oracle=True)
    lat = run(rb3d(*shape, precision=precision), steps)
    ho, hg = o.hydro_fields(), lat.hydro_fields()
    eu, et, ed = rel_l2(hg["u"], ho["u"]), rel_l2(hg["t"], ho["t"]), rel_l2(hg["dens"], ho["dens"])
    print(f"  L2 rel err u {eu:.3e}  T {et:.3e}  rho {ed:.3e}")
    assert eu <= 1e-5 and et <= 1e-5 and ed <= 1e-5


def test_channel_forcing_3d_fp64_bit_identical():
    """LB_FLUID_FORCING_CHANNEL (forcing.c:374-379, F_x = 2 Amp_x^2 / LY) on a
    3-D channel with no-slip y walls, together with the Poiseuille term."""
    cfg = abi.make_config(20, 18, 6, abi.LB_FLUID | abi.FLUID_FORCING_CHANNEL | abi.FLUID_BC_Y, tau_u=0.8,
                          Amp_x=3e-3)
    o, lat = run(cfg, 400, oracle=True), run(cfg, 400)
    po, _ = o.get_populations()
    pg, _ = lat.get_populations()
    assert np.array_equal(pg, po)
    u = lat.hydro_fields()["u"]
    assert np.array_equal(u, o.hydro_fields()["u"])
    assert u[..., 0].max() > 1e-4  # the channel force drives the flow
    both = cfg.copy(flags=cfg.flags | abi.FLUID_FORCING_POISEUILLE)
    o, lat = run(both, 50, oracle=True), run(both, 50)
    assert np.array_equal(lat.get_populations()[0], o.get_populations()[0])


HEAVY = abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS | \
    abi.LAGRANGE_GRAVITY


def test_rb3d_twoway_heavy_particles_fp64():
    """32^3 convection cell with 2000 heavy particles (beta = 0.6, density
    ratio 2) coupled back onto the fluid: every particle field and the fluid
    populations bit-identical to the oracle after 60 steps."""
    c = rb3d(32, 32, 32, Ra=1e5)
    g = 2e-3
    cfg = c.copy(gravity_y=g, beta_t=c.beta_t * c.gravity_y / g)
    pc = abi.make_particle_config(HEAVY | abi.LAGRANGE_TWOWAY_MOMENTUM, 2000, tau_drag=(2, 1.0, 4.0),
                                  beta_coeff=(1, 0.6, 0.6))
    o, lat = Oracle(cfg), Lattice(cfg)
    for x in (o, lat):
        x.initial_conditions()
        x.particles_init(pc)
        x.step(60)
    a, b = o.particles(), lat.particles()
    assert len(a) == len(b) > 1900
    for f in abi.PARTICLE_FIELDS:
        assert np.array_equal(a[f], b[f]), f
    po, go = o.get_populations()
    pg, gg = lat.get_populations()
    assert np.array_equal(po, pg) and np.array_equal(go, gg)
    # the particles act on the fluid: without the feedback the fluid differs
    pc1 = abi.make_particle_config(HEAVY, 2000, tau_drag=(2, 1.0, 4.0), beta_coeff=(1, 0.6, 0.6))
    one = Lattice(cfg)
    one.initial_conditions()
    one.particles_init(pc1)
    one.step(60)
    assert not np.array_equal(one.get_populations()[0], pg)


def test_rb3d_twoway_clustered_particles_fp64():
    """The feedback's per-node merge where it has the most to order: 3000
    heavy particles, 2400 packed in a 2.5-node cube (dozens per stencil cell,
    hundreds of deposits per node, from up to eight cells each), 300 across
    the periodic x and z faces and 300 within half a node of the y walls
    (stencils reaching the ghost rows, whose deposits never reach the force).
    Every particle field and the fluid populations bit-identical to the
    oracle after 12 steps; the float-storage run within the FP32 bar of it."""
    c = rb3d(24, 20, 16, Ra=1e5)
    g = 2e-3
    cfg = c.copy(gravity_y=g, beta_t=c.beta_t * c.gravity_y / g)
    pc = abi.make_particle_config(HEAVY | abi.LAGRANGE_TWOWAY_MOMENTUM, 3000, tau_drag=(2, 1.0, 4.0),
                                  beta_coeff=(1, 0.6, 0.6))
    o = Oracle(cfg)
    o.initial_conditions()
    o.particles_init(pc)
    a = o.particles()
    rng = np.random.default_rng(5)
    n1, n2 = 2400, 300
    a["x"][:n1] = 11.3 + 2.5 * rng.random(n1)
    a["y"][:n1] = 9.7 + 2.5 * rng.random(n1)
    a["z"][:n1] = 7.2 + 2.5 * rng.random(n1)
    a["x"][n1:n1 + n2] = np.where(rng.random(n2) < 0.5, 0.4 * rng.random(n2), 24 - 0.4 * rng.random(n2))
    a["z"][n1:n1 + n2] = np.where(rng.random(n2) < 0.5, 0.4 * rng.random(n2), 16 - 0.4 * rng.random(n2))
    a["y"][n1 + n2:] = np.where(rng.random(3000 - n1 - n2) < 0.5, 0.05 + 0.4 * rng.random(3000 - n1 - n2),
                                20 - 0.05 - 0.4 * rng.random(3000 - n1 - n2))
    o.set_particles(a)
    o.step(12)
    po, fo = o.particles(), o.get_populations()
    ho = o.hydro_fields()
    for prec in (abi.FP64, abi.FP32_DEV):
        lat = Lattice(cfg.copy(precision=prec))
        lat.initial_conditions()
        lat.particles_init(pc)
        lat.set_particles(a)
        lat.step(12)
        pl = lat.particles()
        assert len(pl["x"]) == len(po["x"])
        if prec == abi.FP64:
            for f in abi.PARTICLE_FIELDS:
                assert np.array_equal(pl[f], po[f]), f
            fl = lat.get_populations()
            assert np.array_equal(fl[0], fo[0]) and np.array_equal(fl[1], fo[1])
        else:
            eu = rel_l2(lat.hydro_fields()["u"], ho["u"])
            ev = max(np.abs(pl[f] - po[f]).max() for f in ("vx", "vy", "vz")) / max(
                np.abs(po[f]).max() for f in ("vx", "vy", "vz"))
            print(f"  clustered, float storage: u L2 {eu:.3e}  v {ev:.3e}")
            assert eu <= 1e-5 and ev <= 1e-5
        lat.close()
    # the cluster's feedback moved the fluid: without it the oracle differs
    o2 = Oracle(cfg)
    o2.initial_conditions()
    o2.particles_init(abi.make_particle_config(HEAVY, 3000, tau_drag=(2, 1.0, 4.0), beta_coeff=(1, 0.6, 0.6)))
    o2.set_particles(a)
    o2.step(12)
    assert rel_l2(o2.hydro_fields()["u"], ho["u"]) > 1e-3


def test_rb3d_twoway_feedback_fp32():
    """Float storage takes the feedback as float planes of the same per-node
    sums in particle order (the FP64 path keeps them in double): the same
    32^3 cell with 2000 heavy particles, 60 steps.  Two FP32-deviation runs
    are bit-identical, and against the FP64 oracle the fluid velocity and
    temperature have relative L2 <= 1e-5, positions stay within 1e-6 and
    velocities within 1e-5 of the largest component (the particle velocities
    follow the float fluid's velocity, L2 ~ 4e-7, in max norm).  A run without
    the feedback is far from both: the float planes carry the coupling."""
    c = rb3d(32, 32, 32, Ra=1e5)
    g = 2e-3
    cfg = c.copy(gravity_y=g, beta_t=c.beta_t * c.gravity_y / g)
    pc = abi.make_particle_config(HEAVY | abi.LAGRANGE_TWOWAY_MOMENTUM, 2000, tau_drag=(2, 1.0, 4.0),
                                  beta_coeff=(1, 0.6, 0.6))
    o = Oracle(cfg)
    o.initial_conditions()
    o.particles_init(pc)
    o.step(60)
    po, ho = o.particles(), o.hydro_fields()
    runs = []
    for _ in range(2):
        lat = Lattice(cfg.copy(precision=abi.FP32_DEV))
        lat.initial_conditions()
        lat.particles_init(pc)
        lat.step(60)
        runs.append((lat.particles(), lat.get_populations(), lat.hydro_fields()))
        lat.close()
    (p1, f1, h1), (p2, f2, h2) = runs
    for f in abi.PARTICLE_FIELDS:
        assert np.array_equal(p1[f], p2[f]), f
    assert np.array_equal(f1[0], f2[0]) and np.array_equal(f1[1], f2[1])
    assert len(p1["x"]) == len(po["x"]) > 1900
    eu, et = rel_l2(h1["u"], ho["u"]), rel_l2(h1["t"], ho["t"])
    xmax = max(np.abs(po[f]).max() for f in ("x", "y", "z"))
    vmax = max(np.abs(po[f]).max() for f in ("vx", "vy", "vz"))
    # positions compared as periodic distances (a particle may wrap in one run only)
    ex = max(np.minimum(np.abs(p1[f] - po[f]), 32 - np.abs(p1[f] - po[f])).max() for f in ("x", "y", "z")) / xmax
    ev = max(np.abs(p1[f] - po[f]).max() for f in ("vx", "vy", "vz")) / vmax
    print(f"  FP32 feedback vs FP64 oracle: u {eu:.3e}  T {et:.3e}  x {ex:.3e}  v {ev:.3e}")
    assert eu <= 1e-5 and et <= 1e-5
    assert ex <= 1e-6 and ev <= 1e-5
    # without the feedback the fluid is far from the coupled FP64 run
    one = Lattice(cfg.copy(precision=abi.FP32_DEV))
    one.initial_conditions()
    one.particles_init(abi.make_particle_config(HEAVY, 2000, tau_drag=(2, 1.0, 4.0), beta_coeff=(1, 0.6, 0.6)))
    one.step(60)
    e1 = rel_l2(one.hydro_fields()["u"], ho["u"])
    one.close()
    print(f"  without feedback: u {e1:.3e}")
    assert e1 > 100 * eu


@pytest.mark.parametrize("kind", ["rb", "channel"])
def test_two_steps_per_launch_against_single_steps(kind):
    """A 3-D float-storage lattice of 2^21 nodes runs two steps per launch
    (k_step2z_fp32, x-y tiles marching along z through LDS rings); with
    ch4g_set_fusion(0) it runs one step per launch.  Populations and hydro
    fields bit-identical after 25 steps (an odd count: the last step single
    in both), with graph replay on; RB 128^3 fluid + T and a fluid-only
    channel 128^3."""
    if kind == "rb":
        cfg = rb3d(128, 128, 128, precision=abi.FP32_DEV, Ra=1e6)
    else:
        cfg = configs.channel3d(128, 128, 128, g=1e-5)
    res = []
    for fusion in (True, False):
        lat = Lattice(cfg)
        lat.set_fusion(fusion)
        assert lat.steps_per_launch() == (2 if fusion else 1)
        lat.initial_conditions()
        lat.step(25)
        res.append((lat.get_populations(), lat.hydro_fields()))
        lat.close()
    (f1, h1), (f2, h2) = res
    for a, b in zip(f1, f2):
        if a is not None:
            assert np.array_equal(a, b)
    for k in h1:
        if h1[k] is not None:
            assert np.array_equal(h1[k], h2[k]), k
    assert np.abs(h1["u"]).max() > 0


def test_rb3d_512_full_size():
    """The full 3-D configuration: 512^3, Ra = 1e9, Pr = 1, initial T linear
    plus uniform noise 1e-3.

    Initial averages against closed forms (y = (j + 1/2)/H, H = 512):
    <T> = 1/2, <T^2> = 1/3 - 1/(12 H^2) + a^2/3 (noise amplitude a), and
    Nu = 1 + (beta g / 2) Var(T) H / (kappa dT): the hydro velocity carries the
    Guo half force beta g (T - (T_bot + T_top)/2), so <u_y T> = (beta g / 2)
    Var(T) (a sign slip or another reference temperature changes it).
    Then FP32-deviation against the FP64 device path (bit-identical to the
    oracle on the smaller 3-D lattices) after 100 steps, and mass
    conservation and bounded fields after 400 steps."""
    H, a = 512, 1e-3
    c64 = configs.rb2d(NX=H, NY=H, NZ=H, Ra=1e9)
    kappa = (c64.tau_t - 0.5) / 3.0
    print("  512^3 FP64", flush=True)
    lat = Lattice(c64)
    lat.initial_conditions()
    a0 = lat.dump_averages()
    t2 = 1.0 / 3.0 - 1.0 / (12.0 * H * H) + a * a / 3.0
    var = a0["t2"] - a0["t"] ** 2
    nu0 = 1.0 + 0.5 * c64.beta_t * c64.gravity_y * var * H / kappa
    print(f"  <T> {a0['t']:.12f}  <T^2> {a0['t2']:.12f} (closed form {t2:.12f})  Nu {a0['nuy']:.8f} "
          f"(closed form {nu0:.8f})  <rho> {a0['rho']!r}")
    assert abs(a0["t"] - 0.5) < 1e-6
    assert abs(a0["t2"] - t2) < 4e-7
    assert abs(a0["nuy"] / nu0 - 1) < 1e-4
    assert abs(a0["rho"] - 1.0) < 1e-12
    lat.step(100)
    h64 = lat.hydro_fields()
    lat.close()
    print("  512^3 FP32-deviation", flush=True)
    lat = Lattice(configs.rb2d(NX=H, NY=H, NZ=H, Ra=1e9, precision=abi.FP32_DEV))
    lat.initial_conditions()
    b0 = lat.dump_averages()
    for k in ("t", "t2", "nuy", "rho"):
        assert abs(b0[k] - a0[k]) <= 1e-6 * abs(a0[k]), k
    lat.step(100)
    h32 = lat.hydro_fields()
    eu, et = rel_l2(h32["u"], h64["u"]), rel_l2(h32["t"], h64["t"])
    print(f"  100 steps: L2 rel err FP32-deviation vs FP64: u {eu:.3e}  T {et:.3e}  "
          f"max|u| {np.abs(h64['u']).max():.3e}")
    assert eu <= 1e-5 and et <= 1e-5
    del h64
    lat.step(300)
    b1 = lat.dump_averages()
    h = lat.hydro_fields()
    assert abs(b1["rho"] - 1.0) < 1e-6
    assert np.isfinite(h["u"]).all() and np.isfinite(h["t"]).all()
    assert h["t"].min() > -0.05 and h["t"].max() < 1.05
    lat.close()


def test_rb3d_512_twoway_particles_fp32_against_fp64():
    """The 3-D configuration with its 4e6 two-way coupled heavy particles at
    512^3 (the particle set of bench.py: beta 0.6 i.e. density ratio 2,
    tau_p 516, gradients, added mass, gravity, feedback momentum; the particle
    gravity 2e-5 with beta_t rescaled to keep the fluid's beta_t g): the
    FP32-deviation run (particle rows in cell order, float feedback planes) against the FP64 device path (bit-identical to the oracle on the
    32^3 two-way case above) after 20 steps: every particle alive and in the
    box, positions within 1e-6 (lattice units) and velocities within 1e-5 of
    the largest component; fluid velocity and temperature relative L2 <= 1e-5;
    mass conserved.  (Left to run, this configuration blows up locally after
    about 50 steps in either precision: tau_u = 0.5024 at Ra = 1e9 on 512
    rows, and the 4e6 point forces of particles of 2.7 lattice volumes each.)"""
    H, npart = 512, 4_000_000
    flags = (abi.LAGRANGE_WRAP | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS | abi.LAGRANGE_GRAVITY |
             abi.LAGRANGE_TWOWAY_MOMENTUM)
    pc = abi.make_particle_config(flags, npart, tau_drag=(1, 516.0, 516.0), beta_coeff=(1, 0.6, 0.6))
    res = {}
    for prec in (abi.FP64, abi.FP32_DEV):
        c = configs.rb2d(NX=H, NY=H, NZ=H, Ra=1e9, precision=prec)
        g = 2e-5
        cfg = c.copy(gravity_y=g, beta_t=c.beta_t * c.gravity_y / g)
        print(f"  512^3 two-way, precision {prec}", flush=True)
        lat = Lattice(cfg)
        lat.initial_conditions()
        lat.particles_init(pc)
        lat.step(20)
        assert lat.particles_count() == npart
        res[prec] = (lat.particles(), lat.hydro_fields(), lat.dump_averages())
        lat.close()
    (p64, h64, a64), (p32, h32, a32) = res[abi.FP64], res[abi.FP32_DEV]
    for f, S in (("x", H), ("y", H), ("z", H)):
        assert (p32[f] >= 0).all() and (p32[f] < S).all()
        d = np.abs(p32[f] - p64[f])
        d = np.minimum(d, S - d)
        assert d.max() <= 1e-6, (f, d.max())
    vmax = max(np.abs(p64[f]).max() for f in ("vx", "vy", "vz"))
    for f in ("vx", "vy", "vz"):
        err = np.abs(p32[f] - p64[f]).max() / vmax
        print(f"  {f}: max error / max|v| {err:.3e}")
        assert err <= 1e-5, f
    eu, et = rel_l2(h32["u"], h64["u"]), rel_l2(h32["t"], h64["t"])
    print(f"  fluid after 20 steps: L2 rel err u {eu:.3e}  T {et:.3e}  max|u| {np.abs(h64['u']).max():.3e}")
    assert eu <= 1e-5 and et <= 1e-5
    assert abs(a32["rho"] - 1.0) < 1e-6 and abs(a64["rho"] - 1.0) < 1e-12
