"""GPU parity: the HIP library (through the C-ABI) against the FP64 oracle.

FP64 device path: the bar is max|gpu - oracle| / max|oracle| <= 1e-12 per
field; the kernels follow the reference's evaluation order without
contraction, so the fields are expected (and checked) to be bit-identical.
FP32-deviation path: relative L2 error per field against the FP64 oracle,
velocity <= 1e-5.
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


def both(cfg, steps):
    o = Oracle(cfg)
    o.initial_conditions()
    o.step(steps)
    lat = Lattice(cfg)
    lat.initial_conditions()
    lat.step(steps)
    return o, lat


def compare_fp64(o, lat, exact=True):
    po, go = o.get_populations()
    pg, gg = lat.get_populations()
    ho, hg = o.hydro_fields(), lat.hydro_fields()
    fields = {"p": (pg, po), "dens": (hg["dens"], ho["dens"]), "u": (hg["u"], ho["u"])}
    if lat.cfg.has(abi.LB_TEMPERATURE):
        fields["g"] = (gg, go)
        fields["t"] = (hg["t"], ho["t"])
    for name, (a, b) in fields.items():
        err = rel_max(a, b)
        print(f"  {name}: max rel err {err:.3e}  bit-identical={np.array_equal(a, b)}")
        assert err <= 1e-12, (name, err)
        if exact:
            assert np.array_equal(a, b), name


@pytest.mark.parametrize("steps", [0, 1, 2, 10, 1000])
def test_poiseuille_fp64_matches_oracle(steps):
    o, lat = both(configs.poiseuille(4, 64, 4), steps)
    assert lat.itime == steps
    compare_fp64(o, lat)


def test_rb2d_fp64_matches_oracle_small():
    o, lat = both(configs.rb2d(NX=128, NY=64), 1000)
    compare_fp64(o, lat)


def test_rb2d_fp64_matches_oracle_full_size():
    """512x256 (the 2-D Rayleigh-Benard configuration), 200 steps."""
    o, lat = both(configs.rb2d(), 200)
    compare_fp64(o, lat)


def test_hit_les_fp64_matches_oracle():
    """Periodic HIT forcing (phase walk on the drand48 stream, host sin tables)
    with local-strain Smagorinsky; LES needs sqrt on the device."""
    cfg = configs.hit(N=16, tau=0.51, amp=1e-4)
    o, lat = both(cfg, 100)
    np.testing.assert_array_equal(lat.hit_phases(), o.hit_phases())
    compare_fp64(o, lat, exact=False)


@pytest.mark.parametrize("case", ["rb", "hit_poiseuille"])
def test_nozeromode_fp64_matches_oracle(case):
    """LB_FLUID_FORCING_NOZEROMODE on the FP64 path: the mean of the built
    force is removed at every node.  The device sums it in a fixed tree, the
    reference (and the oracle) node by node, so agreement is to rounding
    (1e-12 relative), not bit-identity; the run without the switch differs."""
    if case == "rb":
        cfg = configs.rb2d(NX=64, NY=32, Ra=1e6)
        steps = 300
    else:
        cfg = configs.hit(N=12, tau=0.55, amp=1e-4, les=False)
        cfg.flags |= abi.FLUID_FORCING_POISEUILLE
        cfg.Amp_x = 2e-3
        steps = 60
    cfg.flags |= abi.FLUID_FORCING_NOZEROMODE
    o, lat = both(cfg, steps)
    po, go = o.get_populations()
    pg, gg = lat.get_populations()
    ho, hg = o.hydro_fields(), lat.hydro_fields()
    assert rel_max(pg, po) <= 1e-12 and rel_max(hg["dens"], ho["dens"]) <= 1e-12
    if cfg.has(abi.LB_TEMPERATURE):
        assert rel_max(gg, go) <= 1e-12 and rel_max(hg["t"], ho["t"]) <= 1e-12
    # u is a difference of populations (plus F/2): before convection sets in
    # |u| is far below |f|, so its rounding scale is the populations'
    scale = max(np.abs(ho["u"]).max(), np.abs(po).max())
    err_u = np.abs(hg["u"] - ho["u"]).max() / scale
    print(f"  p {rel_max(pg, po):.2e}  u {err_u:.2e} (of max(|u|, |f|)), "
          f"u {rel_max(hg['u'], ho['u']):.2e} of max|u| {np.abs(ho['u']).max():.2e}")
    assert err_u <= 1e-12
    u = hg["u"]
    plain = cfg.copy(flags=cfg.flags & ~abi.FLUID_FORCING_NOZEROMODE)
    _, lp = both(plain, steps)
    assert rel_max(lp.hydro_fields()["u"], u) > 1e-6
    if case == "hit_poiseuille":
        # the uniform force is gone: the mean streamwise velocity stays ~0
        assert abs(u[..., 0].mean()) < 1e-3 * abs(lp.hydro_fields()["u"][..., 0].mean())


@pytest.mark.parametrize("precision", [abi.FP32_DEV, abi.FP32_STORE], ids=["fp32dev", "fp32store"])
@pytest.mark.parametrize("case", ["rb", "hit_poiseuille"])
def test_nozeromode_fp32_against_fp64_oracle(case, precision):
    """LB_FLUID_FORCING_NOZEROMODE on float storage: the force-mean pre-pass
    sums the float force the FP32 kernels build, in double, and the kernels
    remove its mean at every node.  Relative L2 error per field against the
    FP64 oracle with the switch <= 1e-5; with a uniform force in the mix the
    mean streamwise velocity stays below 1e-3 of its value without the
    switch."""
    if case == "rb":
        cfg = configs.rb2d(NX=64, NY=32, Ra=1e6)
        steps = 1000
    else:
        cfg = configs.hit(N=12, tau=0.55, amp=1e-4, les=False)
        cfg.flags |= abi.FLUID_FORCING_POISEUILLE
        cfg.Amp_x = 2e-3
        steps = 60
    cfg.flags |= abi.FLUID_FORCING_NOZEROMODE
    o = Oracle(cfg)
    o.initial_conditions()
    o.step(steps)
    c32 = cfg.copy(precision=precision)
    lat = Lattice(c32)
    lat.initial_conditions()
    lat.step(steps)
    assert lat.steps_per_launch() == 1
    ho, hg = o.hydro_fields(), lat.hydro_fields()
    eu, ed = rel_l2(hg["u"], ho["u"]), rel_l2(hg["dens"], ho["dens"])
    et = rel_l2(hg["t"], ho["t"]) if cfg.has(abi.LB_TEMPERATURE) else 0.0
    print(f"  {case}: L2 rel err u {eu:.3e}  T {et:.3e}  rho {ed:.3e}")
    assert eu <= 1e-5 and et <= 1e-5 and ed <= 1e-5
    if case == "hit_poiseuille":
        plain = c32.copy(flags=c32.flags & ~abi.FLUID_FORCING_NOZEROMODE)
        lp = Lattice(plain)
        lp.initial_conditions()
        lp.step(steps)
        assert abs(hg["u"][..., 0].mean()) < 1e-3 * abs(lp.hydro_fields()["u"][..., 0].mean())


def test_hit_les_fd_fp64_matches_oracle():
    """HIT with the finite-difference Smagorinsky closure (the reference's
    default LB_FLUID_LES): a velocity kernel stores u, the step kernel
    differentiates it at the six neighbours."""
    cfg = configs.hit(N=16, tau=0.505, amp=1e-4, les="fd")
    o, lat = both(cfg, 100)
    compare_fp64(o, lat, exact=False)


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_DEV])
def test_rb2d_les_fd_walls_temperature_les(precision):
    """Finite-difference LES with one-sided differences at the y walls and
    the temperature relaxation time following it (LB_TEMPERATURE_LES)."""
    def with_les(c):
        c.flags |= abi.FLUID_LES | abi.TEMPERATURE_LES
        c.les_cs = 0.18
        return c

    o = Oracle(with_les(configs.rb2d(NX=64, NY=32, Ra=1e7)))
    lat = Lattice(with_les(configs.rb2d(NX=64, NY=32, Ra=1e7, precision=precision)))
    for x in (o, lat):
        x.initial_conditions()
        x.step(300)
    if precision == abi.FP64:
        compare_fp64(o, lat, exact=False)
    else:
        ho, hg = o.hydro_fields(), lat.hydro_fields()
        eu, et = rel_l2(hg["u"], ho["u"]), rel_l2(hg["t"], ho["t"])
        print(f"  L2 rel err u {eu:.3e}  T {et:.3e}")
        assert eu <= 1e-5 and et <= 1e-5


def test_dump_averages_fp64_matches_oracle():
    o, lat = both(configs.rb2d(NX=96, NY=48, Ra=1e6), 300)
    ao, ag = o.dump_averages(), lat.dump_averages()
    for k, v in ao.items():
        if k in ("x", "y", "z"):
            continue
        scale = max(abs(v), 1e-300)
        print(f"  {k}: oracle {v:.15e} gpu {ag[k]:.15e}")
        assert abs(ag[k] - v) <= 1e-11 * scale + 1e-16, k  # floor: cancelling sums (periodic means ~ 0)


@pytest.mark.parametrize("cfg,steps", [(configs.rb2d(NX=96, NY=48, Ra=1e6), 300),
                                       (configs.poiseuille(12, 16, 6, tau=0.7, g=1e-5), 50)])
def test_dump_rulers_fp64_matches_oracle(cfg, steps):
    """x, y, z rulers (plane averages of dump_averages) against the oracle's
    sequential plane sums; bar 1e-11 of each column's scale."""
    o, lat = both(cfg, steps)
    _, ox, oy, oz = o.dump_rulers()
    gx, gy, gz = lat.dump_rulers()
    for axis, a, b in (("x", gx, ox), ("y", gy, oy), ("z", gz, oz)):
        for m in b.dtype.names:
            scale = np.abs(b[m]).max()
            err = np.abs(a[m] - b[m]).max()
            assert err <= 1e-11 * scale + 1e-16, (axis, m, err, scale)


def test_set_get_populations_roundtrip():
    rng = np.random.default_rng(5)
    cfg = configs.rb2d(NX=16, NY=8, NZ=3)
    lat = Lattice(cfg)
    W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    p = W * (1 + 0.01 * rng.standard_normal((3, 8, 16, 19)))
    g = W * (0.5 + 0.01 * rng.standard_normal((3, 8, 16, 19)))
    lat.set_populations(p.reshape(-1, 19), g.reshape(-1, 19))
    p2, g2 = lat.get_populations()
    assert np.array_equal(p2, p) and np.array_equal(g2, g)


@pytest.mark.parametrize("name,cfg,steps", [
    ("poiseuille 4x64x4", configs.poiseuille(4, 64, 4, precision=abi.FP32_DEV), 1000),
    # same tau_u = 0.50384 as the 512x256 Ra=1e8 case
    ("rb2d 128x64 Ra=6.25e6", configs.rb2d(NX=128, NY=64, Ra=6.25e6, precision=abi.FP32_DEV), 1000),
    ("rb2d 512x256 Ra=1e8", configs.rb2d(precision=abi.FP32_DEV), 1000),
    ("rb2d 128x64 Ra=6.25e6 store", configs.rb2d(NX=128, NY=64, Ra=6.25e6, precision=abi.FP32_STORE), 1000),
    ("rb2d 512x256 Ra=1e8 store", configs.rb2d(precision=abi.FP32_STORE), 1000),
    ("hit 32^3 fd-les", configs.hit(N=32, tau=0.505, amp=1e-4, les="fd", precision=abi.FP32_DEV), 200),
])
def test_fp32_deviation_against_fp64_oracle(name, cfg, steps):
    """Relative L2 error per field of the FP32-deviation path against the
    FP64 oracle on the same inputs; bar: velocity <= 1e-5."""
    o, lat = both(cfg, steps)
    ho, hg = o.hydro_fields(), lat.hydro_fields()
    eu = rel_l2(hg["u"], ho["u"])
    ed = rel_l2(hg["dens"], ho["dens"])
    print(f"  {name}: L2 rel err u {eu:.3e}  rho {ed:.3e}")
    if cfg.has(abi.LB_TEMPERATURE):
        et = rel_l2(hg["t"], ho["t"])
        print(f"  {name}: L2 rel err T {et:.3e}")
        assert et <= 1e-5
    assert eu <= 1e-5


def test_rb2d_full_size_nusselt_fp32_within_one_percent():
    """512x256 Rayleigh-Benard at Ra=1e8 over 100000 steps: the Nusselt number
    of the FP32-deviation path, averaged over steps 30000..100000 (dumps every
    500 steps), is within 1 % of the FP64 path's (the FP64 device path is
    bit-identical to the oracle; the two trajectories separate once the flow
    is unsteady, so instantaneous values are not compared)."""
    nus = {}
    for prec in (abi.FP64, abi.FP32_DEV):
        lat = Lattice(configs.rb2d(precision=prec))
        lat.initial_conditions()
        series = []
        for _ in range(200):
            lat.step(500)
            series.append(lat.dump_averages()["nuy"])
        nus[prec] = np.array(series)
        lat.close()
    a, b = nus[abi.FP64][60:].mean(), nus[abi.FP32_DEV][60:].mean()
    print(f"  <Nu> 30000..100000: FP64 {a:.5f}  FP32-deviation {b:.5f}  rel {abs(b / a - 1):.2e}")
    assert abs(b / a - 1) < 0.01


def test_poiseuille_fp32_closed_form():
    """FP32-deviation storage reaches the exact lattice Poiseuille solution
    (parabola + bounce-back slip) to FP32 accuracy."""
    H, g, tau = 16, 1e-6, 0.8
    nu = (tau - 0.5) / 3
    lat = Lattice(configs.poiseuille(4, H, 1, tau=tau, g=g, precision=abi.FP32_DEV))
    lat.initial_conditions()
    lat.step(int(40 / (np.pi ** 2 * nu / H ** 2)))
    u = lat.hydro_fields()["u"][0, :, 0, 0]
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y) + (g / nu) * (16 * tau ** 2 - 16 * tau + 1) / 24
    assert np.abs(u - ua).max() / ua.max() < 2e-5


def test_rb2d_full_size_invariants_fp32():
    """512x256 Rayleigh-Benard in FP32-deviation storage: the initial
    averages equal the oracle's (the initial Nusselt number is not 1: the
    hydro velocity carries the Guo half-force, so <u_y T> adds
    (beta g / 2) Var(T) / (kappa dT / H) ~ 0.08), mass is conserved over
    2000 steps and the fields stay finite and bounded."""
    cfg = configs.rb2d(precision=abi.FP32_DEV)
    o = Oracle(configs.rb2d())
    o.initial_conditions()
    ao = o.dump_averages()
    lat = Lattice(cfg)
    lat.initial_conditions()
    a0 = lat.dump_averages()
    assert abs(a0["nuy"] - ao["nuy"]) < 1e-6 * abs(ao["nuy"])
    assert abs(a0["t"] - ao["t"]) < 1e-7
    lat.step(2000)
    a1 = lat.dump_averages()
    assert abs(a1["rho"] - a0["rho"]) < 1e-6
    h = lat.hydro_fields()
    assert np.isfinite(h["u"]).all() and np.isfinite(h["t"]).all()
    assert h["t"].min() > -0.05 and h["t"].max() < 1.05
