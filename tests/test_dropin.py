"""The drop-in layer: reference-named functions over the C-ABI, the
define.h / param.in reader, and the main.c-style host ch4g_run."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from ch4g import abi
from ch4g.lattice import PKG_ROOT

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "poiseuille_4x64x4")
FIXP = os.path.join(HERE, "golden", "rb_particles_32x16")
DROPIN = os.path.join(PKG_ROOT, "lib", "libch4g_dropin.so")
RUN = os.path.join(PKG_ROOT, "bin", "ch4g_run")

REFERENCE_NAMES = ["initial_conditions", "initialization_forcing", "compute_advection", "add_collision",
                   "build_forcing", "add_forcing", "boundary_and_pbc_conditions_for_streaming", "time_stepping",
                   "hydro_fields", "dump_averages", "measure_time", "allocate_particles",
                   "initial_conditions_particles", "add_particle_feedbacks", "boundary_conditions_hydro",
                   "interpolate_vector_at_particles", "interpolate_scalar_at_particles", "move_particles",
                   "output_particles", "dump_particle_averages", "write_pop_h5", "read_pop_h5",
                   "write_point_particle_h5", "read_point_particle_h5"]


class RunParams(C.Structure):
    _fields_ = [("cfg", abi.Config), ("time_dt", C.c_double), ("time_start", C.c_double),
                ("time_end", C.c_double), ("time_dump_field", C.c_double), ("time_dump_diagn", C.c_double),
                ("output_at_start", C.c_int), ("diagn_append", C.c_int), ("output_h5", C.c_int),
                ("resume", C.c_int), ("resume_u", C.c_int), ("resume_t", C.c_int), ("lagrange", C.c_int), ("pcfg", abi.ParticleConfig),
                ("time_dump_lagr", C.c_double), ("lagr_output_fluid", C.c_int)]


def test_dropin_exports_reference_names():
    lib = C.CDLL(DROPIN)
    for n in REFERENCE_NAMES + ["ch4g_read_run_params", "ch4g_dropin_init", "ch4g_dropin_finalize"]:
        assert hasattr(lib, n), n


def test_read_run_params_poiseuille():
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    rc = lib.ch4g_read_run_params(os.path.join(FIX, "define.h").encode(), os.path.join(FIX, "param.in").encode(),
                                  C.byref(rp))
    assert rc == 0
    c = rp.cfg
    assert (c.NX, c.NY, c.NZ) == (4, 64, 4)
    assert c.flags == abi.LB_FLUID | abi.FLUID_FORCING_POISEUILLE | abi.FLUID_BC_Y
    assert c.tau_u == 1.0 and c.Amp_x == 0.003072 and c.seed == 1
    assert rp.time_end == 1000.0 and rp.time_dump_diagn == 250.0 and rp.output_at_start == 1


def test_read_run_params_particles():
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(os.path.join(FIXP, "define.h").encode(),
                                    os.path.join(FIXP, "param.in").encode(), C.byref(rp)) == 0
    assert rp.lagrange == 1
    pc = rp.pcfg
    want = (abi.LAGRANGE_WRAP | abi.LAGRANGE_ADD_TRACER | abi.LAGRANGE_GRADIENT | abi.LAGRANGE_ADDEDMASS |
            abi.LAGRANGE_GRAVITY | abi.LAGRANGE_BC_INELASTIC | abi.LAGRANGE_BC_INELASTIC_REINJECT |
            abi.LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF | abi.LAGRANGE_TWOWAY_MOMENTUM)
    assert pc.flags == want
    assert (pc.particle_number, pc.tau_drag_types, pc.tau_drag_max, pc.beta_coeff_min) == (96, 2.0, 12.0, 0.3)
    assert pc.time_dt_lagr_factor == 1.0 and rp.time_dump_lagr == 40.0
    # without the LAGRANGE_TWOWAY umbrella the feedback switch has no effect
    text = open(os.path.join(FIXP, "define.h")).read().replace("#define LAGRANGE_TWOWAY\n", "")
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        d = os.path.join(td, "define.h")
        open(d, "w").write(text)
        assert lib.ch4g_read_run_params(d.encode(), os.path.join(FIXP, "param.in").encode(), C.byref(rp)) == 0
        assert not rp.pcfg.flags & abi.LAGRANGE_TWOWAY_MOMENTUM
        # orientation: switches nested as in the reference, ladders read from param.in
        open(d, "w").write(text + "#define LAGRANGE_ORIENTATION\n#define LAGRANGE_ORIENTATION_JEFFREY\n"
                           "#define LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS\n#define LAGRANGE_ORIENTATION_ACTIVE\n")
        assert lib.ch4g_read_run_params(d.encode(), os.path.join(FIXP, "param.in").encode(), C.byref(rp)) == -1
        pin = os.path.join(td, "param.in")
        open(pin, "w").write(open(os.path.join(FIXP, "param.in")).read() +
                             "aspect_ratio_types 3\naspect_ratio_min 0.5\naspect_ratio_max 4\n"
                             "gyrotaxis_velocity_types 1\ngyrotaxis_velocity_min 2\ngyrotaxis_velocity_max 2\n"
                             "swim_velocity_types 2\nswim_velocity_min 0\nswim_velocity_max 0.01\n")
        assert lib.ch4g_read_run_params(d.encode(), pin.encode(), C.byref(rp)) == 0
        o = abi.LAGRANGE_ORIENTATION | abi.LAGRANGE_ORIENTATION_JEFFREY | abi.LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS
        assert rp.pcfg.flags & (o | abi.LAGRANGE_ORIENTATION_ACTIVE) == o | abi.LAGRANGE_ORIENTATION_ACTIVE
        assert (rp.pcfg.aspect_ratio_types, rp.pcfg.aspect_ratio_max, rp.pcfg.gyrotaxis_velocity_min,
                rp.pcfg.swim_velocity_max) == (3.0, 4.0, 2.0, 0.01)
        # without LAGRANGE_ORIENTATION its variants have no effect
        open(d, "w").write(text + "#define LAGRANGE_ORIENTATION_JEFFREY\n")
        assert lib.ch4g_read_run_params(d.encode(), os.path.join(FIXP, "param.in").encode(), C.byref(rp)) == 0
        assert not rp.pcfg.flags & abi.LAGRANGE_ORIENTATION_JEFFREY
        # the deterministic swimming modes are read into the second word
        open(d, "w").write(text + "#define LAGRANGE_ORIENTATION\n#define LAGRANGE_ORIENTATION_JEFFREY\n"
                           "#define LAGRANGE_ORIENTATION_ACTIVE\n#define LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC\n")
        assert lib.ch4g_read_run_params(d.encode(), pin.encode(), C.byref(rp)) == 0
        assert rp.pcfg.flags2 == abi.LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC
        # the stochastic orientation variants are not implemented
        open(d, "w").write(text + "#define LAGRANGE_ORIENTATION\n#define LAGRANGE_ORIENTATION_DIFFUSION\n")
        assert lib.ch4g_read_run_params(d.encode(), os.path.join(FIXP, "param.in").encode(), C.byref(rp)) == -1


def test_read_run_params_particle_temperature(tmp_path):
    """LAGRANGE_TEMPERATURE reads the cp_coeff ladder (parameters.c:606-616);
    its log and fluid-start variants act only inside it; it needs
    LB_TEMPERATURE and the added-mass beta_coeff its rate reads."""
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    text = open(os.path.join(FIXP, "define.h")).read()
    d, pin = tmp_path / "define.h", tmp_path / "param.in"
    base = open(os.path.join(FIXP, "param.in")).read()
    d.write_text(text + "#define LAGRANGE_TEMPERATURE\n#define LAGRANGE_TEMPERATURE_INCREMENT_LOG\n"
                 "#define LAGRANGE_INITIAL_TEMPERATURE_FLUID\n")
    pin.write_text(base)
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == -1  # no cp_coeff keys
    pin.write_text(base + "cp_coeff_types 3\ncp_coeff_min 0.5\ncp_coeff_max 2\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    t = abi.LAGRANGE_TEMPERATURE | abi.LAGRANGE_TEMPERATURE_INCREMENT_LOG | abi.LAGRANGE_INITIAL_TEMPERATURE_FLUID
    assert rp.pcfg.flags & t == t
    assert (rp.pcfg.cp_coeff_types, rp.pcfg.cp_coeff_min, rp.pcfg.cp_coeff_max) == (3.0, 0.5, 2.0)
    d.write_text(text + "#define LAGRANGE_TEMPERATURE_INCREMENT_LOG\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    assert not rp.pcfg.flags & t
    d.write_text(text.replace("#define LAGRANGE_ADDEDMASS\n", "") + "#define LAGRANGE_TEMPERATURE\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == -1


def test_read_run_params_initial_velocity(tmp_path):
    """LAGRANGE_INITIAL_VELOCITY_FLUID / _STOKES are read as switches; the
    Stokes start reads beta_coeff, so it needs the added mass."""
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    text = open(os.path.join(FIXP, "define.h")).read()
    d, pin = tmp_path / "define.h", os.path.join(FIXP, "param.in")
    d.write_text(text + "#define LAGRANGE_INITIAL_VELOCITY_FLUID\n#define LAGRANGE_INITIAL_VELOCITY_STOKES\n")
    assert lib.ch4g_read_run_params(str(d).encode(), pin.encode(), C.byref(rp)) == 0
    v = abi.LAGRANGE_INITIAL_VELOCITY_FLUID | abi.LAGRANGE_INITIAL_VELOCITY_STOKES
    assert rp.pcfg.flags & v == v
    d.write_text(text.replace("#define LAGRANGE_ADDEDMASS\n", "") + "#define LAGRANGE_INITIAL_VELOCITY_STOKES\n")
    assert lib.ch4g_read_run_params(str(d).encode(), pin.encode(), C.byref(rp)) == -1


def test_read_run_params_variable_gravity(tmp_path):
    """LAGRANGE_GRAVITY_VARIABLE reads the gravity_coeff ladder
    (parameters.c:589-600) and acts only inside LAGRANGE_GRAVITY."""
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    text = open(os.path.join(FIXP, "define.h")).read()
    base = open(os.path.join(FIXP, "param.in")).read()
    d, pin = tmp_path / "define.h", tmp_path / "param.in"
    d.write_text(text + "#define LAGRANGE_GRAVITY_VARIABLE\n#define LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG\n")
    pin.write_text(base)
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == -1
    pin.write_text(base + "gravity_coeff_types 3\ngravity_coeff_min 0.5\ngravity_coeff_max 2\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    v = abi.LAGRANGE_GRAVITY_VARIABLE | abi.LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG
    assert rp.pcfg.flags & v == v and rp.pcfg.gravity_coeff_max == 2.0
    d.write_text(text.replace("#define LAGRANGE_GRAVITY\n", "") + "#define LAGRANGE_GRAVITY_VARIABLE\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    assert not rp.pcfg.flags & v


def test_read_run_params_probes_grid_pairs(tmp_path):
    """LAGRANGE_ADD_TRACER_STATIC (inside LAGRANGE_ADD_TRACER) in the flag
    word, LAGRANGE_INITIAL_GRID / _PAIRS in the second word."""
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    text = open(os.path.join(FIXP, "define.h")).read()
    d, pin = tmp_path / "define.h", os.path.join(FIXP, "param.in")
    d.write_text(text + "#define LAGRANGE_ADD_TRACER_STATIC\n#define LAGRANGE_INITIAL_GRID\n"
                 "#define LAGRANGE_INITIAL_PAIRS\n")
    assert lib.ch4g_read_run_params(str(d).encode(), pin.encode(), C.byref(rp)) == 0
    assert rp.pcfg.flags & abi.LAGRANGE_ADD_TRACER_STATIC
    assert rp.pcfg.flags2 == abi.LAGRANGE_INITIAL_GRID | abi.LAGRANGE_INITIAL_PAIRS
    d.write_text(text + "#define LAGRANGE_ADDEDMASS_LIFT\n#define LAGRANGE_ADDEDMASS_WAKEDRAG\n")
    assert lib.ch4g_read_run_params(str(d).encode(), pin.encode(), C.byref(rp)) == 0
    assert rp.pcfg.flags2 == abi.LAGRANGE_ADDEDMASS_LIFT | abi.LAGRANGE_ADDEDMASS_WAKEDRAG
    d.write_text(text + "#define LAGRANGE_ORIENTATION_DRAG\n")
    assert lib.ch4g_read_run_params(str(d).encode(), pin.encode(), C.byref(rp)) == -1  # needs the Jeffery switch
    d.write_text(text.replace("#define LAGRANGE_ADDEDMASS\n", "") + "#define LAGRANGE_ADDEDMASS_LIFT\n")
    assert lib.ch4g_read_run_params(str(d).encode(), pin.encode(), C.byref(rp)) == 0
    assert rp.pcfg.flags2 == 0
    d.write_text(text.replace("#define LAGRANGE_ADD_TRACER\n", "") + "#define LAGRANGE_ADD_TRACER_STATIC\n")
    assert lib.ch4g_read_run_params(str(d).encode(), pin.encode(), C.byref(rp)) == 0
    assert not rp.pcfg.flags & abi.LAGRANGE_ADD_TRACER_STATIC


def test_read_run_params_radius_and_density(tmp_path):
    """LAGRANGE_RADIUSandDENSITY reads particle_radius_* and
    particle_density_* instead of tau_drag_* / beta_coeff_*
    (parameters.c:530-556)."""
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    text = open(os.path.join(FIXP, "define.h")).read()
    base = "\n".join(l for l in open(os.path.join(FIXP, "param.in")).read().splitlines()
                     if not l.startswith(("tau_drag", "beta_coeff"))) + "\n"
    d, pin = tmp_path / "define.h", tmp_path / "param.in"
    d.write_text(text + "#define LAGRANGE_RADIUSandDENSITY\n#define LAGRANGE_BC_RADIUSandDENSITY\n")
    pin.write_text(base)
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == -1
    pin.write_text(base + "particle_radius_types 2\nparticle_radius_min 0.5\nparticle_radius_max 1\n"
                   "particle_density_types 1\nparticle_density_min 2\nparticle_density_max 2\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    want = abi.LAGRANGE_RADIUSANDDENSITY | abi.LAGRANGE_BC_RADIUSANDDENSITY
    assert rp.pcfg.flags2 & want == want
    assert (rp.pcfg.particle_radius_types, rp.pcfg.particle_radius_max, rp.pcfg.particle_density_min) == (2.0, 1.0, 2.0)


def test_read_run_params_polymer(tmp_path):
    """LAGRANGE_POLYMER reads tau_polymer (parameters.c:729-733) and acts only
    with LAGRANGE_GRADIENT."""
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    text = open(os.path.join(FIXP, "define.h")).read()
    base = open(os.path.join(FIXP, "param.in")).read()
    d, pin = tmp_path / "define.h", tmp_path / "param.in"
    d.write_text(text + "#define LAGRANGE_POLYMER\n")
    pin.write_text(base)
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == -1
    pin.write_text(base + "tau_polymer 12.5\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    assert rp.pcfg.flags2 & abi.LAGRANGE_POLYMER and rp.pcfg.tau_polymer == 12.5


def test_read_run_params_jump(tmp_path):
    """LAGRANGE_ORIENTATION_ACTIVE_JUMP reads critical_shear_rate_* and
    jump_time_* (parameters.c:700-722); it and its HIGHPASS variant act only
    inside LAGRANGE_ORIENTATION_ACTIVE."""
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    text = open(os.path.join(FIXP, "define.h")).read()
    base = open(os.path.join(FIXP, "param.in")).read() + \
        "aspect_ratio_types 1\naspect_ratio_min 2\naspect_ratio_max 2\n" \
        "swim_velocity_types 1\nswim_velocity_min 0.01\nswim_velocity_max 0.01\n"
    d, pin = tmp_path / "define.h", tmp_path / "param.in"
    active = "#define LAGRANGE_ORIENTATION\n#define LAGRANGE_ORIENTATION_JEFFREY\n#define LAGRANGE_ORIENTATION_ACTIVE\n"
    jump = "#define LAGRANGE_ORIENTATION_ACTIVE_JUMP\n#define LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS\n"
    d.write_text(text + active + jump)
    pin.write_text(base)
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == -1
    pin.write_text(base + "critical_shear_rate_types 2\ncritical_shear_rate_min 1e-3\ncritical_shear_rate_max 2e-3\n"
                   "jump_time_types 1\njump_time_min 4\njump_time_max 4\n")
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    want = abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP | abi.LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS
    assert rp.pcfg.flags2 == want
    assert (rp.pcfg.critical_shear_rate_types, rp.pcfg.critical_shear_rate_max, rp.pcfg.jump_time_min) == \
        (2.0, 2e-3, 4.0)
    d.write_text(text + "#define LAGRANGE_ORIENTATION\n#define LAGRANGE_ORIENTATION_JEFFREY\n" + jump)
    assert lib.ch4g_read_run_params(str(d).encode(), str(pin).encode(), C.byref(rp)) == 0
    assert rp.pcfg.flags2 == 0


def test_read_run_params_rejects_unimplemented_switch(tmp_path):
    d = tmp_path / "define.h"
    d.write_text(open(os.path.join(FIX, "define.h")).read() + "#define LB_FLUID_BC_Y_P_SLIP\n")
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(str(d).encode(), os.path.join(FIX, "param.in").encode(), C.byref(rp)) == -1


def test_read_run_params_missing_key(tmp_path):
    p = tmp_path / "param.in"
    p.write_text("\n".join(l for l in open(os.path.join(FIX, "param.in")) if not l.startswith("tau_u")))
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(os.path.join(FIX, "define.h").encode(), str(p).encode(), C.byref(rp)) == -1


@pytest.mark.gpu
def test_ch4g_run_matches_oracle(tmp_path):
    """ch4g_run on the Poiseuille fixture writes velocity_averages.dat at
    t = 0, 250, ..., 1000 equal to the oracle's averages at those steps."""
    from oracle import Oracle
    import configs
    for f in ("define.h", "param.in"):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    r = subprocess.run([RUN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "The End" in r.stdout
    rows = np.loadtxt(tmp_path / "velocity_averages.dat")
    assert rows.shape == (5, 11)
    assert (tmp_path / "run_time.dat").exists()
    cfg = configs.poiseuille(4, 64, 4)
    cfg.Amp_x = 0.003072
    o = Oracle(cfg)
    o.initial_conditions()
    run = None
    for k, t in enumerate(rows[:, 0]):
        if k:
            o.step(250)
        a = o.dump_averages()
        assert t == 250 * k
        np.testing.assert_allclose(rows[k, 3], a["ux"], rtol=1e-6)  # %e output: 7 significant digits
        np.testing.assert_allclose(rows[k, 1], a["ene"], rtol=1e-6, atol=1e-30)
        _, _, ry, _ = o.dump_rulers()
        run = ry["ux"].copy() if run is None else run + ry["ux"]
    # y ruler of the last dump (rewritten each time) and the running average
    # over the 5 dumps (OUTPUT_AT_START: irun = itime / every + 1)
    ruler = np.loadtxt(tmp_path / "velocity_averages_y.dat")
    assert ruler.shape == (64, 11)
    np.testing.assert_array_equal(ruler[:, 0], np.arange(64) + 0.5)
    np.testing.assert_allclose(ruler[:, 3], ry["ux"], rtol=1e-6)
    running = np.loadtxt(tmp_path / "velocity_averages_y_run.dat")
    np.testing.assert_allclose(running[:, 3], run / 5, rtol=1e-6)
    assert np.loadtxt(tmp_path / "velocity_averages_x.dat").shape == (4, 11)


def particle_type_averages(a, ntypes, thermal, fluid_avg=False):
    """dump_particle_averages (lagrange.c:2598-2930) restated: per type, in
    particle order, sequential FP64 sums of v, v^2, v^4, a, a^2, a^4 (and t,
    t^2, t^4, dt_t, dt_t^2; then u, u^2, u^4 with
    LAGRANGE_OUTPUT_FLUID_AVERAGES), divided by the count."""
    rows = []
    for t in range(ntypes):
        sel = [r for r in a if int(r["name"]) % ntypes == t]
        cols = []
        for f in ("vx", "vy", "vz"):
            cols.append([float(r[f]) for r in sel])
        for f in ("vx", "vy", "vz"):
            cols.append([float(r[f]) * float(r[f]) for r in sel])
        for f in ("vx", "vy", "vz"):
            cols.append([float(r[f]) ** 4.0 for r in sel])
        for f in ("ax", "ay", "az"):
            cols.append([float(r[f]) for r in sel])
        for f in ("ax", "ay", "az"):
            cols.append([float(r[f]) * float(r[f]) for r in sel])
        for f in ("ax", "ay", "az"):
            cols.append([float(r[f]) ** 4.0 for r in sel])
        if thermal:
            cols.append([float(r["t"]) for r in sel])
            cols.append([float(r["t"]) * float(r["t"]) for r in sel])
            cols.append([float(r["t"]) ** 4.0 for r in sel])
            cols.append([float(r["dt_t"]) for r in sel])
            cols.append([float(r["dt_t"]) * float(r["dt_t"]) for r in sel])
        if fluid_avg:
            for f in ("ux", "uy", "uz"):
                cols.append([float(r[f]) for r in sel])
            for f in ("ux", "uy", "uz"):
                cols.append([float(r[f]) * float(r[f]) for r in sel])
            for f in ("ux", "uy", "uz"):
                cols.append([float(r[f]) ** 4.0 for r in sel])
        row = []
        for c in cols:
            acc = 0.0
            for v in c:
                acc += v
            row.append(acc * (1.0 / len(sel)))
        rows.append(row)
    return np.array(rows)


@pytest.mark.gpu
@pytest.mark.parametrize("fluid_avg", [False, True], ids=["plain", "fluid_averages"])
def test_ch4g_run_particles_matches_oracle(tmp_path, fluid_avg):
    """ch4g_run with particles (inelastic re-injection, added mass, two-way
    coupling) writes particle_averages.dat at t = 0, 40, 80, 120 equal (to
    the printed digits) to the oracle's particles at those steps; with
    LAGRANGE_OUTPUT_FLUID_AVERAGES also the moments of u at the particles."""
    from oracle import Oracle
    shutil.copy(os.path.join(FIXP, "param.in"), tmp_path / "param.in")
    (tmp_path / "define.h").write_text(open(os.path.join(FIXP, "define.h")).read() +
                                       ("#define LAGRANGE_OUTPUT_FLUID_AVERAGES\n" if fluid_avg else ""))
    r = subprocess.run([RUN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = np.loadtxt(tmp_path / "particle_averages.dat")
    ntypes = 2 * 2 + 1
    assert rows.shape == (4 * ntypes, 2 + 18 + 5 + (9 if fluid_avg else 0))
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(str(tmp_path / "define.h").encode(),
                                    str(tmp_path / "param.in").encode(), C.byref(rp)) == 0
    o = Oracle(rp.cfg)
    o.initial_conditions()
    o.particles_init(rp.pcfg)
    # particle_properties.dat (lagrange.c:644-700): count, tau_drag, beta_coeff per type
    a0 = o.particles()
    props = [l.split() for l in open(tmp_path / "particle_properties.dat")]
    assert len(props) == ntypes
    for t, row in enumerate(props):
        sel = a0[a0["name"].astype(int) % ntypes == t]
        assert row[:4] == ["type", str(t), "counter_all", str(len(sel))]
        assert row[4] == "tau_drag" and row[6] == "beta_coeff"
        assert float(row[5]) == pytest.approx(sel["tau_drag"][0], rel=1e-6)
        assert float(row[7]) == pytest.approx(sel["beta_coeff"][0], rel=1e-6)
    for k in range(4):
        if k:
            o.step(40)
        want = particle_type_averages(o.particles(), ntypes, True, fluid_avg)
        got = rows[k * ntypes:(k + 1) * ntypes]
        assert np.all(got[:, 0] == 40 * k) and np.array_equal(got[:, 1], np.arange(ntypes))
        np.testing.assert_allclose(got[:, 2:], want, rtol=2e-6, atol=1e-300)


H5LIB = "/opt/conda/lib/libhdf5.so"


def read_compound_h5(path, name, names, shape):
    """Read members `names` (doubles) of the compound dataset `name` through
    the HDF5 C library itself; the dataset's extent must equal `shape`."""
    h5 = C.CDLL(H5LIB)
    hid = C.c_int64
    h5.H5open()
    native_double = hid.in_dll(h5, "H5T_NATIVE_DOUBLE_g").value
    h5.H5Fopen.restype = hid
    h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
    h5.H5Dopen2.restype = hid
    h5.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
    h5.H5Dget_space.restype = hid
    h5.H5Dget_space.argtypes = [hid]
    h5.H5Sget_simple_extent_dims.argtypes = [hid, C.POINTER(C.c_uint64), C.c_void_p]
    h5.H5Tcreate.restype = hid
    h5.H5Tcreate.argtypes = [C.c_int, C.c_size_t]
    h5.H5Tinsert.argtypes = [hid, C.c_char_p, C.c_size_t, hid]
    h5.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    for f in ("H5Dclose", "H5Sclose", "H5Tclose", "H5Fclose"):
        getattr(h5, f).argtypes = [hid]
    f = h5.H5Fopen(str(path).encode(), 0, 0)
    assert f >= 0
    d = h5.H5Dopen2(f, name.encode(), 0)
    assert d >= 0, name
    sp = h5.H5Dget_space(d)
    dims = (C.c_uint64 * len(shape))()
    assert h5.H5Sget_simple_extent_dims(sp, dims, None) == len(shape)
    assert tuple(dims) == shape
    t = h5.H5Tcreate(6, 8 * len(names))  # H5T_COMPOUND
    for i, m in enumerate(names):
        assert h5.H5Tinsert(t, m.encode(), 8 * i, native_double) >= 0, m
    buf = np.empty(shape + (len(names),))
    assert h5.H5Dread(d, t, 0, 0, 0, buf.ctypes.data) >= 0
    for x, fn in ((t, "H5Tclose"), (sp, "H5Sclose"), (d, "H5Dclose"), (f, "H5Fclose")):
        getattr(h5, fn)(x)
    return buf


def read_pop_h5(path, name, shape):
    """One /population dataset: a [NZ][NY][NX] compound of doubles p0..p18."""
    return read_compound_h5(path, name, ["p%d" % i for i in range(19)], shape)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(H5LIB), reason="serial HDF5 not installed")
def test_pop_h5_dump_and_restart(tmp_path):
    """OUTPUT_H5: ch4g_run writes pop.h5 with the reference's layout
    (output_h5.c:347-460) and, with resume 1 / resume_u 1, a second run starts
    from it: 500 + 500 steps give the averages of 1000 straight steps."""
    from oracle import Oracle
    import configs
    d1, d2 = tmp_path / "a", tmp_path / "b"
    define = open(os.path.join(FIX, "define.h")).read() + "#define OUTPUT_H5\n"
    param = open(os.path.join(FIX, "param.in")).read().replace("time_end 1000.", "time_end 500.")
    for d, prm in ((d1, param), (d2, param.replace("resume 0", "resume 1").replace("resume_u 0", "resume_u 1"))):
        d.mkdir()
        (d / "define.h").write_text(define)
        (d / "param.in").write_text(prm)
    r = subprocess.run([RUN], cwd=d1, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    cfg = configs.poiseuille(4, 64, 4)
    cfg.Amp_x = 0.003072
    o = Oracle(cfg)
    o.initial_conditions()
    o.step(500)
    po, _ = o.get_populations()
    pop = read_pop_h5(d1 / "pop.h5", "/population/velocity", (4, 64, 4))
    assert np.array_equal(pop, po)  # FP64 path: bit-identical to the oracle
    shutil.copy(d1 / "pop.h5", d2 / "pop.h5")
    r = subprocess.run([RUN], cwd=d2, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "initialized from file" in r.stderr
    o.step(500)
    a = o.dump_averages()
    last = np.loadtxt(d2 / "velocity_averages.dat")[-1]
    assert last[0] == 500.0
    np.testing.assert_allclose(last[3], a["ux"], rtol=1e-6)
    np.testing.assert_allclose(last[1], a["ene"], rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(H5LIB), reason="serial HDF5 not installed")
@pytest.mark.parametrize("tpart", [False, True], ids=["plain", "particle_temperature"])
def test_part_h5_dump_and_restart(tmp_path, tpart):
    """OUTPUT_H5 with LAGRANGE: part.h5 holds the particle records of the
    last step under the reference's member names (equal to the oracle's, the
    FP64 path being bit-identical; with LAGRANGE_TEMPERATURE also
    temperature_particle and cp_particle), and a resume run reads them back."""
    d1, d2 = tmp_path / "a", tmp_path / "b"
    define = (open(os.path.join(FIXP, "define.h")).read() + "#define OUTPUT_H5\n" +
              ("#define LAGRANGE_TEMPERATURE\n" if tpart else ""))
    param = open(os.path.join(FIXP, "param.in")).read() + ("cp_coeff_types 2\ncp_coeff_min 0.5\ncp_coeff_max 2\n"
                                                            if tpart else "")
    for d, prm in ((d1, param), (d2, param.replace("resume 0", "resume 1").replace("resume_u 0", "resume_u 1")
                                  .replace("resume_t 0", "resume_t 1"))):
        d.mkdir()
        (d / "define.h").write_text(define)
        (d / "param.in").write_text(prm)
    r = subprocess.run([RUN], cwd=d1, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(str(d1 / "define.h").encode(), str(d1 / "param.in").encode(), C.byref(rp)) == 0
    from oracle import Oracle
    o = Oracle(rp.cfg)
    o.initial_conditions()
    o.particles_init(rp.pcfg)
    o.step(int(rp.time_end))
    ref = o.particles()
    names = ["name", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "ux", "uy", "uz", "tau_drag", "t", "dx_t",
             "dy_t", "dz_t"] + (["temperature_particle", "cp_particle"] if tpart else [])
    got = read_compound_h5(d1 / "part.h5", "/particles/point_particle", names, (rp.pcfg.particle_number,))
    n = len(ref)
    rename = {"temperature_particle": "t_p", "cp_particle": "cp_p"}
    for i, m in enumerate(names):
        np.testing.assert_array_equal(got[:n, i], ref[rename.get(m, m)], err_msg=m)
    for f in ("pop.h5", "part.h5"):
        shutil.copy(d1 / f, d2 / f)
    r = subprocess.run([RUN], cwd=d2, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "initialized from file" in r.stderr and "The End" in r.stdout


def read_dataset_h5(path, name, n):
    """One 1-D FP64 dataset of n entries, and the member names of its group."""
    h5 = C.CDLL(H5LIB)
    hid = C.c_int64
    h5.H5open()
    native_double = hid.in_dll(h5, "H5T_NATIVE_DOUBLE_g").value
    h5.H5Fopen.restype = hid
    h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
    h5.H5Dopen2.restype = hid
    h5.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
    h5.H5Dget_space.restype = hid
    h5.H5Dget_space.argtypes = [hid]
    h5.H5Sget_simple_extent_dims.argtypes = [hid, C.POINTER(C.c_uint64), C.c_void_p]
    h5.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    for f in ("H5Dclose", "H5Sclose", "H5Fclose"):
        getattr(h5, f).argtypes = [hid]
    f = h5.H5Fopen(str(path).encode(), 0, 0)
    assert f >= 0
    d = h5.H5Dopen2(f, name.encode(), 0)
    assert d >= 0, name
    sp = h5.H5Dget_space(d)
    dims = (C.c_uint64 * 1)()
    assert h5.H5Sget_simple_extent_dims(sp, dims, None) == 1 and dims[0] == n
    buf = np.empty(n)
    assert h5.H5Dread(d, native_double, 0, 0, 0, buf.ctypes.data) >= 0
    for x, fn in ((sp, "H5Sclose"), (d, "H5Dclose"), (f, "H5Fclose")):
        getattr(h5, fn)(x)
    return buf


ORIENT_DEFINES = ("#define LAGRANGE_ORIENTATION\n#define LAGRANGE_ORIENTATION_JEFFREY\n"
                  "#define LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS\n#define LAGRANGE_ORIENTATION_ACTIVE\n"
                  "#define LAGRANGE_ORIENTATION_SECONDORIENTATION\n")
ORIENT_PARAMS = ("aspect_ratio_types 3\naspect_ratio_min 0.5\naspect_ratio_max 4\n"
                 "gyrotaxis_velocity_types 1\ngyrotaxis_velocity_min 2\ngyrotaxis_velocity_max 2\n"
                 "swim_velocity_types 2\nswim_velocity_min 0\nswim_velocity_max 0.001\n")


TPART_DEFINES = "#define LAGRANGE_TEMPERATURE\n"
TPART_PARAMS = "cp_coeff_types 2\ncp_coeff_min 0.5\ncp_coeff_max 2\n"
GVAR_DEFINES = "#define LAGRANGE_GRAVITY_VARIABLE\n"
GVAR_PARAMS = "gravity_coeff_types 2\ngravity_coeff_min 0.5\ngravity_coeff_max 1.5\n"


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(H5LIB), reason="serial HDF5 not installed")
@pytest.mark.parametrize("orient,tpart,gvar,sed", [(False, False, False, False), (True, False, False, False),
                                                   (False, True, False, False), (False, False, True, False),
                                                   (False, False, False, True)],
                         ids=["plain", "orientation", "particle_temperature", "gravity_variable", "sediment"])
def test_particle_snapshots(tmp_path, orient, tpart, gvar, sed):
    """OUTPUT_H5 with LAGRANGE: output_particles writes RUN/particle_<itime>.h5
    every time_dump_lagr (lagrange.c:1598-2595), one FP64 dataset per member
    under /lagrange in the reference's order, equal to the oracle's records
    (FP64 path: bit-identical), and an XDMF descriptor naming those
    datasets; with the orientation switches also their members and the dp/dt
    moments of particle_averages.dat; with LAGRANGE_TEMPERATURE the
    particle's temperature and heat capacity (lagrange.c:2090-2103,
    :2515-2528) and the t_p column of particle_averages.dat."""
    import xml.etree.ElementTree as ET
    define = open(os.path.join(FIXP, "define.h")).read() + "#define OUTPUT_H5\n#define LAGRANGE_OUTPUT_FLUID\n"
    (tmp_path / "define.h").write_text(define + (ORIENT_DEFINES if orient else "") + (TPART_DEFINES if tpart else "")
                                       + (GVAR_DEFINES if gvar else "")
                                       + ("#define LAGRANGE_BC_INELASTIC_SEDIMENT\n" if sed else ""))
    (tmp_path / "param.in").write_text(open(os.path.join(FIXP, "param.in")).read() + (ORIENT_PARAMS if orient else "")
                                       + (TPART_PARAMS if tpart else "") + (GVAR_PARAMS if gvar else ""))
    r = subprocess.run([RUN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    for it in (0, 40, 80, 120):
        assert (tmp_path / "RUN" / f"particle_{it}.h5").exists() and (tmp_path / "RUN" / f"particle_{it}.xmf").exists()
    lib = C.CDLL(DROPIN)
    rp = RunParams()
    assert lib.ch4g_read_run_params(str(tmp_path / "define.h").encode(), str(tmp_path / "param.in").encode(),
                                    C.byref(rp)) == 0
    from oracle import Oracle
    o = Oracle(rp.cfg)
    o.initial_conditions()
    o.particles_init(rp.pcfg)
    o.step(120)
    ref = o.particles()
    n, npart = len(ref), rp.pcfg.particle_number
    # gradient + added mass + LAGRANGE_OUTPUT_FLUID + temperature, in order
    members = ["name", "tau_drag", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "dx_ux", "dy_ux", "dz_ux",
               "dx_uy", "dy_uy", "dz_uy", "dx_uz", "dy_uz", "dz_uz", "beta_coeff", "ux", "uy", "uz", "Dt_ux", "Dt_uy",
               "Dt_uz", "temperature", "dt_t", "dx_t", "dy_t", "dz_t"]
    if orient:
        i = members.index("temperature")
        members[i:i] = ["px", "py", "pz", "dt_px", "dt_py", "dt_pz", "nx", "ny", "nz", "dt_nx", "dt_ny", "dt_nz",
                        "aspect_ratio", "gyrotaxis_velocity", "swim_velocity"]
        assert np.unique(ref["aspect_ratio"]).size == 4  # the ladder and the tracer's 1
        rows = [l.split() for l in open(tmp_path / "particle_averages.dat")]
        assert {len(r) for r in rows} == {2 + 18 + 9 + 9 + 5}
    if gvar:
        i = members.index("az") + 1
        members[i:i] = ["gravity_coeff"]
        assert np.unique(ref["gravity_coeff"]).size == 3  # the ladder and the tracer's 0
    if tpart:
        members += ["temperature_particle", "cp_particle"]
        assert np.unique(ref["cp_p"]).size == 3  # the ladder and the tracer's 0
        rows = [l.split() for l in open(tmp_path / "particle_averages.dat")]
        assert {len(r) for r in rows} == {2 + 18 + 5 + 1}
    if sed:
        members += ["sediment"]
    path = tmp_path / "RUN" / "particle_120.h5"
    rename = {"temperature": "t", "temperature_particle": "t_p", "cp_particle": "cp_p"}
    for m in members:
        got = read_dataset_h5(path, "/lagrange/" + m, npart)
        want = ref[rename.get(m, m)]
        np.testing.assert_array_equal(got[:n], want, err_msg=m)
    root = ET.parse(tmp_path / "RUN" / "particle_120.xmf").getroot()
    refs = [e.text.strip() for e in root.iter("DataItem") if e.text and e.text.strip()]
    assert all(t.startswith("particle_120.h5:/lagrange/") for t in refs)
    assert {t.split(":/lagrange/")[1] for t in refs} == set(members)
    assert root.find("Domain/Grid/Topology").get("NumberOfElements") == str(npart)
