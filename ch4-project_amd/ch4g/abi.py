"""ctypes mirror of include/ch4g.h (structs, flags, precision modes).

Kept field-for-field identical to the C header; tests/test_abi.py checks the
sizes and offsets against the compiled library.
"""
import ctypes as C

NPOP = 19
HIT_NK = 6

# enum ch4g_flag (reference define.h switch names)
LB_FLUID = 1 << 0
LB_TEMPERATURE = 1 << 1
FLUID_FORCING_POISEUILLE = 1 << 2
FLUID_FORCING_CHANNEL = 1 << 3
FLUID_FORCING_HIT = 1 << 4
TEMPERATURE_BUOYANCY = 1 << 5
FLUID_FORCING_NOZEROMODE = 1 << 6
FLUID_BC_Y = 1 << 7
TEMPERATURE_BC_Y = 1 << 8
FLUID_LES = 1 << 9
FLUID_LES_LOCAL_STRAIN = 1 << 10
TEMPERATURE_BUOYANCY_T0_REF = 1 << 11
TEMPERATURE_INITIAL_LINEAR = 1 << 12
TEMPERATURE_INITIAL_ADD_NOISE = 1 << 13
FLUID_INITIAL_UNIT_DENSITY = 1 << 14
INITIAL_BAROMETRIC = 1 << 15
TEMPERATURE_LES = 1 << 16

# enum ch4g_lagrange_flag (reference LAGRANGE_* switch names)
LAGRANGE_ADD_TRACER = 1 << 0
LAGRANGE_WRAP = 1 << 1
LAGRANGE_GRADIENT = 1 << 2
LAGRANGE_ADDEDMASS = 1 << 3
LAGRANGE_GRAVITY = 1 << 4
LAGRANGE_BC_INELASTIC = 1 << 5
LAGRANGE_BC_INELASTIC_REINJECT = 1 << 6
LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM = 1 << 7
LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO = 1 << 8
LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF = 1 << 9
LAGRANGE_SMALLTAUD_BETA = 1 << 10
LAGRANGE_SMALLTAUD_FLOATER = 1 << 11
LAGRANGE_TWOWAY_MOMENTUM = 1 << 12
LAGRANGE_TWOWAY_SMALLTAUD = 1 << 13
LAGRANGE_TAUDRAG_INCREMENT_LOG = 1 << 14
LAGRANGE_ADDEDMASS_INCREMENT_LOG = 1 << 15
LAGRANGE_ORIENTATION = 1 << 16
LAGRANGE_ORIENTATION_JEFFREY = 1 << 17
LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG = 1 << 18
LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS = 1 << 19
LAGRANGE_ORIENTATION_ACTIVE = 1 << 20
LAGRANGE_ORIENTATION_BC = 1 << 21
LAGRANGE_ORIENTATION_SECONDORIENTATION = 1 << 22
LAGRANGE_TEMPERATURE = 1 << 23
LAGRANGE_TEMPERATURE_INCREMENT_LOG = 1 << 24
LAGRANGE_INITIAL_TEMPERATURE_FLUID = 1 << 25
LAGRANGE_INITIAL_VELOCITY_FLUID = 1 << 26
LAGRANGE_INITIAL_VELOCITY_STOKES = 1 << 27
LAGRANGE_GRAVITY_VARIABLE = 1 << 28
LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG = 1 << 29
LAGRANGE_OUTPUT_FLUID_AVERAGES = 1 << 30
LAGRANGE_ADD_TRACER_STATIC = 1 << 31
# enum ch4g_lagrange_flag2 (ParticleConfig.flags2)
LAGRANGE_INITIAL_GRID = 1 << 0
LAGRANGE_INITIAL_PAIRS = 1 << 1
LAGRANGE_ADDEDMASS_LIFT = 1 << 2
LAGRANGE_ADDEDMASS_WAKEDRAG = 1 << 3
LAGRANGE_ORIENTATION_DRAG = 1 << 4
LAGRANGE_RADIUSANDDENSITY = 1 << 5
LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS = 1 << 6
LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY = 1 << 7
LAGRANGE_BC_RADIUSANDDENSITY = 1 << 8
LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC = 1 << 9
LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE = 1 << 10
LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION = 1 << 11
LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION = 1 << 12
LAGRANGE_BC_INELASTIC_SEDIMENT = 1 << 13
LAGRANGE_ORIENTATION_RANDOM = 1 << 14
LAGRANGE_POLYMER = 1 << 15
LAGRANGE_ORIENTATION_ACTIVE_JUMP = 1 << 16
LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS = 1 << 17
LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE = 1 << 18

FP64 = 0
FP32_DEV = 1
FP32_STORE = 2


class Config(C.Structure):
    _fields_ = [
        ("NX", C.c_int32), ("NY", C.c_int32), ("NZ", C.c_int32),
        ("precision", C.c_int32),
        ("flags", C.c_uint32),
        ("seed", C.c_int32),
        ("tau_u", C.c_double), ("tau_t", C.c_double),
        ("Amp_x", C.c_double), ("Amp_y", C.c_double), ("Amp_z", C.c_double),
        ("gravity_x", C.c_double), ("gravity_y", C.c_double), ("gravity_z", C.c_double),
        ("beta_t", C.c_double), ("beta2_t", C.c_double),
        ("T_bot", C.c_double), ("T_top", C.c_double), ("T_ref", C.c_double),
        ("noise_t", C.c_double),
        ("les_cs", C.c_double),
        ("rho0", C.c_double), ("T0", C.c_double),
    ]

    def copy(self, **kw):
        c = Config()
        C.pointer(c)[0] = self
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def nodes(self):
        return int(self.NX) * int(self.NY) * int(self.NZ)

    def has(self, flag):
        return (int(self.flags) & flag) != 0


OUTPUT_FIELDS = ["x", "y", "z", "ux", "uy", "uz", "ux2", "uy2", "uz2", "rho", "ene", "eps", "rho2",
                 "t", "t2", "epst", "dxt", "dyt", "dzt", "uxt", "uyt", "uzt", "nux", "nuy", "nuz", "lb"]


class Output(C.Structure):
    _fields_ = [(n, C.c_double) for n in OUTPUT_FIELDS]

    def as_dict(self):
        return {n: float(getattr(self, n)) for n in OUTPUT_FIELDS}


def output_dtype():
    """numpy record layout of ch4g_output (rulers are arrays of it)."""
    import numpy as np
    return np.dtype([(n, np.float64) for n in OUTPUT_FIELDS])


def make_config(NX, NY, NZ, flags, precision=FP64, **kw):
    """Config with the reference's defaults (rho0 = 1, T0 = mean wall T)."""
    c = Config()
    c.NX, c.NY, c.NZ = NX, NY, NZ
    c.flags = flags
    c.precision = precision
    c.seed = kw.pop("seed", 12345)
    c.tau_u = kw.pop("tau_u", 1.0)
    c.tau_t = kw.pop("tau_t", 1.0)
    c.rho0 = kw.pop("rho0", 1.0)
    c.les_cs = kw.pop("les_cs", 0.18)
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    if "T0" not in kw:
        c.T0 = 0.5 * (c.T_bot + c.T_top)
    return c


class ParticleConfig(C.Structure):
    _fields_ = [
        ("flags", C.c_uint32),
        ("particle_number", C.c_int32),
        ("tau_drag_types", C.c_double), ("tau_drag_min", C.c_double), ("tau_drag_max", C.c_double),
        ("beta_coeff_types", C.c_double), ("beta_coeff_min", C.c_double), ("beta_coeff_max", C.c_double),
        ("time_dt_lagr_factor", C.c_double),
        ("aspect_ratio_types", C.c_double), ("aspect_ratio_min", C.c_double), ("aspect_ratio_max", C.c_double),
        ("gyrotaxis_velocity_types", C.c_double), ("gyrotaxis_velocity_min", C.c_double),
        ("gyrotaxis_velocity_max", C.c_double),
        ("swim_velocity_types", C.c_double), ("swim_velocity_min", C.c_double), ("swim_velocity_max", C.c_double),
        ("cp_coeff_types", C.c_double), ("cp_coeff_min", C.c_double), ("cp_coeff_max", C.c_double),
        ("gravity_coeff_types", C.c_double), ("gravity_coeff_min", C.c_double), ("gravity_coeff_max", C.c_double),
        ("flags2", C.c_uint32), ("reserved", C.c_int32),
        ("particle_radius_types", C.c_double), ("particle_radius_min", C.c_double),
        ("particle_radius_max", C.c_double),
        ("particle_density_types", C.c_double), ("particle_density_min", C.c_double),
        ("particle_density_max", C.c_double),
        ("tau_polymer", C.c_double),
        ("critical_shear_rate_types", C.c_double), ("critical_shear_rate_min", C.c_double),
        ("critical_shear_rate_max", C.c_double),
        ("jump_time_types", C.c_double), ("jump_time_min", C.c_double), ("jump_time_max", C.c_double),
    ]


def make_particle_config(flags, particle_number, tau_drag=(1, 0.0, 0.0), beta_coeff=(1, 1.0, 1.0),
                         time_dt_lagr_factor=1.0, aspect_ratio=(1, 1.0, 1.0), gyrotaxis_velocity=(1, 0.0, 0.0),
                         swim_velocity=(1, 0.0, 0.0), cp_coeff=(1, 1.0, 1.0), gravity_coeff=(1, 1.0, 1.0),
                         flags2=0, particle_radius=(1, 0.0, 0.0), particle_density=(1, 1.0, 1.0), tau_polymer=1.0,
                         critical_shear_rate=(1, 0.0, 0.0), jump_time=(1, 1.0, 1.0)):
    """ParticleConfig from (types, min, max) ladders as param.in gives them."""
    pc = ParticleConfig()
    pc.flags = flags
    pc.particle_number = particle_number
    pc.tau_drag_types, pc.tau_drag_min, pc.tau_drag_max = (float(v) for v in tau_drag)
    pc.beta_coeff_types, pc.beta_coeff_min, pc.beta_coeff_max = (float(v) for v in beta_coeff)
    pc.time_dt_lagr_factor = float(time_dt_lagr_factor)
    pc.aspect_ratio_types, pc.aspect_ratio_min, pc.aspect_ratio_max = (float(v) for v in aspect_ratio)
    (pc.gyrotaxis_velocity_types, pc.gyrotaxis_velocity_min,
     pc.gyrotaxis_velocity_max) = (float(v) for v in gyrotaxis_velocity)
    pc.swim_velocity_types, pc.swim_velocity_min, pc.swim_velocity_max = (float(v) for v in swim_velocity)
    pc.cp_coeff_types, pc.cp_coeff_min, pc.cp_coeff_max = (float(v) for v in cp_coeff)
    pc.gravity_coeff_types, pc.gravity_coeff_min, pc.gravity_coeff_max = (float(v) for v in gravity_coeff)
    pc.flags2 = flags2
    (pc.particle_radius_types, pc.particle_radius_min,
     pc.particle_radius_max) = (float(v) for v in particle_radius)
    (pc.particle_density_types, pc.particle_density_min,
     pc.particle_density_max) = (float(v) for v in particle_density)
    pc.tau_polymer = float(tau_polymer)
    (pc.critical_shear_rate_types, pc.critical_shear_rate_min,
     pc.critical_shear_rate_max) = (float(v) for v in critical_shear_rate)
    pc.jump_time_types, pc.jump_time_min, pc.jump_time_max = (float(v) for v in jump_time)
    return pc


PARTICLE_FIELDS = ["name", "x", "y", "z", "vx", "vy", "vz", "vx_old", "vy_old", "vz_old",
                   "ax", "ay", "az", "ax_old", "ay_old", "az_old", "ux", "uy", "uz", "ux_old", "uy_old", "uz_old",
                   "dx_ux", "dy_ux", "dz_ux", "dx_uy", "dy_uy", "dz_uy", "dx_uz", "dy_uz", "dz_uz",
                   "Dt_ux", "Dt_uy", "Dt_uz", "t", "t_old", "dt_t", "tau_drag", "beta_coeff", "dx_t", "dy_t", "dz_t",
                   "px", "py", "pz", "dt_px", "dt_py", "dt_pz", "aspect_ratio", "gyrotaxis_velocity", "swim_velocity",
                   "nx", "ny", "nz", "dt_nx", "dt_ny", "dt_nz", "t_p", "cp_p",
                   "gravity_coeff", "sediment", "cxx", "cyy", "czz", "cxy", "cyz", "cxz",
                   "dt_cxx", "dt_cyy", "dt_czz", "dt_cxy", "dt_cyz", "dt_cxz",
                   "critical_shear_rate", "shear_rate", "jump_time", "time_from_jump", "px_jump", "py_jump", "pz_jump"]


class Particle(C.Structure):
    _fields_ = [(n, C.c_double) for n in PARTICLE_FIELDS]


class ParticleType(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("count", "particle_radius", "particle_density", "tau_drag", "beta_coeff",
                                          "gravity_coeff", "cp_coeff", "aspect_ratio", "gyrotaxis_velocity",
                                          "swim_velocity", "jump_time", "critical_shear_rate")]


def particle_dtype():
    """numpy structured dtype with the layout of struct ch4g_particle."""
    import numpy as np
    return np.dtype([(n, np.float64) for n in PARTICLE_FIELDS])


SERVE_BATCHED = 0
SERVE_SEQUENTIAL = 1


class Job(C.Structure):
    _fields_ = [
        ("cfg", Config),
        ("nsteps", C.c_int64),
        ("arrival_s", C.c_double),
        ("admit_s", C.c_double),
        ("finish_s", C.c_double),
        ("out", Output),
        ("status", C.c_int32),
        ("reserved", C.c_int32),
    ]


class ServeParams(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_int64),
        ("mode", C.c_int32),
        ("chunk_steps", C.c_int32),
        ("max_active", C.c_int32),
        ("use_graphs", C.c_int32),
        ("schedule", C.c_int32),
        ("reserved", C.c_int32),
    ]


SCHEDULE_LOCKSTEP = 0
SCHEDULE_PROPORTIONAL = 1
