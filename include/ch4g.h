/*
 * ch4g.h — C-ABI of the MI355X lattice-Boltzmann time step.
 *
 * One lattice = one device-resident D3Q19 fluid (+ optional D3Q19 temperature)
 * population set on a single GPU, advanced by fused HIP kernels.  Every entry
 * point takes plain pointers and sizes, returns 0 on success and a negative
 * status on failure (message via ch4g_last_error()).  All calls for one lattice
 * must come from one host thread; work is ordered on the lattice's own stream.
 *
 * The reference (ch4-project, C/MPI) has no plugin API: its boundary is the set
 * of global-state functions declared in src/functions.h and called in a fixed
 * order by src/main.c:67-196.  Each entry point below names the reference
 * function(s) it replaces; include/ch4g_dropin.h re-exports the reference names
 * themselves on top of this header.
 *
 * Coordinates follow the reference: x is the fastest index
 * (src/typedef.h:342 IDX), y is the wall-normal / vertical axis of channel and
 * Rayleigh-Benard runs (src/forcing.c:665, src/output.c:338), cell centres sit
 * at i+1/2 (src/grid.c:183-185).
 *
 * Host population buffers are AoS, interior nodes only, layout
 * [NZ][NY][NX][19] (x fastest), population order and velocities of
 * src/lb.c:139-181 (D3Q19).
 */
#ifndef CH4G_H
#define CH4G_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CH4G_NPOP 19
#define CH4G_HIT_NK 6

/* Compile-time switches of the reference (src/define.h), carried at run time.
 * Names keep the reference's macro names without the LB_ prefix.          */
enum ch4g_flag {
    CH4G_LB_FLUID                      = 1u << 0,  /* define.h LB_FLUID */
    CH4G_LB_TEMPERATURE                = 1u << 1,  /* LB_TEMPERATURE */
    CH4G_FLUID_FORCING_POISEUILLE      = 1u << 2,  /* forcing.c:366-371 */
    CH4G_FLUID_FORCING_CHANNEL         = 1u << 3,  /* forcing.c:374-379 */
    CH4G_FLUID_FORCING_HIT             = 1u << 4,  /* forcing.c:530-592 default branch */
    CH4G_TEMPERATURE_BUOYANCY          = 1u << 5,  /* forcing.c:620-667 */
    CH4G_FLUID_FORCING_NOZEROMODE      = 1u << 6,  /* forcing.c:1123-1152 */
    CH4G_FLUID_BC_Y                    = 1u << 7,  /* no-slip walls at y, boundary_conditions.c:714-873 */
    CH4G_TEMPERATURE_BC_Y              = 1u << 8,  /* fixed-T walls at y, boundary_conditions.c:966-1078 */
    CH4G_FLUID_LES                     = 1u << 9,  /* Smagorinsky, les.c:5-269 */
    CH4G_FLUID_LES_LOCAL_STRAIN        = 1u << 10, /* |S| from the local stress, les.c:20-30 */
    CH4G_TEMPERATURE_BUOYANCY_T0_REF   = 1u << 11, /* theta = t - T_ref, forcing.c:623-624 */
    CH4G_TEMPERATURE_INITIAL_LINEAR    = 1u << 12, /* initial_conditions.c:303-308 */
    CH4G_TEMPERATURE_INITIAL_ADD_NOISE = 1u << 13, /* initial_conditions.c:421-424 */
    CH4G_FLUID_INITIAL_UNIT_DENSITY    = 1u << 14, /* initial_conditions.c:256-279 */
    CH4G_INITIAL_BAROMETRIC            = 1u << 15, /* initial_conditions.c:72-84 */
    CH4G_TEMPERATURE_LES               = 1u << 16  /* tau_t from tau_u_les, les.c:273-297 (needs FLUID_LES) */
};

/* Storage / arithmetic of the populations.  Reductions (Nusselt, energy,
 * profiles) are always accumulated in FP64.                                  */
enum ch4g_precision {
    CH4G_FP64     = 0,  /* f_i stored and collided in double (reference my_double) */
    CH4G_FP32_DEV = 1,  /* f_i - w_i*rho_b and g_i - w_i*T_b stored and collided in float */
    CH4G_FP32_STORE = 2 /* same float deviation storage, collision arithmetic in double */
};

/* Run parameters: the subset of the reference `property` struct
 * (src/typedef.h:75-224, read by parameters.c:41) used on the streaming step,
 * plus the run-time flag word above.  Lattice spacing and time step are 1
 * (parameters.c:135 requires time_dt == SX/NX under METHOD_STREAMING).      */
typedef struct ch4g_config {
    int32_t  NX, NY, NZ;
    int32_t  precision;        /* enum ch4g_precision */
    uint32_t flags;            /* OR of enum ch4g_flag */
    int32_t  seed;             /* srand48 seed (reference: rank + time(NULL), parallel.c:238-240) */
    double   tau_u, tau_t;     /* relaxation times; nu=(tau_u-1/2)/3, kappa=(tau_t-1/2)/3 (parameters.c:179,285) */
    double   Amp_x, Amp_y, Amp_z;
    double   gravity_x, gravity_y, gravity_z;
    double   beta_t, beta2_t;
    double   T_bot, T_top, T_ref;
    double   noise_t;          /* initial temperature noise amplitude */
    double   les_cs;           /* Smagorinsky constant (reference hard-codes 0.18, les.c:10) */
    double   rho0, T0;         /* base state subtracted in CH4G_FP32_DEV storage */
} ch4g_config;

/* Volume averages written by dump_averages (src/typedef.h:233-258 `output`,
 * fluid + temperature members, same order).  nux/nuy/nuz are Nusselt numbers
 * normalised by kappa*deltaT/SY (output.c:577-579).                          */
typedef struct ch4g_output {
    double x, y, z;
    double ux, uy, uz;
    double ux2, uy2, uz2;
    double rho, ene, eps;
    double rho2;
    double t, t2, epst;
    double dxt, dyt, dzt;
    double uxt, uyt, uzt;
    double nux, nuy, nuz;
    double lb;
} ch4g_output;

typedef struct ch4g_lattice ch4g_lattice;

/* Last error message of the calling thread ("" if none). */
const char *ch4g_last_error(void);

/* Number of GPUs visible to the library (0 if none). */
int ch4g_device_count(void);

/* Allocate a device-resident lattice on the current device.
 * Replaces allocate_fields (parameters.c:917) + design_lb (lb.c:3).
 * Fails with a negative status for unsupported flag combinations.           */
int ch4g_create(const ch4g_config *cfg, ch4g_lattice **out);

/* Release all device memory of the lattice.  Replaces free_fields. */
void ch4g_destroy(ch4g_lattice *lat);

/* Build the initial state on the device from the reference's initial
 * conditions: p = w_i (initial_conditions.c:45, optional unit-density
 * renormalisation :256-279, barometric :72-84), T linear + uniform noise drawn
 * with drand48 seeded by cfg->seed in k,j,i order (:303-308, :421-424),
 * g = w_i T (:486-487); then the first hydro_fields/build_forcing/hydro_fields
 * sequence of main.c:26-48.  HIT phases are drawn after the temperature field
 * (initialization_forcing, initial_conditions.c:866-871).                   */
int ch4g_initial_conditions(ch4g_lattice *lat);

/* Load / read back the pre-collision populations p (and g) of the current
 * step, AoS [NZ][NY][NX][19].  g may be NULL without LB_TEMPERATURE.
 * set = restart from a population dump (read_pop_h5, output_h5.c:540);
 * get = write_pop_h5 (output_h5.c:347) payload.  `set` recomputes
 * hydro_fields/build_forcing/hydro_fields as main.c:26-48 does.            */
int ch4g_set_populations(ch4g_lattice *lat, const double *p, const double *g);
int ch4g_get_populations(ch4g_lattice *lat, double *p, double *g);

/* Advance nsteps time steps of main.c:92-161 (compute_advection, add_collision
 * x2, build_forcing, add_forcing, boundary_and_pbc_conditions_for_streaming,
 * time_stepping x2, hydro_fields x2) with fused kernels: one launch per step,
 * or per two steps where ch4g_steps_per_launch reports 2 (bit-identical).
 * Asynchronous on the lattice stream.                                        */
int ch4g_step(ch4g_lattice *lat, int nsteps);

/* hydro_fields (lb.c:420-540): density, velocity (with the Guo half-force
 * shift) and temperature of the current step, [NZ][NY][NX] (u: x,y,z
 * interleaved).  Any pointer may be NULL.  Synchronous.                     */
int ch4g_hydro_fields(ch4g_lattice *lat, double *dens, double *u, double *t);

/* dump_averages (output.c:63-620): FP64 volume averages of the current step
 * including the Nusselt numbers.  Synchronous.                               */
int ch4g_dump_averages(ch4g_lattice *lat, ch4g_output *out);

/* No counterpart in the reference.  ch4g_dump_averages for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): the same averages, bit-identical,
 * with no device allocation.  The caller brings the FP64 field buffers the
 * averages are computed from, as device pointers: dens[N] and u[3N] (x,y,z
 * interleaved) for a fluid lattice, t[N] for a thermal one (others may be
 * NULL); on return they hold the fields ch4g_hydro_fields reports.  The
 * reduction's block partials and sums live in a static device array of
 * (2048 + 1) x 26 doubles (426 KB) per GPU, held until the process ends;
 * the reduction grid is capped at 2048 blocks, so every lattice size fits.
 * Calls that overlap on one GPU (several host threads, or lattices whose
 * streams run concurrently) race on that array.  Synchronous.               */
int ch4g_dump_averages_noalloc(ch4g_lattice *lat, double *dens, double *u, double *t, ch4g_output *out);

/* dump_averages rulers (output.c:212-491, normalisation :520-690): averages
 * over the y-z planes (rx[NX]), x-z planes (ry[NY]) and x-y planes (rz[NZ])
 * of the same members, with the reference's normalisation factors (its
 * temperature members of the z ruler are scaled by 1/(SY SZ), as the
 * reference does).  Each ruler's own coordinate member holds the plane
 * centre, the two other coordinate members are 0.  Any pointer may be
 * NULL.  These are the rows of velocity_averages_{x,y,z}.dat and
 * temperature_averages_{x,y,z}.dat.  Synchronous.                           */
int ch4g_dump_rulers(ch4g_lattice *lat, ch4g_output *rx, ch4g_output *ry, ch4g_output *rz);

/* HIT forcing phases phi_u[6][3] of the current step (forcing.c:49-79).
 * get/set let a host that owns the random stream drive the phase walk.      */
int ch4g_get_hit_phases(ch4g_lattice *lat, double *phi);
int ch4g_set_hit_phases(ch4g_lattice *lat, const double *phi);

/* ch4g_step bracketed by HIP events recorded on the lattice stream; waits
 * for the steps and returns their device time in *elapsed_ms.  Replaces the
 * MPI_Wtime bracket of main.c:54-56,198-201 and measure_time
 * (parallel.c:618-638, run_time.dat "time per step and grid point").      */
int ch4g_step_timed(ch4g_lattice *lat, int nsteps, double *elapsed_ms);

/* Replay steady-state steps from captured HIP graphs (default on).  Graphs
 * hold GRAPH_STEPS steps and are used once the first step has run, for runs
 * without per-step host forcing tables (HIT) or particles; results are
 * identical either way.  No reference counterpart (launch mechanics of this
 * library).                                                                 */
int ch4g_set_graphs(ch4g_lattice *lat, int enable);

/* Two steps per launch where ch4g_steps_per_launch says so (default on);
 * off: one step per launch.  Results are identical either way.  No
 * reference counterpart (launch mechanics of this library).                 */
int ch4g_set_fusion(ch4g_lattice *lat, int enable);

/* Steps one steady-state launch advances: 2 for float-storage lattices
 * without HIT forcing, LES, NOZEROMODE or two-way particles that are 2-D
 * (k_step2_fp32: two steps through an LDS tile) or 3-D with at least 2^21
 * nodes and rows that fill their 64-wide tiles (k_step2z_fp32: x-y tiles
 * marching along z through an LDS ring of planes), else 1.  Odd remainders
 * run as single steps; results are identical either way.  No reference
 * counterpart (launch mechanics of this library).                          */
int ch4g_steps_per_launch(const ch4g_lattice *lat, int *steps);

/* Wait for all work queued on the lattice stream. */
int ch4g_synchronize(ch4g_lattice *lat);

/* Current step counter (itime of main.c:69). */
int64_t ch4g_itime(const ch4g_lattice *lat);

/* Bytes of device memory held by the lattice. */
int64_t ch4g_device_bytes(const ch4g_lattice *lat);

/* ------------------------------------------------------------------------ */
/* Lagrangian point particles (src/lagrange.c, src/lagrange_feedback.c).
 *
 * Particles live on the lattice's device next to the populations and advance
 * inside ch4g_step after the fluid step, in main.c's order (main.c:163-183):
 * boundary_conditions_hydro, then time_dt_lagr_factor sub-steps of
 * interpolate_vector_at_particles(u) [+ interpolate_scalar_at_particles(t)]
 * and move_particles (with sendrecv_particles' wrap).  With two-way coupling
 * add_particle_feedbacks enters the force of every step from the second on
 * (main.c:125-130): per-node sums in the reference's particle order, kept in
 * double on FP64 lattices and rounded to float on float storage; it needs
 * (NX+1)(NY+1)(NZ+1) < 2^32.  Positions are FP64 in every precision mode.
 * Interpolated fields use the two-node halo of the reference's particle build
 * (define.h:333-334 LAGRANGE_GRADIENT_BRD2, typedef.h:315-316).            */

/* Particle switches (define.h:316-382), carried at run time. */
enum ch4g_lagrange_flag {
    CH4G_LAGRANGE_ADD_TRACER                          = 1u << 0,  /* type 0 is a fluid tracer, parameters.c:743-755 */
    CH4G_LAGRANGE_WRAP                                = 1u << 1,  /* positions wrapped into the box, lagrange.c:4075-4079 */
    CH4G_LAGRANGE_GRADIENT                            = 1u << 2,  /* grad u at particles, lagrange.c:1331-1436 */
    CH4G_LAGRANGE_ADDEDMASS                           = 1u << 3,  /* beta Du/Dt, lagrange.c:3325-3380 (needs GRADIENT) */
    CH4G_LAGRANGE_GRAVITY                             = 1u << 4,  /* -g (+beta g), lagrange.c:3312-3340 */
    CH4G_LAGRANGE_BC_INELASTIC                        = 1u << 5,  /* wall rule, lagrange.c:3760-3771 */
    CH4G_LAGRANGE_BC_INELASTIC_REINJECT               = 1u << 6,  /* lagrange.c:3806-3845, :3887-3925 */
    CH4G_LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM      = 1u << 7,
    CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO = 1u << 8,
    CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF = 1u << 9,
    CH4G_LAGRANGE_SMALLTAUD_BETA                      = 1u << 10, /* lagrange.c:3426-3445 */
    CH4G_LAGRANGE_SMALLTAUD_FLOATER                   = 1u << 11, /* lagrange.c:3446-3465 */
    CH4G_LAGRANGE_TWOWAY_MOMENTUM                     = 1u << 12, /* lagrange_feedback.c:11-256 */
    CH4G_LAGRANGE_TWOWAY_SMALLTAUD                    = 1u << 13, /* lagrange_feedback.c:106-110 */
    CH4G_LAGRANGE_TAUDRAG_INCREMENT_LOG               = 1u << 14, /* lagrange.c:271-297 */
    CH4G_LAGRANGE_ADDEDMASS_INCREMENT_LOG             = 1u << 15, /* lagrange.c:313-336 */
    CH4G_LAGRANGE_ORIENTATION                         = 1u << 16, /* orientation p, lagrange.c:877-901, :3470-3752 (needs GRADIENT) */
    CH4G_LAGRANGE_ORIENTATION_JEFFREY                 = 1u << 17, /* Jeffery equation, lagrange.c:3490-3680 */
    CH4G_LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG   = 1u << 18, /* aspect_ratio ladder, lagrange.c:399-421 */
    CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS       = 1u << 19, /* lagrange.c:445-470, :3533-3566 */
    CH4G_LAGRANGE_ORIENTATION_ACTIVE                  = 1u << 20, /* tracers swim along p, lagrange.c:502-521, :3179-3186 */
    CH4G_LAGRANGE_ORIENTATION_BC                      = 1u << 21, /* p_y turned away from a y wall, lagrange.c:3866-3870, :3942-3946 */
    CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION       = 1u << 22, /* second vector n normal to p, lagrange.c:904-924, :3586-3740 */
    CH4G_LAGRANGE_TEMPERATURE                         = 1u << 23, /* particle temperature t_p relaxing to t, lagrange.c:4007-4020 (needs LB_TEMPERATURE) */
    CH4G_LAGRANGE_TEMPERATURE_INCREMENT_LOG           = 1u << 24, /* cp_coeff ladder, lagrange.c:579-617 */
    CH4G_LAGRANGE_INITIAL_TEMPERATURE_FLUID           = 1u << 25, /* t_p starts at the fluid's t, lagrange.c:1004-1015 */
    CH4G_LAGRANGE_INITIAL_VELOCITY_FLUID              = 1u << 26, /* v starts at the fluid's u, lagrange.c:968-980 */
    CH4G_LAGRANGE_INITIAL_VELOCITY_STOKES             = 1u << 27, /* v starts at -g (1 - beta) tau, lagrange.c:981-991 (needs ADDEDMASS) */
    CH4G_LAGRANGE_GRAVITY_VARIABLE                    = 1u << 28, /* gravity scaled per type, lagrange.c:348-384, :3314-3335 (needs GRAVITY) */
    CH4G_LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG      = 1u << 29, /* gravity_coeff ladder, lagrange.c:356-376 */
    CH4G_LAGRANGE_OUTPUT_FLUID_AVERAGES               = 1u << 30, /* u, u^2, u^4 columns of particle_averages.dat, lagrange.c:2683-2916 (output only) */
    CH4G_LAGRANGE_ADD_TRACER_STATIC                   = 1u << 31  /* type 0 an Eulerian probe (tau_drag -1), type 1 the tracer, parameters.c:744-750, lagrange.c:280-282 */
};

/* Particle switches of the initial positions (define.h:321, :325), of the
 * radius/density parametrisation (define.h:328-330, :345), of the
 * added-mass forces (define.h:337-338) and of the spheroid drag (define.h:368),
 * a second word
 * (ch4g_particle_config.flags2).  The wake drag calls pow(): its results
 * follow the reference's to the last bits of the device's pow, not
 * bit for bit.                                                              */
enum ch4g_lagrange_flag2 {
    CH4G_LAGRANGE_INITIAL_GRID  = 1u << 0,  /* positions on a regular grid, lagrange.c:774-795 */
    CH4G_LAGRANGE_INITIAL_PAIRS = 1u << 1,  /* particle 2k-1 at distance 1 from particle 2k, lagrange.c:798-814 */
    CH4G_LAGRANGE_ADDEDMASS_LIFT     = 1u << 2,  /* lift (beta/3) (u - v) x omega, lagrange.c:3365-3380 (needs ADDEDMASS) */
    CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG = 1u << 3,  /* Schiller-Naumann drag 0.15 Re_p^0.687 / tau, lagrange.c:3242-3264 (needs ADDEDMASS) */
    CH4G_LAGRANGE_ORIENTATION_DRAG   = 1u << 4,  /* Stokes drag of a spheroid along p, lagrange.c:3267-3310 (needs ORIENTATION_JEFFREY) */
    CH4G_LAGRANGE_RADIUSANDDENSITY   = 1u << 5,  /* radius and density ladders define tau_drag (and beta), lagrange.c:150-260 */
    CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS  = 1u << 6,  /* lagrange.c:160-164, :177-180 */
    CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY = 1u << 7,  /* lagrange.c:199-203, :216-219 */
    CH4G_LAGRANGE_BC_RADIUSANDDENSITY = 1u << 8, /* walls at the particle radius sqrt(3 nu beta tau), lagrange.c:3777-3788 */
    /* swimming modes of LAGRANGE_ORIENTATION_ACTIVE (lagrange.c:3134-3178), the first set wins */
    CH4G_LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC            = 1u << 9,   /* v = swim p, no fluid velocity */
    CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE          = 1u << 10,  /* swim only while dt_t > 0 (needs LB_TEMPERATURE) */
    CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION    = 1u << 11,  /* swim along n */
    CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION = 1u << 12,  /* along p if aspect_ratio >= 1, else along n */
    CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT = 1u << 13, /* count of bottom-wall depositions, lagrange.c:3845-3863 (needs BC_INELASTIC) */
    CH4G_LAGRANGE_ORIENTATION_RANDOM    = 1u << 14, /* p drawn anew (random_vector) every sub-step, lagrange.c:3698-3713;
                                                       the device's acos / sin / cos: parity to their last bits */
    CH4G_LAGRANGE_POLYMER               = 1u << 15, /* conformation tensor along the trajectory, polymer.c:150-240 (needs GRADIENT) */
    /* copepod-like jumps of active tracers (lagrange.c:3048-3132; the first of the swimming modes):
       needs LAGRANGE_GRADIENT; the decay exp(-t / jump_time) is the device's exp, parity to its last bits */
    CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP             = 1u << 16,
    CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS    = 1u << 17,  /* jump below the critical shear rate */
    CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE = 1u << 18   /* react to t instead of the shear rate */
};

/* Particle parameters of param.in (parameters.c:155-167, :524-584). */
typedef struct ch4g_particle_config {
    uint32_t flags;            /* OR of enum ch4g_lagrange_flag */
    int32_t  particle_number;
    double   tau_drag_types, tau_drag_min, tau_drag_max;
    double   beta_coeff_types, beta_coeff_min, beta_coeff_max; /* read with LAGRANGE_ADDEDMASS */
    double   time_dt_lagr_factor;                              /* sub-steps per lattice step */
    double   aspect_ratio_types, aspect_ratio_min, aspect_ratio_max;                   /* ORIENTATION_JEFFREY */
    double   gyrotaxis_velocity_types, gyrotaxis_velocity_min, gyrotaxis_velocity_max; /* ..._GYROTAXIS */
    double   swim_velocity_types, swim_velocity_min, swim_velocity_max;                /* ORIENTATION_ACTIVE */
    double   cp_coeff_types, cp_coeff_min, cp_coeff_max;                               /* LAGRANGE_TEMPERATURE */
    double   gravity_coeff_types, gravity_coeff_min, gravity_coeff_max;                /* LAGRANGE_GRAVITY_VARIABLE */
    uint32_t flags2;           /* OR of enum ch4g_lagrange_flag2 */
    int32_t  reserved;         /* 0 */
    /* LAGRANGE_RADIUSandDENSITY: read instead of tau_drag_* / beta_coeff_* (parameters.c:530-556) */
    double   particle_radius_types, particle_radius_min, particle_radius_max;
    double   particle_density_types, particle_density_min, particle_density_max;
    double   tau_polymer;      /* LAGRANGE_POLYMER relaxation time (parameters.c:730-733) */
    double   critical_shear_rate_types, critical_shear_rate_min, critical_shear_rate_max;  /* ..._ACTIVE_JUMP */
    double   jump_time_types, jump_time_min, jump_time_max;
} ch4g_particle_config;

/* One particle: the members of the reference's point_particle
 * (typedef.h:391-481) that the supported switches read or write.           */
typedef struct ch4g_particle {
    double name;
    double x, y, z;
    double vx, vy, vz, vx_old, vy_old, vz_old;
    double ax, ay, az, ax_old, ay_old, az_old;
    double ux, uy, uz, ux_old, uy_old, uz_old;
    double dx_ux, dy_ux, dz_ux, dx_uy, dy_uy, dz_uy, dx_uz, dy_uz, dz_uz;
    double Dt_ux, Dt_uy, Dt_uz;
    double t, t_old, dt_t;
    double tau_drag, beta_coeff;
    double dx_t, dy_t, dz_t;   /* temperature gradient (LB_TEMPERATURE with LAGRANGE_GRADIENT) */
    double px, py, pz, dt_px, dt_py, dt_pz;             /* orientation and its last dp/dt (LAGRANGE_ORIENTATION) */
    double aspect_ratio, gyrotaxis_velocity, swim_velocity;
    double nx, ny, nz, dt_nx, dt_ny, dt_nz;             /* second orientation (..._SECONDORIENTATION) */
    double t_p, cp_p;                                   /* particle temperature and specific heat (LAGRANGE_TEMPERATURE) */
    double gravity_coeff;                               /* gravity factor of the particle's type (LAGRANGE_GRAVITY_VARIABLE) */
    double sediment;                                    /* bottom-wall depositions (LAGRANGE_BC_INELASTIC_SEDIMENT) */
    double cxx, cyy, czz, cxy, cyz, cxz;                /* polymer conformation tensor (LAGRANGE_POLYMER) */
    double dt_cxx, dt_cyy, dt_czz, dt_cxy, dt_cyz, dt_cxz;   /* and its last dc/dt */
    double critical_shear_rate, shear_rate, jump_time, time_from_jump;   /* LAGRANGE_ORIENTATION_ACTIVE_JUMP */
    double px_jump, py_jump, pz_jump;                   /* direction of the current jump */
} ch4g_particle;

/* No counterpart in the reference.  sizeof(ch4g_particle) as this library
 * was built: callers that exchange particle records check it against their
 * own header before ch4g_get_particles / ch4g_set_particles (the record grew
 * with the orientation, particle-temperature, gravity, sediment, polymer and jump members; a stale header
 * would mis-stride the arrays).                                             */
int64_t ch4g_particle_size(void);

/* allocate_particles (lagrange.c:6-49) + initial_conditions_particles
 * without restart (lagrange.c:72-1021): type ladders, name % types, positions
 * drawn x, y, z per particle from the lattice's drand48 stream (which has
 * already served initial_conditions and the first build_forcing, as in
 * main.c:23-51), then with LAGRANGE_ORIENTATION the particle's random unit
 * vector p (random_vector, random.c:139-150: two more draws), zero velocity
 * (or with LAGRANGE_INITIAL_VELOCITY_FLUID the fluid velocity interpolated
 * at the particle, u and with the gradient grad u set with it; with
 * LAGRANGE_INITIAL_VELOCITY_STOKES, which wins over it, -g (1 - beta) tau);
 * with LAGRANGE_TEMPERATURE the cp_coeff ladder (after the orientation
 * ones, lagrange.c:579-617) and t_p = T_bot, or with
 * LAGRANGE_INITIAL_TEMPERATURE_FLUID the fluid temperature interpolated at
 * the particle (t and, with the gradient, dx_t..dz_t set with it).  Call
 * after ch4g_initial_conditions (or ch4g_set_populations) and before the
 * first step.                                                               */
int ch4g_particles_init(ch4g_lattice *lat, const ch4g_particle_config *pcfg);

/* No counterpart in the reference, whose initial_conditions_particles keeps
 * its per-type tables local and prints them to particle_properties.dat
 * (lagrange.c:644-700): the values of each particle type (name % types) of
 * the last ch4g_particles_init, for writing that file.  *ntypes receives the
 * number of types; at most `cap` entries are written.                      */
typedef struct ch4g_particle_type {
    double count;                               /* particles of the type */
    double particle_radius, particle_density;   /* LAGRANGE_RADIUSandDENSITY */
    double tau_drag, beta_coeff, gravity_coeff, cp_coeff;
    double aspect_ratio, gyrotaxis_velocity, swim_velocity;
    double jump_time, critical_shear_rate;
} ch4g_particle_type;
int ch4g_particle_types(ch4g_lattice *lat, ch4g_particle_type *out, int cap, int *ntypes);

/* Number of particles alive (sendrecv_particles drops a particle whose
 * wrapped position rounds onto the upper box face, lagrange.c:4090-4116).
 * Synchronous.                                                              */
int ch4g_particles_count(ch4g_lattice *lat, int64_t *n);

/* Read back the particles in the reference's array order (output_particles,
 * lagrange.c:1598; write_point_particle_h5 payload).  At most `cap` records;
 * *n receives the live count.  Synchronous.                                 */
int ch4g_get_particles(ch4g_lattice *lat, ch4g_particle *out, int64_t cap, int64_t *n);

/* Replace the particle state (read_point_particle_h5 restart,
 * lagrange.c:128-131).  Requires ch4g_particles_init first; the count may
 * differ from particle_number.                                              */
int ch4g_set_particles(ch4g_lattice *lat, const ch4g_particle *in, int64_t n);

/* ------------------------------------------------------------------------ */
/* Serving: many independent lattices on one GPU.  No reference counterpart
 * (the reference runs one job per MPI launch, main.c); these entry points
 * are for hosts that run ensembles of the reference's jobs.                 */

/* Advance n lattices by nsteps steps together: one kernel launch per step
 * (per storage/population variant) covers every lattice through a table of
 * 256-node tiles.  Each lattice's fields are bit-identical to ch4g_step on
 * it alone.  Lattices must be past their first step (ch4g_step(lat, 1)),
 * without HIT forcing or particles, on the current device.                  */
int ch4g_step_batch(ch4g_lattice *const *lats, int n, int nsteps);

/* One job of a served ensemble. */
typedef struct ch4g_job {
    ch4g_config cfg;      /* lattice of the job */
    int64_t nsteps;       /* steps to run (main.c time loop length) */
    double arrival_s;     /* release time, seconds after ch4g_serve starts */
    double admit_s;       /* out: state resident, first step queued */
    double finish_s;      /* out: last step done, averages read back */
    ch4g_output out;      /* out: dump_averages of the last step */
    int32_t status;       /* out: 0 done, negative: not run (message in ch4g_last_error) */
    int32_t reserved;
} ch4g_job;

enum ch4g_serve_mode {
    CH4G_SERVE_BATCHED    = 0, /* continuous batching: admit while others run, one launch per step for all */
    CH4G_SERVE_SEQUENTIAL = 1  /* one job at a time through ch4g_step (the plain single-job path) */
};

enum ch4g_schedule {
    CH4G_SCHEDULE_LOCKSTEP     = 0, /* every job takes every step of the chunk (default) */
    CH4G_SCHEDULE_PROPORTIONAL = 1  /* each job's steps per chunk proportional to its remaining steps */
};

typedef struct ch4g_serve_params {
    int64_t hbm_budget;   /* device bytes the jobs may hold at once (0: 90% of free memory) */
    int32_t mode;         /* enum ch4g_serve_mode */
    int32_t chunk_steps;  /* launches between admission points (0: 128); replayed from a captured graph */
    int32_t max_active;   /* lattices resident at once (0: no limit but the budget) */
    int32_t use_graphs;   /* capture the chunk's launches in a hipGraph and replay it */
    int32_t schedule;     /* enum ch4g_schedule */
    int32_t reserved;
} ch4g_serve_params;

/* Run a request stream: jobs are released at arrival_s (wall clock from the
 * call), admitted in arrival order while their device footprint fits the
 * budget, advanced together, and retired (averages read, memory freed) as
 * they finish.  Returns when every job is done or rejected.                */
int ch4g_serve(const ch4g_serve_params *sp, ch4g_job *jobs, int32_t njobs);

/* Device bytes a lattice with this configuration would hold (admission). */
int64_t ch4g_footprint_bytes(const ch4g_config *cfg);

#ifdef __cplusplus
}
#endif

#endif /* CH4G_H */
