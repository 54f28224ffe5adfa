/*
 * ch4g_dropin.h — the reference's per-step functions, with the reference's
 * names and signatures (src/functions.h), implemented on one device-resident
 * lattice (include/ch4g.h).  A main program written against the reference's
 * call sequence (src/main.c:23-201) links against libch4g_dropin.so instead
 * of lb.o fluid.o forcing.o boundary_conditions.o output.o parallel.o.
 *
 * The reference works on global arrays (p, rhs_p, g, rhs_g, u, dens, t,
 * force; src/common_main.h); here the device lattice owns the state and the
 * pop* arguments are accepted as handles only (they may be NULL).  Host copies
 * of the fields are produced on demand by ch4g_hydro_fields / dump_averages.
 *
 * Per-step semantics of the calls (each has the reference's order contract):
 *   compute_advection, add_collision, build_forcing, add_forcing,
 *   boundary_and_pbc_conditions_for_streaming : record the stage; the work
 *       they name is fused into the step kernel;
 *   time_stepping : the first call of a step launches the fused step
 *       (collision + forcing + walls + streaming of both population sets),
 *       the second (the other population set) completes it;
 *   hydro_fields : marks the fields of the new step current;
 *   dump_averages : at the reference's cadence, FP64 averages on the device,
 *       appended to velocity_averages.dat / temperature_averages.dat;
 *   measure_time : writes run_time.dat like parallel.c:618-638.
 * Calls out of the reference's order are reported on stderr and exit, as the
 * reference does for its own consistency checks.
 */
#ifndef CH4G_DROPIN_H
#define CH4G_DROPIN_H

#include "ch4g.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef double my_double;                    /* src/typedef.h:7-14 default */
typedef struct { my_double p[CH4G_NPOP]; } pop; /* src/typedef.h:66-68, D3Q19 */
typedef struct { my_double x, y, z; } vector;   /* src/typedef.h vector */

/* Run description: the lattice configuration plus the time-loop and output
 * parameters of the reference's `property` (src/typedef.h:75-224).        */
typedef struct ch4g_run_params {
    ch4g_config cfg;
    double time_dt, time_start, time_end, time_dump_field, time_dump_diagn;
    int output_at_start; /* define.h OUTPUT_AT_START */
    int diagn_append;    /* define.h OUTPUT_DIAGN_APPEND: ruler files appended, not rewritten */
    int output_h5;       /* define.h OUTPUT_H5: pop.h5 written at the end of the run */
    int resume, resume_u, resume_t; /* param.in resume flags (parameters.c:78-88) */
    int lagrange;        /* define.h LAGRANGE */
    ch4g_particle_config pcfg;
    double time_dump_lagr;
    int lagr_output_fluid; /* define.h LAGRANGE_OUTPUT_FLUID: u (and Dt_u) in the particle snapshots */
} ch4g_run_params;

/* Read a reference-style define.h (active #define names) and param.in
 * (name value pairs, first match wins; parameters.c:3-39, :41-800).
 * Returns 0, or -1 after printing the reason on stderr as the reference does.
 * Optional keys of this implementation: precision (0 FP64, 1 FP32-deviation,
 * 2 float storage with double arithmetic), les_cs, seed.  With LAGRANGE the
 * particle keys of parameters.c:155-167, :524-584 are read as well.        */
int ch4g_read_run_params(const char *define_h, const char *param_in, ch4g_run_params *out);

/* allocate_fields (parameters.c:917) + design_lb (lb.c:3) on the device. */
int ch4g_dropin_init(const ch4g_run_params *rp);
/* free_fields */
void ch4g_dropin_finalize(void);
/* the lattice behind the drop-in functions (for callers mixing both APIs) */
ch4g_lattice *ch4g_dropin_lattice(void);
/* step counter and simulated time of the drop-in run (main.c itime, time_now) */
int ch4g_dropin_itime(void);

/* src/functions.h, same names and signatures */
void initial_conditions(int restart);                                              /* initial_conditions.c:3 */
void initialization_forcing(void);                                                 /* initial_conditions.c:822 */
void compute_advection(pop *f, pop *rhs_f, my_double tau, pop *f_eq, char which_pop); /* fluid.c:4 */
void add_collision(pop *f, pop *rhs_f, my_double tau, pop *f_eq, char which_pop);     /* fluid.c:1033 */
void build_forcing(void);                                                          /* forcing.c:5 */
void add_forcing(void);                                                            /* forcing.c:1250 */
void boundary_and_pbc_conditions_for_streaming(void);                              /* boundary_conditions.c:419 */
void time_stepping(pop *f, pop *rhs_f, pop *old_rhs_f, pop *old_old_rhs_f, my_double tau, pop *f_eq,
                   char which_pop);                                                /* lb.c:588 */
void hydro_fields(char which_pop);                                                 /* lb.c:420 */
void dump_averages(void);                                                          /* output.c:63 */
/* pop.h5 (output_h5.c:347-540, :540-729): /population/velocity and
 * /population/temperature, compound p0..p18 of doubles over [NZ][NY][NX],
 * and with HIT forcing the /forcing group ("number of modes (nk)", phi_u).
 * Serial HDF5; without HDF5 at build time both report an error and exit. */
void write_pop_h5(void);
void read_pop_h5(void);
void measure_time(void);                                                           /* parallel.c:618 */

/* Lagrangian particles (src/functions.h).  The particles advance inside the
 * fused step launched by time_stepping (feedback before the collision,
 * interpolation + move after the streaming, in main.c's order); these calls
 * check the call order and do the host-side work the reference does there. */
void allocate_particles(void);                                          /* lagrange.c:6 */
void initial_conditions_particles(int restart);                         /* lagrange.c:72 */
void add_particle_feedbacks(void);                                      /* lagrange_feedback.c:11 */
void boundary_conditions_hydro(void);                                   /* boundary_conditions.c:2786 */
void interpolate_vector_at_particles(vector *f, char which_vector);      /* lagrange.c:1211 */
void interpolate_scalar_at_particles(my_double *f, char which_scalar);  /* lagrange.c:1442 */
void move_particles(void);                                              /* lagrange.c:2934 */
void output_particles(void);                                            /* lagrange.c:1598 */
void dump_particle_averages(void);                                      /* lagrange.c:2598 */
/* part.h5 (lagrange.c:4231, :4548): /particles/point_particle compound of
 * the carried point_particle members over [particle_number], /properties
 * particle_types.  Serial HDF5; without it both report an error and exit. */
void write_point_particle_h5(void);                                     /* lagrange.c:4231 */
void read_point_particle_h5(void);                                      /* lagrange.c:4548 */

#ifdef __cplusplus
}
#endif

#endif /* CH4G_DROPIN_H */
