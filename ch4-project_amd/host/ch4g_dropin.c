/*
 * Drop-in implementation of the reference's per-step functions
 * (include/ch4g_dropin.h) on one device-resident lattice.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../include/ch4g_dropin.h"

/* stages of one reference time step (main.c:92-183) */
enum { ST_IDLE = 0, ST_ADVECTION, ST_COLLISION, ST_FORCING, ST_BC, ST_STREAMED, ST_HYDRO, ST_PARTICLES };

static struct {
    ch4g_run_params rp;
    ch4g_lattice *lat;
    int stage;
    int itime;
    int launched;   /* the fused step of the current itime has been queued */
    double t1, t2;  /* wall-clock bracket of the time loop (main.c:54-56, 198-201) */
    int timing;
    int particles;  /* initial_conditions_particles done */
    int moves;      /* move_particles calls of the current step */
    ch4g_output *ruler[3];   /* x, y, z rulers of the last dump_averages */
    ch4g_output *running[3]; /* their running sums (ruler_*_running) */
} S;

static void write_particle_snapshot(void);

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void die(const char *what)
{
    const char *e = ch4g_last_error();
    fprintf(stderr, "Error message -> %s%s%s. Exit.\n", what, (e && *e) ? ": " : "", (e && *e) ? e : "");
    exit(-1);
}

static void need_lattice(const char *fn)
{
    if (!S.lat) {
        fprintf(stderr, "Error message -> %s called before ch4g_dropin_init. Exit.\n", fn);
        exit(-1);
    }
}

static void order(const char *fn, int allowed_from, int allowed_to, int next)
{
    need_lattice(fn);
    if (S.stage < allowed_from || S.stage > allowed_to) {
        fprintf(stderr, "Error message -> %s called out of the main.c step order (stage %d). Exit.\n", fn, S.stage);
        exit(-1);
    }
    if (next > S.stage) S.stage = next;
}

int ch4g_dropin_init(const ch4g_run_params *rp)
{
    if (S.lat) ch4g_dropin_finalize();
    memset(&S, 0, sizeof S);
    S.rp = *rp;
    if (ch4g_particle_size() != (int64_t)sizeof(ch4g_particle)) {
        fprintf(stderr, "Error message -> ch4g_particle is %lld bytes in libch4g, %zu in this build. Exit.\n",
                (long long)ch4g_particle_size(), sizeof(ch4g_particle));
        return -1;
    }
    return ch4g_create(&rp->cfg, &S.lat);
}

void ch4g_dropin_finalize(void)
{
    if (S.lat) ch4g_destroy(S.lat);
    S.lat = NULL;
    for (int a = 0; a < 3; a++) {
        free(S.ruler[a]);
        free(S.running[a]);
        S.ruler[a] = S.running[a] = NULL;
    }
}

ch4g_lattice *ch4g_dropin_lattice(void) { return S.lat; }

int ch4g_dropin_itime(void) { return S.itime; }

/* initial_conditions + the first hydro_fields / build_forcing / hydro_fields
 * of main.c:26-48 (the device lattice builds all of them together).  With
 * restart the populations (resume_u, resume_t) and the HIT phases (resume_u)
 * then come from pop.h5 when it exists (initial_conditions.c:599-612); the
 * 18 phase draws the reference skips with resume_u are still taken from the
 * drand48 stream here (the reference's stream is time-seeded in any case). */
void initial_conditions(int restart)
{
    need_lattice("initial_conditions");
    if (ch4g_initial_conditions(S.lat)) die("initial_conditions");
    S.itime = 0;
    S.stage = ST_IDLE;
    FILE *fin = fopen("pop.h5", "r");
    if (restart && fin != NULL) {
        fclose(fin);
        read_pop_h5();
        fprintf(stderr, "The pop.h5 file is present!\n Populations are initialized from file.\n");
    } else {
        if (fin) fclose(fin);
        if (restart) fprintf(stderr, "Warning message -> pop.h5 file is missing!\n Populations are initialized from scratch.\n");
    }
}

/* HIT phases are drawn inside ch4g_initial_conditions in the reference's
 * order (after the temperature noise), so this call has nothing left to do. */
void initialization_forcing(void) { need_lattice("initialization_forcing"); }

void compute_advection(pop *f, pop *rhs_f, my_double tau, pop *f_eq, char which_pop)
{
    (void)f, (void)rhs_f, (void)tau, (void)f_eq, (void)which_pop;
    if (S.stage == ST_HYDRO || S.stage == ST_STREAMED || S.stage == ST_PARTICLES) S.stage = ST_IDLE; /* new step */
    order("compute_advection", ST_IDLE, ST_COLLISION, ST_ADVECTION);
    if (!S.timing) {
        S.t1 = now_s();
        S.timing = 1;
    }
}

void add_collision(pop *f, pop *rhs_f, my_double tau, pop *f_eq, char which_pop)
{
    (void)f, (void)rhs_f, (void)tau, (void)f_eq, (void)which_pop;
    order("add_collision", ST_ADVECTION, ST_COLLISION, ST_COLLISION);
}

void build_forcing(void) { order("build_forcing", ST_COLLISION, ST_FORCING, ST_FORCING); }

void add_forcing(void) { order("add_forcing", ST_FORCING, ST_FORCING, ST_FORCING); }

void boundary_and_pbc_conditions_for_streaming(void)
{
    order("boundary_and_pbc_conditions_for_streaming", ST_COLLISION, ST_FORCING, ST_BC);
}

void time_stepping(pop *f, pop *rhs_f, pop *old_rhs_f, pop *old_old_rhs_f, my_double tau, pop *f_eq,
                   char which_pop)
{
    (void)f, (void)rhs_f, (void)old_rhs_f, (void)old_old_rhs_f, (void)tau, (void)f_eq, (void)which_pop;
    if (S.stage == ST_STREAMED) return; /* second population set of the same step */
    order("time_stepping", ST_BC, ST_BC, ST_STREAMED);
    if (ch4g_step(S.lat, 1)) die("time_stepping");
    S.itime++;
}

void hydro_fields(char which_pop)
{
    (void)which_pop;
    order("hydro_fields", ST_STREAMED, ST_HYDRO, ST_HYDRO);
}

/* one ruler file: rows of the ruler's coordinate and the members listed
 * (offsets into ch4g_output), divided by irun for the running averages */
static void write_ruler(const char *name, int append, const ch4g_output *r, int n, int coord, const size_t *m,
                        int nm, double irun)
{
    FILE *fo = fopen(name, append ? "a" : "w");
    if (!fo) return;
    for (int i = 0; i < n; i++) {
        const double *v = (const double *)&r[i];
        fprintf(fo, "%e", v[coord] / irun);
        for (int k = 0; k < nm; k++) fprintf(fo, " %e", v[m[k] / sizeof(double)] / irun);
        fprintf(fo, "\n");
    }
    fclose(fo);
}

#define OFF(f) offsetof(ch4g_output, f)
static const size_t VEL_COLS[] = {OFF(ene), OFF(rho), OFF(ux), OFF(uy), OFF(uz), OFF(ux2),
                                  OFF(uy2), OFF(uz2), OFF(eps), OFF(rho2)};
static const size_t TEMP_COLS[] = {OFF(t), OFF(t2), OFF(epst), OFF(dxt), OFF(dyt), OFF(dzt), OFF(uxt),
                                   OFF(uyt), OFF(uzt), OFF(nux), OFF(nuy), OFF(nuz), OFF(lb)};
#undef OFF

/* dump_averages (output.c:63-1101): every time_dump_diagn the FP64 volume
 * averages appended to velocity_averages.dat / temperature_averages.dat
 * (:773, :864), the x, y, z rulers to *_averages_{x,y,z}.dat (rewritten, or
 * appended with OUTPUT_DIAGN_APPEND), their running sums divided by irun to
 * *_averages_{x,y,z}_run.dat (:765-768, :802-829), and with HIT forcing the
 * turbulence scales to velocity_turbulence.dat (:832-857).                  */
void dump_averages(void)
{
    need_lattice("dump_averages");
    const ch4g_run_params *rp = &S.rp;
    int every = (int)(rp->time_dump_diagn / rp->time_dt);
    if (every <= 0 || S.itime % every != 0) return;
    int irun = (int)(S.itime / every);
    if (rp->output_at_start) irun += 1;
    ch4g_output o;
    if (ch4g_dump_averages(S.lat, &o)) die("dump_averages");
    const int n[3] = {rp->cfg.NX, rp->cfg.NY, rp->cfg.NZ};
    if (!S.ruler[0]) {
        for (int a = 0; a < 3; a++) {
            S.ruler[a] = (ch4g_output *)calloc((size_t)n[a], sizeof(ch4g_output));
            S.running[a] = (ch4g_output *)calloc((size_t)n[a], sizeof(ch4g_output));
            if (!S.ruler[a] || !S.running[a]) die("dump_averages");
        }
    }
    if (ch4g_dump_rulers(S.lat, S.ruler[0], S.ruler[1], S.ruler[2])) die("dump_averages");
    for (int a = 0; a < 3; a++) /* add_output(ruler_*_running, ruler_*) */
        for (int i = 0; i < n[a]; i++) {
            double *d = (double *)&S.running[a][i];
            const double *v = (const double *)&S.ruler[a][i];
            for (size_t k = 0; k < sizeof(ch4g_output) / sizeof(double); k++) d[k] += v[k];
        }
    double time_now = rp->time_start + S.itime * rp->time_dt;
    static const char *AX[3] = {"x", "y", "z"};
    char name[64];
    if (rp->cfg.flags & CH4G_LB_FLUID) {
        FILE *fo = fopen("velocity_averages.dat", "a");
        if (fo) {
            fprintf(fo, "%e %e %e %e %e %e %e %e %e %e %e\n", time_now, o.ene, o.rho, o.ux, o.uy, o.uz, o.ux2, o.uy2,
                    o.uz2, o.eps, o.rho2);
            fclose(fo);
        }
        for (int a = 0; a < 3; a++) {
            snprintf(name, sizeof name, "velocity_averages_%s.dat", AX[a]);
            write_ruler(name, rp->diagn_append, S.ruler[a], n[a], a, VEL_COLS, 10, 1.0);
            snprintf(name, sizeof name, "velocity_averages_%s_run.dat", AX[a]);
            write_ruler(name, rp->diagn_append, S.running[a], n[a], a, VEL_COLS, 10, (double)irun);
        }
        if (rp->cfg.flags & CH4G_FLUID_FORCING_HIT) {
            const double nu = (rp->cfg.tau_u - 0.5 * rp->time_dt) / 3.0;
            double u_prime = sqrt(((o.ux2 - o.ux * o.ux) + (o.uy2 - o.uy * o.uy) + (o.uz2 - o.uz * o.uz)) / 3.0);
            double lambda = sqrt(15.0 * nu * u_prime * u_prime / o.eps);
            double Re_lambda = lambda * u_prime / nu;
            double lk = pow(pow(nu, 3.0) / o.eps, 0.25);
            double tk = pow(nu / o.eps, 0.5);
            double vk = pow(nu * o.eps, 0.25);
            double k_turb = 1.5 * pow(u_prime, 2.0);
            double lL = pow(u_prime, 3.0) / o.eps;
            double tL = k_turb / o.eps;
            double vL = u_prime;
            FILE *ft = fopen("velocity_turbulence.dat", "a");
            if (ft) {
                fprintf(ft, "%e %e %e %e %e %e %e %e %e\n", time_now, Re_lambda, lambda, lk, tk, vk, lL, tL, vL);
                fclose(ft);
            }
        }
    }
    if (rp->cfg.flags & CH4G_LB_TEMPERATURE) {
        FILE *fo = fopen("temperature_averages.dat", "a");
        if (fo) {
            fprintf(fo, "%e %e %e %e %e %e %e %e %e %e %e %e %e %e\n", time_now, o.t, o.t2, o.epst, o.dxt, o.dyt,
                    o.dzt, o.uxt, o.uyt, o.uzt, o.nux, o.nuy, o.nuz, o.lb);
            fclose(fo);
        }
        for (int a = 0; a < 3; a++) {
            snprintf(name, sizeof name, "temperature_averages_%s.dat", AX[a]);
            write_ruler(name, rp->diagn_append, S.ruler[a], n[a], a, TEMP_COLS, 13, 1.0);
            snprintf(name, sizeof name, "temperature_averages_%s_run.dat", AX[a]);
            write_ruler(name, rp->diagn_append, S.running[a], n[a], a, TEMP_COLS, 13, (double)irun);
        }
    }
}

/* ---- Lagrangian particles ---------------------------------------------- */

static void need_particles(const char *fn)
{
    need_lattice(fn);
    if (!S.rp.lagrange) {
        fprintf(stderr, "Error message -> %s called without LAGRANGE in define.h. Exit.\n", fn);
        exit(-1);
    }
}

/* the particle arrays are allocated with the initial conditions */
void allocate_particles(void) { need_particles("allocate_particles"); }

/* initial_conditions_particles (lagrange.c:72-1021): types, positions from
 * the lattice's drand48 stream; with restart the records then come from
 * part.h5 (read_point_particle_h5, lagrange.c:128-131).                    */
/* particle_properties.dat (lagrange.c:644-700): one line per type with its
 * count and the property values the switches define.                      */
static void write_particle_properties(void)
{
    int nt = 0;
    if (ch4g_particle_types(S.lat, NULL, 0, &nt)) die("initial_conditions_particles");
    ch4g_particle_type *y = (ch4g_particle_type *)calloc((size_t)(nt > 0 ? nt : 1), sizeof *y);
    if (!y) die("initial_conditions_particles: out of memory");
    if (ch4g_particle_types(S.lat, y, nt, &nt)) die("initial_conditions_particles");
    const unsigned fl = S.rp.pcfg.flags, f2 = S.rp.pcfg.flags2;
    const int grad = (fl & CH4G_LAGRANGE_GRADIENT) != 0;
    FILE *fo = fopen("particle_properties.dat", "w");
    if (fo) {
        for (int i = 0; i < nt; i++) {
            fprintf(fo, "type %d counter_all %d ", i, (int)y[i].count);
            if (f2 & CH4G_LAGRANGE_RADIUSANDDENSITY) {
                fprintf(fo, "particle_radius %e ", y[i].particle_radius);
                fprintf(fo, "particle_density %e ", y[i].particle_density);
            }
            fprintf(fo, "tau_drag %e ", y[i].tau_drag);
            if (grad && (fl & CH4G_LAGRANGE_ADDEDMASS)) fprintf(fo, "beta_coeff %e ", y[i].beta_coeff);
            if (fl & CH4G_LAGRANGE_GRAVITY_VARIABLE) fprintf(fo, "gravity_coeff %e ", y[i].gravity_coeff);
            if (fl & CH4G_LAGRANGE_TEMPERATURE) fprintf(fo, "cp_coeff %e ", y[i].cp_coeff);
            if (grad && (fl & CH4G_LAGRANGE_ORIENTATION)) {
                if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY) {
                    fprintf(fo, "aspect_ratio %e ", y[i].aspect_ratio);
                    if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS)
                        fprintf(fo, "gyrotaxis_velocity %e ", y[i].gyrotaxis_velocity);
                }
                if (fl & CH4G_LAGRANGE_ORIENTATION_ACTIVE) {
                    fprintf(fo, "swim_velocity %e ", y[i].swim_velocity);
                    if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) {
                        fprintf(fo, "jump_time %e ", y[i].jump_time);
                        fprintf(fo, "critical_shear_rate %e ", y[i].critical_shear_rate);
                    }
                }
            }
            fprintf(fo, "\n");
        }
        fclose(fo);
    }
    free(y);
}

void initial_conditions_particles(int restart)
{
    need_particles("initial_conditions_particles");
    if (S.itime != 0) {
        fprintf(stderr, "Error message -> initial_conditions_particles after the first step. Exit.\n");
        exit(-1);
    }
    /* lagrange.c:123-134: with restart and part.h5 present the records come
     * from the file and none of the initial draws are taken; otherwise the
     * full initialisation */
    FILE *fin = restart ? fopen("part.h5", "r") : NULL;
    if (fin) {
        fclose(fin);
        ch4g_particle_config pc = S.rp.pcfg;
        pc.particle_number = 0;
        pc.flags2 &= ~(uint32_t)CH4G_LAGRANGE_INITIAL_GRID;
        if (ch4g_particles_init(S.lat, &pc)) die("initial_conditions_particles");
        S.particles = 1;
        read_point_particle_h5();
        fprintf(stderr, "The part.h5 file is present!\n Particles will be initialized from file.\n");
        return;
    }
    if (ch4g_particles_init(S.lat, &S.rp.pcfg)) die("initial_conditions_particles");
    S.particles = 1;
    write_particle_properties();
}

/* from the second step on, between build_forcing and add_forcing
 * (main.c:125-130); the deposit is part of the fused step */
void add_particle_feedbacks(void)
{
    need_particles("add_particle_feedbacks");
    order("add_particle_feedbacks", ST_FORCING, ST_FORCING, ST_FORCING);
    if (!(S.rp.pcfg.flags & CH4G_LAGRANGE_TWOWAY_MOMENTUM)) {
        fprintf(stderr, "Error message -> add_particle_feedbacks without LAGRANGE_TWOWAY(_MOMENTUM). Exit.\n");
        exit(-1);
    }
}

/* the halo rules are applied where the particle kernel reads the fields */
void boundary_conditions_hydro(void)
{
    need_lattice("boundary_conditions_hydro");
    order("boundary_conditions_hydro", ST_HYDRO, ST_PARTICLES, ST_PARTICLES);
    S.moves = 0;
}

void interpolate_vector_at_particles(vector *f, char which_vector)
{
    (void)f, (void)which_vector;
    need_particles("interpolate_vector_at_particles");
    order("interpolate_vector_at_particles", ST_PARTICLES, ST_PARTICLES, ST_PARTICLES);
}

void interpolate_scalar_at_particles(my_double *f, char which_scalar)
{
    (void)f, (void)which_scalar;
    need_particles("interpolate_scalar_at_particles");
    order("interpolate_scalar_at_particles", ST_PARTICLES, ST_PARTICLES, ST_PARTICLES);
}

/* one of the time_dt_lagr_factor sub-steps (main.c:169-182) */
void move_particles(void)
{
    need_particles("move_particles");
    order("move_particles", ST_PARTICLES, ST_PARTICLES, ST_PARTICLES);
    if (++S.moves > (int)S.rp.pcfg.time_dt_lagr_factor) {
        fprintf(stderr, "Error message -> more move_particles calls than time_dt_lagr_factor. Exit.\n");
        exit(-1);
    }
}

/* output_particles (lagrange.c:1598-2595): with OUTPUT_H5, every
 * time_dump_lagr, the particle snapshot RUN/particle_<itime>.h5 and its XDMF
 * descriptor (write_particle_snapshot) */
void output_particles(void)
{
    need_particles("output_particles");
    const ch4g_run_params *rp = &S.rp;
    const int every = (int)(rp->time_dump_lagr / rp->time_dt);
    if (!rp->output_h5 || every <= 0 || S.itime % every != 0) return;
    write_particle_snapshot();
}

/* property.particle_types (parameters.c:559-584, :621-695, :743-755): the
 * product of the ladder sizes the switches read, plus the fluid tracer. */
static double ptypes(const ch4g_particle_config *pc)
{
    const unsigned fl = pc->flags;
    double types = pc->tau_drag_types;
    if ((fl & CH4G_LAGRANGE_GRADIENT) && (fl & CH4G_LAGRANGE_ADDEDMASS)) types *= pc->beta_coeff_types;
    if (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY) /* parameters.c:541, :554 */
        types = pc->particle_radius_types * pc->particle_density_types;
    if (fl & CH4G_LAGRANGE_GRAVITY_VARIABLE) types *= pc->gravity_coeff_types; /* parameters.c:600 */
    if ((fl & CH4G_LAGRANGE_GRADIENT) && (fl & CH4G_LAGRANGE_ORIENTATION)) {
        if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY) {
            types *= pc->aspect_ratio_types;
            if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS) types *= pc->gyrotaxis_velocity_types;
        }
        if (fl & CH4G_LAGRANGE_ORIENTATION_ACTIVE) types *= pc->swim_velocity_types;
        if (pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) /* parameters.c:709, :720 */
            types *= pc->critical_shear_rate_types * pc->jump_time_types;
    }
    if (fl & CH4G_LAGRANGE_TEMPERATURE) types *= pc->cp_coeff_types; /* parameters.c:615 */
    if (fl & CH4G_LAGRANGE_ADD_TRACER) types += (fl & CH4G_LAGRANGE_ADD_TRACER_STATIC) ? 2.0 : 1.0;
    return types;
}

/* dump_particle_averages (lagrange.c:2598-2930): per particle type, FP64
 * averages of v, v^2, v^4, a, a^2, a^4 (and t, t^2, t^4, dt_t, dt_t^2 with
 * LB_TEMPERATURE, then t_p with LAGRANGE_TEMPERATURE; the fluid velocity at
 * the particles, u, u^2, u^4, with LAGRANGE_OUTPUT_FLUID_AVERAGES),
 * accumulated in particle order, appended to particle_averages.dat every
 * time_dump_diagn.                                                         */
void dump_particle_averages(void)
{
    need_particles("dump_particle_averages");
    const ch4g_run_params *rp = &S.rp;
    int every = (int)(rp->time_dump_diagn / rp->time_dt);
    if (every <= 0 || S.itime % every != 0) return;
    int64_t n = 0;
    if (ch4g_particles_count(S.lat, &n)) die("dump_particle_averages");
    ch4g_particle *a = (ch4g_particle *)malloc(sizeof(ch4g_particle) * (size_t)(n > 0 ? n : 1));
    if (!a) die("dump_particle_averages: out of memory");
    if (ch4g_get_particles(S.lat, a, n, &n)) die("dump_particle_averages");
    const int nt = (int)ptypes(&rp->pcfg);
    const unsigned fl = rp->pcfg.flags;
    const int orient = (fl & CH4G_LAGRANGE_GRADIENT) && (fl & CH4G_LAGRANGE_ORIENTATION);
    const int second = orient && (fl & CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION);
    const int tpart = (fl & CH4G_LAGRANGE_TEMPERATURE) != 0;
    const int favg = (fl & CH4G_LAGRANGE_OUTPUT_FLUID_AVERAGES) != 0;
    enum { NV = 51 };
    double *acc = (double *)calloc((size_t)nt * NV, sizeof(double));
    double *cnt = (double *)calloc((size_t)nt, sizeof(double));
    if (!acc || !cnt) die("dump_particle_averages: out of memory");
    for (int64_t i = 0; i < n; i++) {
        const ch4g_particle *p = &a[i];
        int type = ((int)p->name) % nt;
        double *o = acc + (size_t)type * NV;
        cnt[type] += 1.0;
        o[0] += p->vx;
        o[1] += p->vy;
        o[2] += p->vz;
        o[3] += p->vx * p->vx;
        o[4] += p->vy * p->vy;
        o[5] += p->vz * p->vz;
        o[6] += pow(p->vx, 4.0);
        o[7] += pow(p->vy, 4.0);
        o[8] += pow(p->vz, 4.0);
        o[9] += p->ax;
        o[10] += p->ay;
        o[11] += p->az;
        o[12] += p->ax * p->ax;
        o[13] += p->ay * p->ay;
        o[14] += p->az * p->az;
        o[15] += pow(p->ax, 4.0);
        o[16] += pow(p->ay, 4.0);
        o[17] += pow(p->az, 4.0);
        o[18] += p->t;
        o[19] += p->t * p->t;
        o[20] += pow(p->t, 4.0);
        o[21] += p->dt_t;
        o[22] += p->dt_t * p->dt_t;
        o[23] += p->dt_px;
        o[24] += p->dt_py;
        o[25] += p->dt_pz;
        o[26] += p->dt_px * p->dt_px;
        o[27] += p->dt_py * p->dt_py;
        o[28] += p->dt_pz * p->dt_pz;
        o[29] += pow(p->dt_px, 4.0);
        o[30] += pow(p->dt_py, 4.0);
        o[31] += pow(p->dt_pz, 4.0);
        o[32] += p->dt_nx;
        o[33] += p->dt_ny;
        o[34] += p->dt_nz;
        o[35] += p->dt_nx * p->dt_nx;
        o[36] += p->dt_ny * p->dt_ny;
        o[37] += p->dt_nz * p->dt_nz;
        o[38] += pow(p->dt_nx, 4.0);
        o[39] += pow(p->dt_ny, 4.0);
        o[40] += pow(p->dt_nz, 4.0);
        o[41] += p->t_p;
        o[42] += p->ux; /* lagrange.c:2769-2782 */
        o[43] += p->uy;
        o[44] += p->uz;
        o[45] += p->ux * p->ux;
        o[46] += p->uy * p->uy;
        o[47] += p->uz * p->uz;
        o[48] += pow(p->ux, 4.0);
        o[49] += pow(p->uy, 4.0);
        o[50] += pow(p->uz, 4.0);
    }
    FILE *fo = fopen("particle_averages.dat", "a");
    if (fo) {
        double time_now = rp->time_start + S.itime * rp->time_dt;
        for (int type = 0; type < nt; type++) {
            double norm = 1.0 / cnt[type];
            double *o = acc + (size_t)type * NV;
            fprintf(fo, "%e %e ", time_now, (double)type);
            for (int k = 0; k < 18; k++) fprintf(fo, "%e ", o[k] * norm);
            if (orient) /* lagrange.c:2891-2896: dp/dt moments */
                for (int k = 23; k < 32; k++) fprintf(fo, "%e ", o[k] * norm);
            if (second) /* lagrange.c:2897-2902: dn/dt moments */
                for (int k = 32; k < 41; k++) fprintf(fo, "%e ", o[k] * norm);
            if (rp->cfg.flags & CH4G_LB_TEMPERATURE) {
                for (int k = 18; k < 23; k++) fprintf(fo, "%e ", o[k] * norm);
                if (tpart) fprintf(fo, "%e ", o[41] * norm); /* lagrange.c:2908-2911 */
            }
            if (favg) /* lagrange.c:2913-2917 */
                for (int k = 42; k < 51; k++) fprintf(fo, "%e ", o[k] * norm);
            fprintf(fo, "\n");
        }
        fclose(fo);
    }
    free(acc);
    free(cnt);
    free(a);
}

/* measure_time (parallel.c:618-638): run_time.dat of the time loop. */
void measure_time(void)
{
    need_lattice("measure_time");
    if (ch4g_synchronize(S.lat)) die("measure_time");
    S.t2 = now_s();
    double dt = S.timing ? S.t2 - S.t1 : 0.0;
    const ch4g_config *c = &S.rp.cfg;
    double npts = (double)c->NX * (double)c->NY * (double)c->NZ;
    int it = S.itime > 0 ? S.itime : 1;
    remove("run_time.dat");
    FILE *fo = fopen("run_time.dat", "w");
    if (!fo) return;
    fprintf(fo, "Total execution time %e\n", dt);
    fprintf(stdout, "Total execution time %e\n", dt);
    fprintf(fo, "Time steps %d\n", S.itime);
    fprintf(fo, "Execution time per time step %e\n", dt / (double)it);
    fprintf(fo, "Execution time per grid point %e\n", dt / npts);
    fprintf(fo, "Execution time per time step and grid point %e\n", dt / ((double)it * npts));
    fprintf(fo, "Time tick on this machine %e\n", 1e-9);
    fprintf(fo, "Number of processes for this run %d\n", 1);
    fprintf(fo, "Execution time per process %e\n", dt);
    fclose(fo);
}

/* ---- pop.h5 (output_h5.c:347-729) ------------------------------------- */

#ifdef CH4G_HDF5
#include <hdf5.h>

/* compound type of one pop record: p0..p18 doubles */
static hid_t pop_type(void)
{
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(pop));
    char label[16];
    for (int i = 0; i < CH4G_NPOP; i++) {
        snprintf(label, sizeof label, "p%d", i);
        H5Tinsert(t, label, HOFFSET(pop, p[i]), H5T_NATIVE_DOUBLE);
    }
    return t;
}

static hid_t vector_type(void)
{
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(vector));
    H5Tinsert(t, "x", HOFFSET(vector, x), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "y", HOFFSET(vector, y), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "z", HOFFSET(vector, z), H5T_NATIVE_DOUBLE);
    return t;
}

void write_pop_h5(void)
{
    need_lattice("write_pop_h5");
    const ch4g_config *c = &S.rp.cfg;
    const size_t n = (size_t)c->NX * c->NY * c->NZ;
    const int fluid = (c->flags & CH4G_LB_FLUID) != 0, thermal = (c->flags & CH4G_LB_TEMPERATURE) != 0;
    pop *p = fluid ? (pop *)malloc(n * sizeof(pop)) : NULL;
    pop *g = thermal ? (pop *)malloc(n * sizeof(pop)) : NULL;
    if ((fluid && !p) || (thermal && !g)) die("write_pop_h5");
    if (ch4g_get_populations(S.lat, (double *)p, (double *)g)) die("write_pop_h5");
    fprintf(stderr, "Writing file pop.h5\n");
    hsize_t dims[3] = {(hsize_t)c->NZ, (hsize_t)c->NY, (hsize_t)c->NX};
    hid_t file = H5Fcreate("pop.h5", H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) die("write_pop_h5");
    hid_t group = H5Gcreate2(file, "/population", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hid_t space = H5Screate_simple(3, dims, NULL);
    hid_t pt = pop_type();
    const char *names[2] = {"velocity", "temperature"};
    pop *src[2] = {p, g};
    for (int k = 0; k < 2; k++) {
        if (!src[k]) continue;
        hid_t d = H5Dcreate2(group, names[k], pt, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (d < 0 || H5Dwrite(d, pt, H5S_ALL, H5S_ALL, H5P_DEFAULT, src[k]) < 0) die("write_pop_h5");
        H5Dclose(d);
    }
    H5Tclose(pt);
    H5Sclose(space);
    H5Gclose(group);
    if (c->flags & CH4G_FLUID_FORCING_HIT) {
        /* /forcing: scalar attribute nk and the phase array phi_u */
        double phi[CH4G_HIT_NK][3];
        if (ch4g_get_hit_phases(S.lat, &phi[0][0])) die("write_pop_h5");
        vector v[CH4G_HIT_NK];
        for (int i = 0; i < CH4G_HIT_NK; i++) {
            v[i].x = phi[i][0];
            v[i].y = phi[i][1];
            v[i].z = phi[i][2];
        }
        int nk = CH4G_HIT_NK;
        hid_t fg = H5Gcreate2(file, "/forcing", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        hid_t a1s = H5Screate(H5S_SCALAR);
        hid_t a1 = H5Acreate2(fg, "number of modes (nk)", H5T_NATIVE_INT, a1s, H5P_DEFAULT, H5P_DEFAULT);
        H5Awrite(a1, H5T_NATIVE_INT, &nk);
        H5Aclose(a1);
        H5Sclose(a1s);
        hsize_t adim[1] = {(hsize_t)nk};
        hid_t a2s = H5Screate_simple(1, adim, NULL);
        hid_t vt = vector_type();
        hid_t a2 = H5Acreate2(fg, "phi_u", vt, a2s, H5P_DEFAULT, H5P_DEFAULT);
        H5Awrite(a2, vt, v);
        H5Aclose(a2);
        H5Tclose(vt);
        H5Sclose(a2s);
        H5Gclose(fg);
    }
    H5Fclose(file);
    free(p);
    free(g);
}

void read_pop_h5(void)
{
    need_lattice("read_pop_h5");
    const ch4g_run_params *rp = &S.rp;
    const ch4g_config *c = &rp->cfg;
    const size_t n = (size_t)c->NX * c->NY * c->NZ;
    const int fluid = (c->flags & CH4G_LB_FLUID) && rp->resume_u;
    const int thermal = (c->flags & CH4G_LB_TEMPERATURE) && rp->resume_t;
    hid_t file = H5Fopen("pop.h5", H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) {
        fprintf(stderr, "Error message -> pop.h5 cannot be opened. Exit.\n");
        exit(-1);
    }
    pop *buf[2] = {fluid ? (pop *)malloc(n * sizeof(pop)) : NULL, thermal ? (pop *)malloc(n * sizeof(pop)) : NULL};
    const char *names[2] = {"/population/velocity", "/population/temperature"};
    hid_t pt = pop_type();
    for (int k = 0; k < 2; k++) {
        if (!buf[k]) continue;
        hid_t d = H5Dopen2(file, names[k], H5P_DEFAULT);
        hsize_t dims[3] = {0, 0, 0};
        hid_t sp = d >= 0 ? H5Dget_space(d) : -1;
        if (d < 0 || sp < 0 || H5Sget_simple_extent_ndims(sp) != 3 || H5Sget_simple_extent_dims(sp, dims, NULL) != 3 ||
            dims[0] != (hsize_t)c->NZ || dims[1] != (hsize_t)c->NY || dims[2] != (hsize_t)c->NX) {
            fprintf(stderr, "Error message -> %s in pop.h5 is missing or not [NZ][NY][NX]. Exit.\n", names[k]);
            exit(-1);
        }
        if (H5Dread(d, pt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf[k]) < 0) die("read_pop_h5");
        H5Sclose(sp);
        H5Dclose(d);
    }
    H5Tclose(pt);
    if (ch4g_set_populations(S.lat, (const double *)buf[0], (const double *)buf[1])) die("read_pop_h5");
    /* /forcing (output_h5.c:653-700): HIT phases when resume_u */
    if ((c->flags & CH4G_FLUID_FORCING_HIT) && rp->resume_u && H5Lexists(file, "/forcing", H5P_DEFAULT) > 0) {
        hid_t fg = H5Gopen2(file, "/forcing", H5P_DEFAULT);
        int nk = 0;
        hid_t a1 = H5Aopen(fg, "number of modes (nk)", H5P_DEFAULT);
        H5Aread(a1, H5T_NATIVE_INT, &nk);
        H5Aclose(a1);
        if (nk != CH4G_HIT_NK) {
            fprintf(stderr, "Error message -> pop.h5 holds %d forcing modes, this lattice %d. Exit.\n", nk, CH4G_HIT_NK);
            exit(-1);
        }
        vector v[CH4G_HIT_NK];
        hid_t vt = vector_type();
        hid_t a2 = H5Aopen(fg, "phi_u", H5P_DEFAULT);
        if (a2 < 0 || H5Aread(a2, vt, v) < 0) die("read_pop_h5");
        H5Aclose(a2);
        H5Tclose(vt);
        H5Gclose(fg);
        double phi[CH4G_HIT_NK][3];
        for (int i = 0; i < CH4G_HIT_NK; i++) {
            phi[i][0] = v[i].x;
            phi[i][1] = v[i].y;
            phi[i][2] = v[i].z;
        }
        if (ch4g_set_hit_phases(S.lat, &phi[0][0])) die("read_pop_h5");
    }
    H5Fclose(file);
    free(buf[0]);
    free(buf[1]);
}
/* part.h5 (lagrange.c:4231-4546, :4548-4827): /particles/point_particle, a
 * compound of the point_particle members this lattice carries (name,
 * position, velocity, acceleration and their previous values, tau_drag, the
 * fluid velocity and its previous value, gravity_coeff with
 * LAGRANGE_GRAVITY_VARIABLE; the velocity gradient with
 * LAGRANGE_GRADIENT, then beta_coeff with added mass and the orientation
 * members with LAGRANGE_ORIENTATION; t, and with the gradient dx_t..dz_t,
 * with LB_TEMPERATURE, then temperature_particle and cp_particle with
 * LAGRANGE_TEMPERATURE) over [particle_number] records, the particles in
 * array order first (the reference sizes the dataset by particle_number and
 * writes the npart it holds); /properties carries particle_types.          */
static hid_t particle_type(const ch4g_run_params *rp)
{
    const unsigned fl = rp->pcfg.flags;
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(ch4g_particle));
#define M(f) H5Tinsert(t, #f, HOFFSET(ch4g_particle, f), H5T_NATIVE_DOUBLE)
    M(name); M(x); M(y); M(z);
    M(vx); M(vy); M(vz); M(vx_old); M(vy_old); M(vz_old);
    M(ax); M(ay); M(az); M(ax_old); M(ay_old); M(az_old);
    M(tau_drag);
    M(ux); M(uy); M(uz); M(ux_old); M(uy_old); M(uz_old);
    if (fl & CH4G_LAGRANGE_GRAVITY_VARIABLE) M(gravity_coeff); /* lagrange.c:4336-4341 */
    if (fl & CH4G_LAGRANGE_GRADIENT) {
        M(dx_ux); M(dy_ux); M(dz_ux); M(dx_uy); M(dy_uy); M(dz_uy); M(dx_uz); M(dy_uz); M(dz_uz);
        if (fl & CH4G_LAGRANGE_ADDEDMASS) M(beta_coeff);
        if (fl & CH4G_LAGRANGE_ORIENTATION) { /* lagrange.c:4365-4427 */
            M(px); M(py); M(pz); M(dt_px); M(dt_py); M(dt_pz);
            if (fl & CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION) {
                M(nx); M(ny); M(nz); M(dt_nx); M(dt_ny); M(dt_nz);
            }
            if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY) {
                M(aspect_ratio);
                if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS) M(gyrotaxis_velocity);
            }
            if (fl & CH4G_LAGRANGE_ORIENTATION_ACTIVE) M(swim_velocity);
            if (rp->pcfg.flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) { /* lagrange.c:4411-4424 */
                M(critical_shear_rate); M(shear_rate); M(jump_time); M(px_jump); M(py_jump); M(pz_jump);
            }
        }
    }
    if (rp->cfg.flags & CH4G_LB_TEMPERATURE) {
        M(t);
        if (fl & CH4G_LAGRANGE_GRADIENT) {
            M(dx_t); M(dy_t); M(dz_t);
        }
        if (fl & CH4G_LAGRANGE_TEMPERATURE) { /* lagrange.c:4442-4447 */
            H5Tinsert(t, "temperature_particle", HOFFSET(ch4g_particle, t_p), H5T_NATIVE_DOUBLE);
            H5Tinsert(t, "cp_particle", HOFFSET(ch4g_particle, cp_p), H5T_NATIVE_DOUBLE);
        }
    }
#undef M
    return t;
}

void write_point_particle_h5(void)
{
    need_particles("write_point_particle_h5");
    const ch4g_run_params *rp = &S.rp;
    int64_t n = 0;
    if (ch4g_particles_count(S.lat, &n)) die("write_point_particle_h5");
    const int64_t total = rp->pcfg.particle_number > n ? rp->pcfg.particle_number : n;
    ch4g_particle *a = (ch4g_particle *)malloc(sizeof(ch4g_particle) * (size_t)(n > 0 ? n : 1));
    if (!a) die("write_point_particle_h5: out of memory");
    if (ch4g_get_particles(S.lat, a, n, &n)) die("write_point_particle_h5");
    fprintf(stderr, "Writing file part.h5\n");
    hid_t file = H5Fcreate("part.h5", H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) die("write_point_particle_h5");
    hid_t group = H5Gcreate2(file, "/particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t dims[1] = {(hsize_t)total}, count[1] = {(hsize_t)n}, offset[1] = {0};
    hid_t fspace = H5Screate_simple(1, dims, NULL);
    hid_t mspace = H5Screate_simple(1, count, NULL);
    H5Sselect_hyperslab(fspace, H5S_SELECT_SET, offset, NULL, count, NULL);
    hid_t pt = particle_type(rp);
    hid_t d = H5Dcreate2(group, "point_particle", pt, fspace, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (d < 0 || (n > 0 && H5Dwrite(d, pt, mspace, fspace, H5P_DEFAULT, a) < 0)) die("write_point_particle_h5");
    H5Dclose(d);
    H5Tclose(pt);
    H5Sclose(mspace);
    H5Sclose(fspace);
    H5Gclose(group);
    hid_t pg = H5Gcreate2(file, "/properties", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hid_t as = H5Screate(H5S_SCALAR);
    hid_t at = H5Acreate2(pg, "particle_types", H5T_NATIVE_DOUBLE, as, H5P_DEFAULT, H5P_DEFAULT);
    const double types = ptypes(&rp->pcfg);
    H5Awrite(at, H5T_NATIVE_DOUBLE, &types);
    H5Aclose(at);
    H5Sclose(as);
    H5Gclose(pg);
    H5Fclose(file);
    free(a);
}

void read_point_particle_h5(void)
{
    need_particles("read_point_particle_h5");
    const ch4g_run_params *rp = &S.rp;
    hid_t file = H5Fopen("part.h5", H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) {
        fprintf(stderr, "Error message -> part.h5 cannot be opened. Exit.\n");
        exit(-1);
    }
    hid_t d = H5Dopen2(file, "/particles/point_particle", H5P_DEFAULT);
    hid_t sp = d >= 0 ? H5Dget_space(d) : -1;
    hsize_t dims[1] = {0};
    if (d < 0 || sp < 0 || H5Sget_simple_extent_ndims(sp) != 1 || H5Sget_simple_extent_dims(sp, dims, NULL) != 1) {
        fprintf(stderr, "Error message -> /particles/point_particle in part.h5 is missing or not 1-D. Exit.\n");
        exit(-1);
    }
    /* the reference reads particle_number records (npart of lagrange.c:128) */
    const int64_t n = rp->pcfg.particle_number < (int64_t)dims[0] ? rp->pcfg.particle_number : (int64_t)dims[0];
    ch4g_particle *a = (ch4g_particle *)calloc((size_t)(n > 0 ? n : 1), sizeof(ch4g_particle));
    if (!a) die("read_point_particle_h5: out of memory");
    hsize_t count[1] = {(hsize_t)n}, offset[1] = {0};
    hid_t mspace = H5Screate_simple(1, count, NULL);
    H5Sselect_hyperslab(sp, H5S_SELECT_SET, offset, NULL, count, NULL);
    hid_t pt = particle_type(rp);
    if (n > 0 && H5Dread(d, pt, mspace, sp, H5P_DEFAULT, a) < 0) die("read_point_particle_h5");
    H5Tclose(pt);
    H5Sclose(mspace);
    H5Sclose(sp);
    H5Dclose(d);
    H5Fclose(file);
    if (ch4g_set_particles(S.lat, a, n)) die("read_point_particle_h5");
    free(a);
}
/* Datasets of a particle snapshot, in the order of output_particles
 * (lagrange.c:1700-2160), for the switches of this build: every record
 * member is an FP64 array of particle_number entries under /lagrange.      */
struct snap_member {
    const char *name;
    size_t off;
};

static int snapshot_members(const ch4g_run_params *rp, struct snap_member *m)
{
    const unsigned fl = rp->pcfg.flags;
    const int grad = (fl & CH4G_LAGRANGE_GRADIENT) != 0, am = grad && (fl & CH4G_LAGRANGE_ADDEDMASS);
    int n = 0;
#define SM(nm, f) (m[n].name = nm, m[n].off = offsetof(ch4g_particle, f), n++)
    SM("name", name); SM("tau_drag", tau_drag);
    SM("x", x); SM("y", y); SM("z", z);
    SM("vx", vx); SM("vy", vy); SM("vz", vz);
    SM("ax", ax); SM("ay", ay); SM("az", az);
    if (fl & CH4G_LAGRANGE_GRAVITY_VARIABLE) SM("gravity_coeff", gravity_coeff); /* lagrange.c:1772-1781 */
    if (grad) {
        SM("dx_ux", dx_ux); SM("dy_ux", dy_ux); SM("dz_ux", dz_ux);
        SM("dx_uy", dx_uy); SM("dy_uy", dy_uy); SM("dz_uy", dz_uy);
        SM("dx_uz", dx_uz); SM("dy_uz", dy_uz); SM("dz_uz", dz_uz);
        if (am) SM("beta_coeff", beta_coeff);
        if (rp->lagr_output_fluid) {
            SM("ux", ux); SM("uy", uy); SM("uz", uz);
            if (am) {
                SM("Dt_ux", Dt_ux); SM("Dt_uy", Dt_uy); SM("Dt_uz", Dt_uz);
            }
        }
        if (fl & CH4G_LAGRANGE_ORIENTATION) { /* lagrange.c:1889-2012 */
            SM("px", px); SM("py", py); SM("pz", pz);
            SM("dt_px", dt_px); SM("dt_py", dt_py); SM("dt_pz", dt_pz);
            if (fl & CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION) {
                SM("nx", nx); SM("ny", ny); SM("nz", nz);
                SM("dt_nx", dt_nx); SM("dt_ny", dt_ny); SM("dt_nz", dt_nz);
            }
            if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY) {
                SM("aspect_ratio", aspect_ratio);
                if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS) SM("gyrotaxis_velocity", gyrotaxis_velocity);
            }
            if (fl & CH4G_LAGRANGE_ORIENTATION_ACTIVE) SM("swim_velocity", swim_velocity);
            if (rp->pcfg.flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) { /* lagrange.c:1993-2010 */
                SM("critical_shear_rate", critical_shear_rate); SM("shear_rate", shear_rate);
                SM("jump_time", jump_time);
            }
        }
        if (rp->pcfg.flags2 & CH4G_LAGRANGE_POLYMER) { /* lagrange.c:2015-2052 */
            SM("cxx", cxx); SM("cyy", cyy); SM("czz", czz); SM("cxy", cxy); SM("cyz", cyz); SM("cxz", cxz);
        }
    }
    if (rp->cfg.flags & CH4G_LB_TEMPERATURE) {
        SM("temperature", t); SM("dt_t", dt_t);
        if (grad) {
            SM("dx_t", dx_t); SM("dy_t", dy_t); SM("dz_t", dz_t);
        }
        if (fl & CH4G_LAGRANGE_TEMPERATURE) { /* lagrange.c:2090-2103 */
            SM("temperature_particle", t_p); SM("cp_particle", cp_p);
        }
    }
    if (rp->pcfg.flags2 & CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT) SM("sediment", sediment); /* lagrange.c:2106-2115 */
#undef SM
    return n;
}

/* XDMF items of the descriptor (lagrange.c:2176-2590): a scalar attribute,
 * or a vector / tensor joined from k datasets */
static void xmf_scalar(FILE *fo, const char *h5, const char *attr, const char *ds, int np)
{
    fprintf(fo, "<Attribute Name=\"%s\" AttributeType=\"Scalar\" Center=\"Node\"> \n", attr);
    fprintf(fo, "<DataItem Dimensions=\"%d\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n", np);
    fprintf(fo, "%s:/lagrange/%s\n</DataItem>\n</Attribute>\n", h5, ds);
}

static void xmf_joined(FILE *fo, const char *h5, const char *attr, const char *type, const char *const *ds, int k,
                       int np)
{
    fprintf(fo, "<Attribute Name=\"%s\" AttributeType=\"%s\" Center=\"Node\"> \n", attr, type);
    fprintf(fo, "<DataItem ItemType=\"Function\" Dimensions=\"%d %d\" \n   Function=\"JOIN(", np, k);
    /* the reference's spelling: "$0 , $1, $2" for vectors, "$0,$1,...,$8" for the tensor */
    for (int i = 0; i < k; i++) fprintf(fo, "%s$%d", i == 0 ? "" : (k == 3 ? (i == 1 ? " , " : ", ") : ","), i);
    fprintf(fo, ")\">\n");
    for (int i = 0; i < k; i++) {
        fprintf(fo, "<DataItem Dimensions=\"%d 1\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n", np);
        fprintf(fo, "%s:/lagrange/%s\n</DataItem>\n", h5, ds[i]);
    }
    fprintf(fo, "</DataItem>\n</Attribute>\n");
}

static void write_particle_snapshot(void)
{
    const ch4g_run_params *rp = &S.rp;
    int64_t n = 0;
    if (ch4g_particles_count(S.lat, &n)) die("output_particles");
    ch4g_particle *a = (ch4g_particle *)malloc(sizeof(ch4g_particle) * (size_t)(n > 0 ? n : 1));
    double *aux = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!a || !aux) die("output_particles: out of memory");
    if (ch4g_get_particles(S.lat, a, n, &n)) die("output_particles");
    const int np = rp->pcfg.particle_number;
    const int64_t total = np > n ? np : n;
    char h5[64], path[96], xmf[96];
    snprintf(h5, sizeof h5, "particle_%d.h5", S.itime);
    snprintf(path, sizeof path, "RUN/%s", h5);
    snprintf(xmf, sizeof xmf, "RUN/particle_%d.xmf", S.itime);
    mkdir("RUN", S_IWUSR | S_IXUSR | S_IRUSR); /* parameters.c:46-48 */
    fprintf(stderr, "Writing file %s\n", path);
    hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) die("output_particles");
    hid_t group = H5Gcreate2(file, "/lagrange", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t dims[1] = {(hsize_t)total}, count[1] = {(hsize_t)n}, offset[1] = {0};
    hid_t fspace = H5Screate_simple(1, dims, NULL);
    hid_t mspace = H5Screate_simple(1, count, NULL);
    H5Sselect_hyperslab(fspace, H5S_SELECT_SET, offset, NULL, count, NULL);
    struct snap_member m[64];
    const int nm = snapshot_members(rp, m);
    for (int k = 0; k < nm; k++) {
        for (int64_t i = 0; i < n; i++) aux[i] = *(const double *)((const char *)&a[i] + m[k].off);
        hid_t d = H5Dcreate2(group, m[k].name, H5T_NATIVE_DOUBLE, fspace, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (d < 0 || (n > 0 && H5Dwrite(d, H5T_NATIVE_DOUBLE, mspace, fspace, H5P_DEFAULT, aux) < 0))
            die("output_particles");
        H5Dclose(d);
    }
    H5Sclose(mspace);
    H5Sclose(fspace);
    H5Gclose(group);
    H5Fclose(file);
    free(aux);
    free(a);

    FILE *fo = fopen(xmf, "w");
    if (!fo) die("output_particles: xmf");
    const unsigned fl = rp->pcfg.flags;
    const int grad = (fl & CH4G_LAGRANGE_GRADIENT) != 0, am = grad && (fl & CH4G_LAGRANGE_ADDEDMASS);
    fprintf(fo, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n<Domain>\n");
    fprintf(fo, "<Grid Name=\"Points\" GridType=\"Uniform\">\n");
    fprintf(fo, "<Topology TopologyType=\"Polyvertex\" NumberOfElements=\"%d\"/>\n", np);
    fprintf(fo, "<Geometry GeometryType=\"X_Y_Z\">\n");
    const char *xyz[3] = {"x", "y", "z"};
    for (int c = 0; c < 3; c++) {
        fprintf(fo, "<DataItem Name=\"%s\" Dimensions=\"%d\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n",
                xyz[c], np);
        fprintf(fo, "%s:/lagrange/%s\n</DataItem>\n", h5, xyz[c]);
    }
    fprintf(fo, "</Geometry>\n");
    xmf_scalar(fo, h5, "name", "name", np);
    xmf_scalar(fo, h5, "tau_drag", "tau_drag", np);
    const char *v[3] = {"vx", "vy", "vz"}, *ac[3] = {"ax", "ay", "az"};
    xmf_joined(fo, h5, "velocity", "Vector", v, 3, np);
    xmf_joined(fo, h5, "acceleration", "Vector", ac, 3, np);
    if (fl & CH4G_LAGRANGE_GRAVITY_VARIABLE) xmf_scalar(fo, h5, "gravity_coeff", "gravity_coeff", np);
    if (grad) {
        const char *g[9] = {"dx_ux", "dy_ux", "dz_ux", "dx_uy", "dy_uy", "dz_uy", "dx_uz", "dy_uz", "dz_uz"};
        xmf_joined(fo, h5, "fluid velocity gradient", "Tensor", g, 9, np);
        if (am) xmf_scalar(fo, h5, "beta_coeff", "beta_coeff", np);
        if (rp->lagr_output_fluid) {
            const char *u[3] = {"ux", "uy", "uz"}, *du[3] = {"Dt_ux", "Dt_uy", "Dt_uz"};
            xmf_joined(fo, h5, "fluid_velocity", "Vector", u, 3, np);
            if (am) xmf_joined(fo, h5, "fluid_acceleration", "Vector", du, 3, np);
        }
        if (fl & CH4G_LAGRANGE_ORIENTATION) { /* lagrange.c:2344-2440 */
            const char *po[3] = {"px", "py", "pz"}, *dp[3] = {"dt_px", "dt_py", "dt_pz"};
            xmf_joined(fo, h5, "orientation", "Vector", po, 3, np);
            xmf_joined(fo, h5, "angular velocity", "Vector", dp, 3, np);
            if (fl & CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION) {
                const char *no[3] = {"nx", "ny", "nz"}, *dn[3] = {"dt_nx", "dt_ny", "dt_nz"};
                xmf_joined(fo, h5, "second orientation", "Vector", no, 3, np);
                xmf_joined(fo, h5, "second angular velocity", "Vector", dn, 3, np);
            }
            if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY) {
                xmf_scalar(fo, h5, "aspect_ratio", "aspect_ratio", np);
                if (fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS)
                    xmf_scalar(fo, h5, "gyrotaxis_velocity", "gyrotaxis_velocity", np);
            }
            if (fl & CH4G_LAGRANGE_ORIENTATION_ACTIVE) xmf_scalar(fo, h5, "swim_velocity", "swim_velocity", np);
        }
        if (rp->pcfg.flags2 & CH4G_LAGRANGE_POLYMER) { /* lagrange.c:2457-2482 */
            const char *c6[6] = {"cxx", "cxy", "cxz", "cyy", "cyz", "czz"};
            xmf_joined(fo, h5, "polymer conformation tensor", "Tensor6", c6, 6, np);
        }
    }
    if (rp->cfg.flags & CH4G_LB_TEMPERATURE) {
        xmf_scalar(fo, h5, "temperature", "temperature", np);
        xmf_scalar(fo, h5, "dt_t", "dt_t", np);
        if (grad) {
            const char *gt[3] = {"dx_t", "dy_t", "dz_t"};
            xmf_joined(fo, h5, "temperature gradient", "Vector", gt, 3, np);
        }
        if (fl & CH4G_LAGRANGE_TEMPERATURE) { /* lagrange.c:2515-2528 */
            xmf_scalar(fo, h5, "temperature_particle", "temperature_particle", np);
            xmf_scalar(fo, h5, "cp_particle", "cp_particle", np);
        }
    }
    if (rp->pcfg.flags2 & CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT) /* lagrange.c:2531-2540 */
        xmf_scalar(fo, h5, "sediment", "sediment", np);
    fprintf(fo, "<Time Value=\" %e\" />\n</Grid>\n</Domain>\n</Xdmf>\n", rp->time_start + S.itime * rp->time_dt);
    fclose(fo);
}
#else
static void write_particle_snapshot(void)
{
    fprintf(stderr, "Error message -> this drop-in layer was built without HDF5 (particle snapshots). Exit.\n");
    exit(-1);
}

void write_point_particle_h5(void)
{
    fprintf(stderr, "Error message -> this drop-in layer was built without HDF5 (part.h5). Exit.\n");
    exit(-1);
}

void read_point_particle_h5(void)
{
    fprintf(stderr, "Error message -> this drop-in layer was built without HDF5 (part.h5). Exit.\n");
    exit(-1);
}

void write_pop_h5(void)
{
    fprintf(stderr, "Error message -> this drop-in layer was built without HDF5 (pop.h5). Exit.\n");
    exit(-1);
}

void read_pop_h5(void)
{
    fprintf(stderr, "Error message -> this drop-in layer was built without HDF5 (pop.h5). Exit.\n");
    exit(-1);
}
#endif
