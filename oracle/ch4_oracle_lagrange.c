/*
 * ch4_oracle_lagrange.c — FP64 CPU restatement of the reference's Lagrangian
 * point particles.  TEST INFRASTRUCTURE ONLY (the checker of the HIP particle
 * kernels); nothing in the product links, loads or calls it.
 *
 * Restated with the reference's structure: particle records in one array,
 * fields with a ghost halo filled by boundary_conditions_hydro, cell lookup
 * by a linear scan of the cell centres, trilinear sums and finite
 * differences in the reference's evaluation order, sequential feedback
 * accumulation in particle order.  The halo is two nodes wide
 * (LAGRANGE_GRADIENT_BRD2, define.h:334, typedef.h:315-316): with one node
 * and no fluid walls the reference reads unset gradient components at the
 * ghost corners (gradients.c:83-95 has no branch for them), which has no
 * defined value to follow.
 *
 * Sources followed:
 *   initial_conditions_particles  lagrange.c:72-1021 (no restart branch; the
 *                                 fluid-velocity, Stokes and fluid-temperature
 *                                 starts :968-991, :1004-1015)
 *   boundary_conditions_hydro     boundary_conditions.c:2786-2997
 *   interpolate_*_at_particles    lagrange.c:1211-1593
 *   gradient_vector (LAGRANGE)    gradients.c:78-499
 *   move_particles                lagrange.c:2934-4025 (thermal relaxation of
 *                                 t_p :4007-4020, wall temperatures of a
 *                                 re-injected particle :3839-3842, :3915-3918;
 *                                 spheroid drag :3267-3310 with glibc log /
 *                                 acos / sqrt, which the device library's host
 *                                 code evaluates on the same operands)
 *   sendrecv_particles (1 rank)   lagrange.c:4028-4221
 *   add_particle_feedbacks        lagrange_feedback.c:11-256, 525-836
 * Pins: tests/test_oracle_particles.py.  Parity unpinned: none of these
 * functions calls libm except pow(r^2, 3/2) of the particle volume and
 * exp(-fac1) of the particle temperature, which the device library's host
 * code computes with the same glibc calls on the same operands, and
 * pow(Re_p, 0.687) of the wake drag, which the device evaluates with its own
 * pow (agreement to its last bits, not bit for bit).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ch4_oracle_internal.h"

#define B2 2 /* halo width of the particle-side fields */

typedef struct olagr {
    ch4g_particle_config pc;
    int npart;
    ch4g_particle *tr;
    int GX, GY, GZ;
    size_t ng;
    ovec *u, *ftw;     /* velocity with halo, force_twoway */
    double *t;         /* temperature with halo */
    double *cx, *cy, *cz; /* cell centres incl. ghosts (grid.c:183-185,498) */
    double dt;         /* time_dt_lagr = time_dt / time_dt_lagr_factor (parameters.c:164) */
} olagr;

#define LHAS(L, f) (((L)->pc.flags & (f)) != 0)
#define IDX2(L, i, j, k) ((((size_t)(k)) * (size_t)(L)->GY + (size_t)(j)) * (size_t)(L)->GX + (size_t)(i))
#define WRAP(x, s) ((x) - floor((x) / (s)) * (s)) /* typedef.h:378 */

int olagr_twoway(const oracle *o) { return o->lagr && LHAS(o->lagr, CH4G_LAGRANGE_TWOWAY_MOMENTUM); }

void olagr_free(oracle *o)
{
    olagr *L = o->lagr;
    if (!L) return;
    free(L->tr);
    free(L->u);
    free(L->ftw);
    free(L->t);
    free(L->cx);
    free(L->cy);
    free(L->cz);
    free(L);
    o->lagr = NULL;
}

/* One property ladder of lagrange.c:264-298 (and its copies for beta):
 * values vary with period `rep` over the particle types that follow the
 * fluid tracers.                                                            */
static void ladder(double *tab, int ntypes, int ntr, double trval, double types, double vmin, double vmax, int logs,
                   double repetitions)
{
    double cycles = (ntypes - ntr) / types;
    double step;
    if (logs)
        step = types > 1 ? pow(vmax / vmin, 1.0 / (types - 1.0)) : 0;
    else
        step = types > 1 ? (vmax - vmin) / (types - 1.0) : 0;
    for (int n = 0; n < ntr; n++) tab[n] = trval;
    for (int k = 0; k < (int)(cycles / repetitions); k++)
        for (int i = 0; i < (int)types; i++)
            for (int j = 0; j < (int)repetitions; j++) {
                int at = ntr + j + i * (int)repetitions + k * (int)repetitions * (int)types;
                tab[at] = logs ? vmin * pow(step, (double)i) : vmin + i * step;
            }
}

static void bc_hydro(oracle *o);
static void interpolate_fields(oracle *o, int fluid, int thermal);

int oracle_particles_init(oracle *o, const ch4g_particle_config *pc)
{
    olagr_free(o);
    olagr *L = (olagr *)calloc(1, sizeof(olagr));
    if (!L) return -1;
    o->lagr = L;
    L->pc = *pc;
    L->GX = o->NX + 2 * B2;
    L->GY = o->NY + 2 * B2;
    L->GZ = o->NZ + 2 * B2;
    L->ng = (size_t)L->GX * L->GY * L->GZ;
    L->u = (ovec *)calloc(L->ng, sizeof(ovec));
    L->ftw = (ovec *)calloc(L->ng, sizeof(ovec));
    L->t = (double *)calloc(L->ng, sizeof(double));
    L->cx = (double *)malloc(sizeof(double) * L->GX);
    L->cy = (double *)malloc(sizeof(double) * L->GY);
    L->cz = (double *)malloc(sizeof(double) * L->GZ);
    if (!L->u || !L->ftw || !L->t || !L->cx || !L->cy || !L->cz) return -1;
    for (int i = 0; i < L->GX; i++) L->cx[i] = (double)(i - B2) + 0.5;
    for (int j = 0; j < L->GY; j++) L->cy[j] = (double)(j - B2) + 0.5;
    for (int k = 0; k < L->GZ; k++) L->cz[k] = (double)(k - B2) + 0.5;
    L->dt = 1.0 / pc->time_dt_lagr_factor;

    /* particle_types (parameters.c:559-584, :621-695, :743-755) */
    const int am = LHAS(L, CH4G_LAGRANGE_GRADIENT) && LHAS(L, CH4G_LAGRANGE_ADDEDMASS);
    const int orient = LHAS(L, CH4G_LAGRANGE_ORIENTATION);
    const int jeff = orient && LHAS(L, CH4G_LAGRANGE_ORIENTATION_JEFFREY);
    const int gyro = jeff && LHAS(L, CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS);
    const int swim = orient && LHAS(L, CH4G_LAGRANGE_ORIENTATION_ACTIVE);
    const int tpart = LHAS(L, CH4G_LAGRANGE_TEMPERATURE);
    const int gvar = LHAS(L, CH4G_LAGRANGE_GRAVITY) && LHAS(L, CH4G_LAGRANGE_GRAVITY_VARIABLE);
    const int rad = (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY) != 0;
    double types = pc->tau_drag_types;
    if (am) types *= pc->beta_coeff_types;
    if (rad) types = pc->particle_radius_types * pc->particle_density_types; /* parameters.c:530-556 */
    if (gvar) types *= pc->gravity_coeff_types; /* parameters.c:600 */
    if (tpart) types *= pc->cp_coeff_types; /* parameters.c:615 */
    if (jeff) types *= pc->aspect_ratio_types;
    if (gyro) types *= pc->gyrotaxis_velocity_types;
    if (swim) types *= pc->swim_velocity_types;
    if (swim && (pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP))
        types *= pc->critical_shear_rate_types * pc->jump_time_types; /* parameters.c:709, :720 */
    /* fluid_tracers: 2 with the Eulerian probe type (parameters.c:744-755) */
    int ntr = LHAS(L, CH4G_LAGRANGE_ADD_TRACER) ? (LHAS(L, CH4G_LAGRANGE_ADD_TRACER_STATIC) ? 2 : 1) : 0;
    types += ntr;
    const int nt = (int)types;
    double *tau = (double *)calloc(nt, sizeof(double)), *beta = (double *)calloc(nt, sizeof(double));
    double *asp = (double *)calloc(nt, sizeof(double)), *gyv = (double *)calloc(nt, sizeof(double));
    double *swv = (double *)calloc(nt, sizeof(double)), *cpv = (double *)calloc(nt, sizeof(double));
    double *gcv = (double *)calloc(nt, sizeof(double));
    if (!tau || !beta || !asp || !gyv || !swv || !cpv || !gcv) return -1;
    double rep = 1.0;
    if (rad) { /* lagrange.c:150-260 */
        double *radius = (double *)calloc(nt, sizeof(double)), *density = (double *)calloc(nt, sizeof(double));
        if (!radius || !density) return -1;
        ladder(radius, nt, ntr, 0.0, pc->particle_radius_types, pc->particle_radius_min, pc->particle_radius_max,
               (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS) != 0, rep);
        rep *= pc->particle_radius_types;
        ladder(density, nt, ntr, 1.0, pc->particle_density_types, pc->particle_density_min, pc->particle_density_max,
               (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY) != 0, rep);
        rep *= pc->particle_density_types;
        /* tau = (2/9) (rho_p / rho_f) r^2 / nu with rho_f = 1 */
        for (int i = 0; i < nt; i++) tau[i] = (2. / 9.) * density[i] * pow(radius[i], 2.0) / (o->nu);
        if (am)
            for (int i = 0; i < nt; i++) {
                /* beta from the density; tau = r^2 / (3 beta nu) with the added mass */
                beta[i] = 3.0 / (1.0 + 2.0 * density[i]);
                tau[i] = pow(radius[i], 2.0) / (3.0 * beta[i] * o->nu);
            }
        free(radius);
        free(density);
    } else {
        ladder(tau, nt, ntr, 0.0, pc->tau_drag_types, pc->tau_drag_min, pc->tau_drag_max,
               LHAS(L, CH4G_LAGRANGE_TAUDRAG_INCREMENT_LOG), rep);
        if (LHAS(L, CH4G_LAGRANGE_ADD_TRACER_STATIC)) tau[0] = -1.0; /* the probe type, lagrange.c:280-282 */
        rep *= pc->tau_drag_types;
        if (am) {
            ladder(beta, nt, ntr, 1.0, pc->beta_coeff_types, pc->beta_coeff_min, pc->beta_coeff_max,
                   LHAS(L, CH4G_LAGRANGE_ADDEDMASS_INCREMENT_LOG), rep);
            rep *= pc->beta_coeff_types;
        }
    }
    /* lagrange.c:347-384: gravity_coeff (linear or geometric), 0 for the tracer */
    if (gvar) {
        ladder(gcv, nt, ntr, 0.0, pc->gravity_coeff_types, pc->gravity_coeff_min, pc->gravity_coeff_max,
               LHAS(L, CH4G_LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG), rep);
        rep *= pc->gravity_coeff_types;
    }
    /* lagrange.c:389-423 (linear or geometric), :445-470 (geometric), :502-521 (linear) */
    if (jeff) {
        ladder(asp, nt, ntr, 1.0, pc->aspect_ratio_types, pc->aspect_ratio_min, pc->aspect_ratio_max,
               LHAS(L, CH4G_LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG), rep);
        rep *= pc->aspect_ratio_types;
    }
    if (gyro) {
        ladder(gyv, nt, ntr, 0.0, pc->gyrotaxis_velocity_types, pc->gyrotaxis_velocity_min,
               pc->gyrotaxis_velocity_max, 1, rep);
        rep *= pc->gyrotaxis_velocity_types;
    }
    if (swim) {
        ladder(swv, nt, ntr, 0.0, pc->swim_velocity_types, pc->swim_velocity_min, pc->swim_velocity_max, 0, rep);
        rep *= pc->swim_velocity_types;
    }
    /* lagrange.c:526-570: critical_shear_rate (tracer 1e10) and jump_time (tracer 0), linear */
    const int jump = swim && (pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP);
    double *csr = (double *)calloc(nt, sizeof(double)), *jtv = (double *)calloc(nt, sizeof(double));
    if (!csr || !jtv) return -1;
    if (jump) {
        ladder(csr, nt, ntr, 1.e+10, pc->critical_shear_rate_types, pc->critical_shear_rate_min,
               pc->critical_shear_rate_max, 0, rep);
        rep *= pc->critical_shear_rate_types;
        ladder(jtv, nt, ntr, 0.0, pc->jump_time_types, pc->jump_time_min, pc->jump_time_max, 0, rep);
        rep *= pc->jump_time_types;
    }
    /* lagrange.c:579-617: cp_coeff after the orientation ladders, 0 for the tracer */
    if (tpart) {
        ladder(cpv, nt, ntr, 0.0, pc->cp_coeff_types, pc->cp_coeff_min, pc->cp_coeff_max,
               LHAS(L, CH4G_LAGRANGE_TEMPERATURE_INCREMENT_LOG), rep);
        rep *= pc->cp_coeff_types;
    }

    L->npart = pc->particle_number;
    L->tr = (ch4g_particle *)calloc(L->npart > 0 ? L->npart : 1, sizeof(ch4g_particle));
    if (!L->tr) return -1;
    for (int i = 0; i < L->npart; i++) L->tr[i].name = i; /* lagrange.c:634-641, one rank */
    const double LNX = o->NX, LNY = o->NY, LNZ = o->NZ;
    int nxpart = 1, nypart = 1, nzpart = 1;
    double dpart = 1.0;
    for (int i = 0; i < L->npart; i++) {
        ch4g_particle *p = &L->tr[i];
        int type = ((int)p->name) % nt;
        p->tau_drag = tau[type];
        if (am) p->beta_coeff = beta[type];
        if (gvar) p->gravity_coeff = gcv[type]; /* lagrange.c:719-722 */
        /* lagrange.c:767-769: three draws per particle, particle order */
        p->x = 0 + oracle_drand48(o) * LNX;
        p->y = 0 + oracle_drand48(o) * LNY;
        p->z = 0 + oracle_drand48(o) * LNZ;
        if (pc->flags2 & CH4G_LAGRANGE_INITIAL_GRID) { /* lagrange.c:774-795 */
            if (i == 0) {
                int cells = o->NX * o->NY * o->NZ;
                dpart = pow(cells / L->npart, 1. / 3.);
                nxpart = (int)floor(o->NX / dpart);
                nypart = (int)floor(o->NY / dpart);
                nzpart = (int)floor(o->NZ / dpart);
                if (nxpart == 0) nxpart = 1;
                if (nypart == 0) nypart = 1;
                if (nzpart == 0) nzpart = 1;
            }
            int ipart = i % nxpart;
            int jpart = (int)ceil(i / nxpart) % nypart;
            int kpart = (int)ceil(i / (nxpart * nypart)) % nzpart;
            p->x = 0 + dpart / 2. + ipart * dpart;
            p->y = 0 + dpart / 2. + jpart * dpart;
            p->z = 0 + dpart / 2. + kpart * dpart;
        }
        if ((pc->flags2 & CH4G_LAGRANGE_INITIAL_PAIRS) && i % 2 == 0 && i > 0) { /* lagrange.c:798-814 */
            double theta = acos(2.0 * oracle_drand48(o) - 1.0);
            double phi = 2 * 3.14159265359 * oracle_drand48(o);
            double dx = sin(theta) * sin(phi), dy = sin(theta) * cos(phi), dz = cos(theta);
            double pair_distance = 1.0;
            ch4g_particle *q = &L->tr[i - 1];
            q->x = p->x + dx * pair_distance;
            q->y = p->y + dy * pair_distance;
            q->z = p->z + dz * pair_distance;
            if (q->x < 0 || q->x >= 0 + LNX) q->x = p->x - dx * pair_distance;
            if (q->y < 0 || q->y >= 0 + LNY) q->y = p->y - dy * pair_distance;
            if (q->z < 0 || q->z >= 0 + LNZ) q->z = p->z - dz * pair_distance;
        }
        p->vx = p->vy = p->vz = 0.0;
        p->t = p->t_old = p->dt_t = 0.0;
        if (jeff) p->aspect_ratio = asp[type];
        if (gyro) p->gyrotaxis_velocity = gyv[type];
        if (swim) p->swim_velocity = swv[type];
        if (jump) {
            p->critical_shear_rate = csr[type];
            p->jump_time = jtv[type];
        }
        if (tpart) {
            p->cp_p = cpv[type];     /* lagrange.c:730-732 */
            p->t_p = o->cfg.T_bot;   /* lagrange.c:862-864 */
        }
        if (orient) { /* random_vector (random.c:139-150), lagrange.c:891-901 */
            double theta = acos(2.0 * oracle_drand48(o) - 1.0);
            double phi = 2 * 3.14159265359 * oracle_drand48(o);
            p->px = sin(theta) * sin(phi);
            p->py = sin(theta) * cos(phi);
            p->pz = cos(theta);
            p->dt_px = p->dt_py = p->dt_pz = 0.0;
            if (LHAS(L, CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION)) {
                /* n = p x v / |p x v|, v a second random_vector (lagrange.c:904-917) */
                double t2 = acos(2.0 * oracle_drand48(o) - 1.0);
                double f2 = 2 * 3.14159265359 * oracle_drand48(o);
                double vx = sin(t2) * sin(f2), vy = sin(t2) * cos(f2), vz = cos(t2);
                p->nx = p->py * vz - p->pz * vy;
                p->ny = -p->px * vz + p->pz * vx;
                p->nz = p->px * vy - p->py * vx;
                double val = sqrt(p->nx * p->nx + p->ny * p->ny + p->nz * p->nz);
                p->nx /= val;
                p->ny /= val;
                p->nz /= val;
            }
        }
        if (jump) { /* lagrange.c:927-934 */
            p->time_from_jump = 10. * oracle_drand48(o) * p->jump_time;
            p->px_jump = p->px;
            p->py_jump = p->py;
            p->pz_jump = p->pz;
        }
        if (pc->flags2 & CH4G_LAGRANGE_POLYMER) { /* lagrange.c:937-955 */
            double theta = acos(2.0 * oracle_drand48(o) - 1.0);
            double phi = 2 * 3.14159265359 * oracle_drand48(o);
            p->cxx = sin(theta) * sin(phi);
            p->cyy = sin(theta) * cos(phi);
            p->czz = cos(theta);
        }
    }
    free(tau);
    free(beta);
    free(asp);
    free(gyv);
    free(swv);
    free(cpv);
    free(gcv);
    free(csr);
    free(jtv);
    /* lagrange.c:968-980: boundary_conditions_hydro, then
     * interpolate_vector_at_particles(u) (u and, with the gradient, grad u)
     * and v = u */
    if (LHAS(L, CH4G_LAGRANGE_INITIAL_VELOCITY_FLUID)) {
        bc_hydro(o);
        interpolate_fields(o, 1, 0);
        for (int i = 0; i < L->npart; i++) {
            ch4g_particle *p = &L->tr[i];
            p->vx = p->ux;
            p->vy = p->uy;
            p->vz = p->uz;
        }
    }
    /* lagrange.c:981-991: the settling velocity -g (1 - beta) tau */
    if (LHAS(L, CH4G_LAGRANGE_INITIAL_VELOCITY_STOKES))
        for (int i = 0; i < L->npart; i++) {
            ch4g_particle *p = &L->tr[i];
            p->vx = -o->cfg.gravity_x * (1.0 - p->beta_coeff) * p->tau_drag;
            p->vy = -o->cfg.gravity_y * (1.0 - p->beta_coeff) * p->tau_drag;
            p->vz = -o->cfg.gravity_z * (1.0 - p->beta_coeff) * p->tau_drag;
        }
    /* lagrange.c:1004-1015: interpolate_scalar_at_particles(t) (t and, with
     * the gradient, dx_t..dz_t) and t_p = t.  At that point of main.c
     * (:31-51) hydro_fields has written the interior only; the halo is
     * filled here by boundary_conditions_hydro's rules, as for every
     * sub-step, so the values of particles within one node of a face follow
     * those rules rather than memory the reference has not set.             */
    if (tpart && LHAS(L, CH4G_LAGRANGE_INITIAL_TEMPERATURE_FLUID)) {
        bc_hydro(o);
        interpolate_fields(o, 0, 1);
        for (int i = 0; i < L->npart; i++) L->tr[i].t_p = L->tr[i].t;
    }
    return 0;
}

/* boundary_conditions_hydro (boundary_conditions.c:2786-2997) on one rank:
 * sendrecv_borders_vector/_scalar fill every ghost node with its periodic
 * image, then the no-slip rule u_g = -u and the fixed-temperature rule
 * t_g = (2 (T_wall - T_ref) / t - 1) t set the first ghost plane of each
 * y wall, over all i, k including ghosts.                                  */
static void bc_hydro(oracle *o)
{
    olagr *L = o->lagr;
    const int NX = o->NX, NY = o->NY, NZ = o->NZ;
    for (int k = 0; k < L->GZ; k++)
        for (int j = 0; j < L->GY; j++)
            for (int i = 0; i < L->GX; i++) {
                int si = ((i - B2) % NX + NX) % NX, sj = ((j - B2) % NY + NY) % NY, sk = ((k - B2) % NZ + NZ) % NZ;
                size_t m = IDX(o, si + BRD, sj + BRD, sk + BRD), n = IDX2(L, i, j, k);
                if (HAS(o, CH4G_LB_FLUID)) L->u[n] = o->u[m];
                if (HAS(o, CH4G_LB_TEMPERATURE)) L->t[n] = o->t[m];
            }
    if (HAS(o, CH4G_LB_FLUID) && HAS(o, CH4G_FLUID_BC_Y))
        for (int i = 0; i < L->GX; i++)
            for (int k = 0; k < L->GZ; k++) {
                const double cm = -1.0, cp = -1.0;
                int j = B2;
                ovec a = L->u[IDX2(L, i, j, k)];
                L->u[IDX2(L, i, j - 1, k)] = (ovec){cm * a.x, cm * a.y, cm * a.z};
                j = NY + B2 - 1;
                a = L->u[IDX2(L, i, j, k)];
                L->u[IDX2(L, i, j + 1, k)] = (ovec){cp * a.x, cp * a.y, cp * a.z};
            }
    if (HAS(o, CH4G_LB_TEMPERATURE) && HAS(o, CH4G_TEMPERATURE_BC_Y))
        for (int i = 0; i < L->GX; i++)
            for (int k = 0; k < L->GZ; k++) {
                int j = B2;
                double T_wall = o->cfg.T_bot;
                double fac = 2.0 * (T_wall - o->cfg.T_ref) / L->t[IDX2(L, i, j, k)] - 1.0;
                L->t[IDX2(L, i, j - 1, k)] = fac * L->t[IDX2(L, i, j, k)];
                j = NY + B2 - 1;
                T_wall = o->cfg.T_top;
                fac = 2.0 * (T_wall - o->cfg.T_ref) / L->t[IDX2(L, i, j, k)] - 1.0;
                L->t[IDX2(L, i, j + 1, k)] = fac * L->t[IDX2(L, i, j, k)];
            }
    (void)NZ;
}

/* linear scan of lagrange.c:1253-1264 (last match wins) */
static int locate(const double *c, int n, double x)
{
    int im = -1;
    for (int i = 0; i < n - 1; i++)
        if (c[i] <= x && x < c[i + 1]) im = i;
    return im;
}

typedef struct {
    int im, jm, km;
    double w[8];
    size_t n[8];
} ostencil;

/* cell and volume coefficients of lagrange.c:1232-1292, corners in the
 * order of the reference's sums                                            */
static int stencil(const oracle *o, const ch4g_particle *p, ostencil *s)
{
    const olagr *L = o->lagr;
    double px = WRAP(p->x, (double)o->NX), py = WRAP(p->y, (double)o->NY), pz = WRAP(p->z, (double)o->NZ);
    int im = locate(L->cx, L->GX, px), jm = locate(L->cy, L->GY, py), km = locate(L->cz, L->GZ, pz);
    if (im < 0 || jm < 0 || km < 0) return -1;
    int ip = im + 1, jp = jm + 1, kp = km + 1;
    double dxm = px - L->cx[im], dxp = L->cx[ip] - px;
    double dym = py - L->cy[jm], dyp = L->cy[jp] - py;
    double dzm = pz - L->cz[km], dzp = L->cz[kp] - pz;
    s->im = im;
    s->jm = jm;
    s->km = km;
    s->w[0] = dxp * dyp * dzp; /* (im,jm,km) */
    s->w[1] = dxm * dyp * dzp; /* (ip,jm,km) */
    s->w[2] = dxp * dym * dzp; /* (im,jp,km) */
    s->w[3] = dxp * dyp * dzm; /* (im,jm,kp) */
    s->w[4] = dxm * dym * dzp; /* (ip,jp,km) */
    s->w[5] = dxp * dym * dzm; /* (im,jp,kp) */
    s->w[6] = dxm * dyp * dzm; /* (ip,jm,kp) */
    s->w[7] = dxm * dym * dzm; /* (ip,jp,kp) */
    s->n[0] = IDX2(L, im, jm, km);
    s->n[1] = IDX2(L, ip, jm, km);
    s->n[2] = IDX2(L, im, jp, km);
    s->n[3] = IDX2(L, im, jm, kp);
    s->n[4] = IDX2(L, ip, jp, km);
    s->n[5] = IDX2(L, im, jp, kp);
    s->n[6] = IDX2(L, ip, jm, kp);
    s->n[7] = IDX2(L, ip, jp, kp);
    return 0;
}

static const int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};

/* gradient_vector with LAGRANGE (gradients.c:78-499): centred differences
 * where both neighbours exist; with fluid walls (LB_FLUID_BC) one-sided
 * second order at the first and last fluid rows of y (:201-225) and the
 * LAGRANGE_INTERPOLATION_FIX at the ghost cells next to the interior in
 * every direction (:288-368), which moves the evaluation point one cell
 * inwards and keeps the moved index for the directions that follow; without
 * fluid walls the fix evaluates the one-sided difference at the ghost cell
 * itself (:413-488); then components along a size-one axis are zeroed
 * (:494-496).                                                              */
static oten grad_vec(const oracle *o, int i, int j, int k)
{
    const olagr *L = o->lagr;
    const ovec *t = L->u;
    const int NX = o->NX, NY = o->NY, NZ = o->NZ;
    oten g;
    memset(&g, 0, sizeof g);
    double h1, h2, a0, a1, a2;
    if (i > 0 && i < L->GX - 1) {
        size_t p = IDX2(L, i + 1, j, k), m = IDX2(L, i - 1, j, k);
        double d = L->cx[i + 1] - L->cx[i - 1];
        g.xx = (t[p].x - t[m].x) / d;
        g.yx = (t[p].y - t[m].y) / d;
        g.zx = (t[p].z - t[m].z) / d;
    }
    if (j > 0 && j < L->GY - 1) {
        size_t p = IDX2(L, i, j + 1, k), m = IDX2(L, i, j - 1, k);
        double d = L->cy[j + 1] - L->cy[j - 1];
        g.xy = (t[p].x - t[m].x) / d;
        g.yy = (t[p].y - t[m].y) / d;
        g.zy = (t[p].z - t[m].z) / d;
    }
    if (k > 0 && k < L->GZ - 1) {
        size_t p = IDX2(L, i, j, k + 1), m = IDX2(L, i, j, k - 1);
        double d = L->cz[k + 1] - L->cz[k - 1];
        g.xz = (t[p].x - t[m].x) / d;
        g.yz = (t[p].y - t[m].y) / d;
        g.zz = (t[p].z - t[m].z) / d;
    }
#define FWD(c, a)                                                              \
    h1 = c[a + 1] - c[a];                                                      \
    h2 = c[a + 2] - c[a + 1];                                                  \
    a0 = -(2.0 * h1 + h2) / (h1 * (h1 + h2));                                  \
    a1 = (h1 + h2) / (h1 * h2);                                                \
    a2 = -(h1) / (h2 * (h1 + h2));
#define BWD(c, a)                                                              \
    h1 = c[a] - c[a - 1];                                                      \
    h2 = c[a - 1] - c[a - 2];                                                  \
    a0 = (2.0 * h1 + h2) / (h1 * (h1 + h2));                                   \
    a1 = -(h1 + h2) / (h1 * h2);                                               \
    a2 = (h1) / (h2 * (h1 + h2));
#define COMB(n0, n1, n2, X, Y, Z)                                              \
    X = a0 * t[n0].x + a1 * t[n1].x + a2 * t[n2].x;                            \
    Y = a0 * t[n0].y + a1 * t[n1].y + a2 * t[n2].y;                            \
    Z = a0 * t[n0].z + a1 * t[n1].z + a2 * t[n2].z;
    if (HAS(o, CH4G_FLUID_BC_Y)) {
        /* LB_FLUID_BC_Y, SECOND_ORDER_GRAD */
        if (j == B2) {
            FWD(L->cy, j)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j + 1, k), IDX2(L, i, j + 2, k), g.xy, g.yy, g.zy)
        }
        if (j == NY + B2 - 1) {
            BWD(L->cy, j)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j - 1, k), IDX2(L, i, j - 2, k), g.xy, g.yy, g.zy)
        }
        /* LAGRANGE_INTERPOLATION_FIX, SECOND_ORDER_GRAD */
        if (i == B2 - 1) {
            i = i + 1;
            FWD(L->cx, i)
            COMB(IDX2(L, i, j, k), IDX2(L, i + 1, j, k), IDX2(L, i + 2, j, k), g.xx, g.yx, g.zx)
        }
        if (i == NX + B2) {
            i = i - 1;
            BWD(L->cx, i)
            COMB(IDX2(L, i, j, k), IDX2(L, i - 1, j, k), IDX2(L, i - 2, j, k), g.xx, g.yx, g.zx)
        }
        if (j == B2 - 1) {
            j = j + 1;
            FWD(L->cy, j)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j + 1, k), IDX2(L, i, j + 2, k), g.xy, g.yy, g.zy)
        }
        if (j == NY + B2) {
            j = j - 1;
            BWD(L->cy, j)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j - 1, k), IDX2(L, i, j - 2, k), g.xy, g.yy, g.zy)
        }
        if (k == B2 - 1) {
            k = k + 1;
            FWD(L->cz, k)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j, k + 1), IDX2(L, i, j, k + 2), g.xz, g.yz, g.zz)
        }
        if (k == NZ + B2) {
            k = k - 1;
            BWD(L->cz, k)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j, k - 1), IDX2(L, i, j, k - 2), g.xz, g.yz, g.zz)
        }
    } else {
        /* LAGRANGE_INTERPOLATION_FIX, SECOND_ORDER_GRAD, no fluid walls */
        if (i == B2 - 1) {
            FWD(L->cx, i)
            COMB(IDX2(L, i, j, k), IDX2(L, i + 1, j, k), IDX2(L, i + 2, j, k), g.xx, g.yx, g.zx)
        }
        if (i == NX + B2) {
            BWD(L->cx, i)
            COMB(IDX2(L, i, j, k), IDX2(L, i - 1, j, k), IDX2(L, i - 2, j, k), g.xx, g.yx, g.zx)
        }
        if (j == B2 - 1) {
            FWD(L->cy, j)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j + 1, k), IDX2(L, i, j + 2, k), g.xy, g.yy, g.zy)
        }
        if (j == NY + B2) {
            BWD(L->cy, j)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j - 1, k), IDX2(L, i, j - 2, k), g.xy, g.yy, g.zy)
        }
        if (k == B2 - 1) {
            FWD(L->cz, k)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j, k + 1), IDX2(L, i, j, k + 2), g.xz, g.yz, g.zz)
        }
        if (k == NZ + B2) {
            BWD(L->cz, k)
            COMB(IDX2(L, i, j, k), IDX2(L, i, j, k - 1), IDX2(L, i, j, k - 2), g.xz, g.yz, g.zz)
        }
    }
#undef COMB
    if (NX == 1) g.xx = g.yx = g.zx = 0.0;
    if (NY == 1) g.xy = g.yy = g.zy = 0.0;
    if (NZ == 1) g.xz = g.yz = g.zz = 0.0;
    return g;
}

/* gradient_scalar with LAGRANGE (gradients.c:505-827) of the temperature
 * with its halo: centred differences where both neighbours exist; with
 * temperature walls (LB_TEMPERATURE_BC_Y) one-sided second order at the
 * first and last rows of y (:560-626) and the LAGRANGE_INTERPOLATION_FIX
 * that moves the evaluation point one cell inwards (:657-725); without them
 * the fix evaluates the one-sided difference at the ghost cell itself
 * (:762-825); components along a size-one axis are zeroed (:829-831).     */
static ovec grad_sca(const oracle *o, int i, int j, int k)
{
    const olagr *L = o->lagr;
    const double *t = L->t;
    const int NX = o->NX, NY = o->NY, NZ = o->NZ;
    ovec g;
    memset(&g, 0, sizeof g);
    double h1, h2, a0, a1, a2;
    if (i > 0 && i < L->GX - 1) g.x = (t[IDX2(L, i + 1, j, k)] - t[IDX2(L, i - 1, j, k)]) / (L->cx[i + 1] - L->cx[i - 1]);
    if (j > 0 && j < L->GY - 1) g.y = (t[IDX2(L, i, j + 1, k)] - t[IDX2(L, i, j - 1, k)]) / (L->cy[j + 1] - L->cy[j - 1]);
    if (k > 0 && k < L->GZ - 1) g.z = (t[IDX2(L, i, j, k + 1)] - t[IDX2(L, i, j, k - 1)]) / (L->cz[k + 1] - L->cz[k - 1]);
#define S3(n0, n1, n2) (a0 * t[n0] + a1 * t[n1] + a2 * t[n2])
    if (HAS(o, CH4G_TEMPERATURE_BC_Y)) {
        if (j == B2) {
            FWD(L->cy, j)
            g.y = S3(IDX2(L, i, j, k), IDX2(L, i, j + 1, k), IDX2(L, i, j + 2, k));
        }
        if (j == NY + B2 - 1) {
            BWD(L->cy, j)
            g.y = S3(IDX2(L, i, j, k), IDX2(L, i, j - 1, k), IDX2(L, i, j - 2, k));
        }
        if (i == B2 - 1) {
            i = i + 1;
            FWD(L->cx, i)
            g.x = S3(IDX2(L, i, j, k), IDX2(L, i + 1, j, k), IDX2(L, i + 2, j, k));
        }
        if (i == NX + B2) {
            i = i - 1;
            BWD(L->cx, i)
            g.x = S3(IDX2(L, i, j, k), IDX2(L, i - 1, j, k), IDX2(L, i - 2, j, k));
        }
        if (j == B2 - 1) {
            j = j + 1;
            FWD(L->cy, j)
            g.y = S3(IDX2(L, i, j, k), IDX2(L, i, j + 1, k), IDX2(L, i, j + 2, k));
        }
        if (j == NY + B2) {
            j = j - 1;
            BWD(L->cy, j)
            g.y = S3(IDX2(L, i, j, k), IDX2(L, i, j - 1, k), IDX2(L, i, j - 2, k));
        }
        if (k == B2 - 1) {
            k = k + 1;
            FWD(L->cz, k)
            g.z = S3(IDX2(L, i, j, k), IDX2(L, i, j, k + 1), IDX2(L, i, j, k + 2));
        }
        if (k == NZ + B2) {
            k = k - 1;
            BWD(L->cz, k)
            g.z = S3(IDX2(L, i, j, k), IDX2(L, i, j, k - 1), IDX2(L, i, j, k - 2));
        }
    } else {
        if (i == B2 - 1) {
            FWD(L->cx, i)
            g.x = S3(IDX2(L, i, j, k), IDX2(L, i + 1, j, k), IDX2(L, i + 2, j, k));
        }
        if (i == NX + B2) {
            BWD(L->cx, i)
            g.x = S3(IDX2(L, i, j, k), IDX2(L, i - 1, j, k), IDX2(L, i - 2, j, k));
        }
        if (j == B2 - 1) {
            FWD(L->cy, j)
            g.y = S3(IDX2(L, i, j, k), IDX2(L, i, j + 1, k), IDX2(L, i, j + 2, k));
        }
        if (j == NY + B2) {
            BWD(L->cy, j)
            g.y = S3(IDX2(L, i, j, k), IDX2(L, i, j - 1, k), IDX2(L, i, j - 2, k));
        }
        if (k == B2 - 1) {
            FWD(L->cz, k)
            g.z = S3(IDX2(L, i, j, k), IDX2(L, i, j, k + 1), IDX2(L, i, j, k + 2));
        }
        if (k == NZ + B2) {
            BWD(L->cz, k)
            g.z = S3(IDX2(L, i, j, k), IDX2(L, i, j, k - 1), IDX2(L, i, j, k - 2));
        }
    }
#undef S3
    if (NX == 1) g.x = 0.0;
    if (NY == 1) g.y = 0.0;
    if (NZ == 1) g.z = 0.0;
    return g;
}
#undef FWD
#undef BWD

/* interpolate_vector_at_particles(u,'u') and interpolate_scalar_at_particles
 * (t,'t') (lagrange.c:1211-1593), with the gradients of both under
 * LAGRANGE_GRADIENT; `fluid` / `thermal` select which of the two runs.    */
static void interpolate_fields(oracle *o, int fluid, int thermal)
{
    olagr *L = o->lagr;
    const int grad = LHAS(L, CH4G_LAGRANGE_GRADIENT);
    for (int ip = 0; ip < L->npart; ip++) {
        ch4g_particle *p = &L->tr[ip];
        ostencil s;
        if (stencil(o, p, &s)) continue;
        if (fluid && HAS(o, CH4G_LB_FLUID)) {
            double vx = L->u[s.n[0]].x * s.w[0], vy = L->u[s.n[0]].y * s.w[0], vz = L->u[s.n[0]].z * s.w[0];
            for (int c = 1; c < 8; c++) {
                vx += L->u[s.n[c]].x * s.w[c];
                vy += L->u[s.n[c]].y * s.w[c];
                vz += L->u[s.n[c]].z * s.w[c];
            }
            p->ux = vx;
            p->uy = vy;
            p->uz = vz;
            if (grad) {
                oten G[8];
                for (int c = 0; c < 8; c++) G[c] = grad_vec(o, s.im + CORNER[c][0], s.jm + CORNER[c][1], s.km + CORNER[c][2]);
                double a[9];
                for (int m = 0; m < 9; m++) {
                    double acc = ((const double *)&G[0])[m] * s.w[0];
                    for (int c = 1; c < 8; c++) acc += ((const double *)&G[c])[m] * s.w[c];
                    a[m] = acc;
                }
                /* oten order xx,xy,xz,yx,yy,yz,zx,zy,zz; dy_ux = grad.xy (lagrange.c:1420-1428) */
                p->dx_ux = a[0];
                p->dy_ux = a[1];
                p->dz_ux = a[2];
                p->dx_uy = a[3];
                p->dy_uy = a[4];
                p->dz_uy = a[5];
                p->dx_uz = a[6];
                p->dy_uz = a[7];
                p->dz_uz = a[8];
            }
        }
        if (thermal && HAS(o, CH4G_LB_TEMPERATURE)) {
            double v = L->t[s.n[0]] * s.w[0];
            for (int c = 1; c < 8; c++) v += L->t[s.n[c]] * s.w[c];
            p->t = v;
            if (grad) { /* lagrange.c:1500-1576 */
                ovec G[8];
                for (int c = 0; c < 8; c++) G[c] = grad_sca(o, s.im + CORNER[c][0], s.jm + CORNER[c][1], s.km + CORNER[c][2]);
                double gx = G[0].x * s.w[0], gy = G[0].y * s.w[0], gz = G[0].z * s.w[0];
                for (int c = 1; c < 8; c++) {
                    gx += G[c].x * s.w[c];
                    gy += G[c].y * s.w[c];
                    gz += G[c].z * s.w[c];
                }
                p->dx_t = gx;
                p->dy_t = gy;
                p->dz_t = gz;
            }
        }
    }
}

static void interpolate(oracle *o) { interpolate_fields(o, 1, 1); }

/* move_particles (lagrange.c:2934-4025) for the switches of ch4g.h, then
 * sendrecv_particles (lagrange.c:4028-4221) on one rank.                   */
/* Orientation update of move_particles (lagrange.c:3470-3752) with
 * LAGRANGE_ORIENTATION_JEFFREY (+ _GYROTAXIS): Jeffery's equation for the
 * axis p of a spheroid of aspect ratio alpha,
 *   dp/dt = W p + f S p - (p . f S p) p,   f = (alpha^2 - 1) / (alpha^2 + 1),
 * S and W the symmetric and antisymmetric parts of A[i][j] = d u_i / d x_j,
 * Euler on the first step, Adams-Bashforth 2 after; p renormalised. */
static void orient_step(oracle *o, const olagr *L, ch4g_particle *p, int first, double gx, double gy, double gz)
{
    double P[3] = {p->px, p->py, p->pz};
    double Fold[3] = {p->dt_px, p->dt_py, p->dt_pz};
    const int second = LHAS(L, CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION);
    double Nv[3] = {p->nx, p->ny, p->nz};
    double FNold[3] = {p->dt_nx, p->dt_ny, p->dt_nz};
    if (LHAS(L, CH4G_LAGRANGE_ORIENTATION_JEFFREY)) {
        const double al = p->aspect_ratio;
        const double fa = (al * al - 1.0) / (1.0 + al * al);
        const double A[3][3] = {{p->dx_ux, p->dy_ux, p->dz_ux}, {p->dx_uy, p->dy_uy, p->dz_uy},
                                {p->dx_uz, p->dy_uz, p->dz_uz}};
        double S[3][3], W[3][3], F[3], T[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                S[i][j] = 0.5 * (A[i][j] + A[j][i]);
                W[i][j] = 0.5 * (A[i][j] - A[j][i]);
            }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) S[i][j] *= fa;
        if (LHAS(L, CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS) && p->gyrotaxis_velocity != 0) {
            /* lagrange.c:3533-3566: the first column of S takes a_i p_0; the
             * six other products land in A, which is not read again */
            const double g = -0.5 / p->gyrotaxis_velocity;
            double a[3];
            if (LHAS(L, CH4G_LAGRANGE_GRAVITY)) {
                a[0] = g * (-gx + p->ax);
                a[1] = g * (-gy + p->ay);
                a[2] = g * (-gz + p->az);
            } else {
                a[0] = g * 0.0;
                a[1] = g * 0.0;
                a[2] = g * 1.0;
            }
            S[0][0] += a[0] * P[0];
            S[1][0] += a[1] * P[0];
            S[2][0] += a[2] * P[0];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) T[i] += S[i][j] * P[j];
        double oso = 0.0;
        for (int i = 0; i < 3; i++) oso += P[i] * T[i];
        for (int i = 0; i < 3; i++) {
            F[i] = 0.0;
            for (int j = 0; j < 3; j++) F[i] += W[i][j] * P[j] + (S[i][j] * P[j]);
            F[i] -= P[i] * oso;
        }
        /* second orientation (lagrange.c:3586-3616): W n - (p . S n) p */
        double FN[3] = {0, 0, 0}, TN[3] = {0, 0, 0}, osn = 0.0;
        if (second) {
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) TN[i] += S[i][j] * Nv[j];
            for (int i = 0; i < 3; i++) osn += P[i] * TN[i];
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++) FN[i] += W[i][j] * Nv[j];
                FN[i] -= P[i] * osn;
            }
        }
        for (int i = 0; i < 3; i++) {
            P[i] = first ? P[i] + L->dt * F[i] : P[i] + 0.5 * L->dt * (3. * F[i] - Fold[i]);
            Fold[i] = F[i];
        }
        if (second)
            for (int i = 0; i < 3; i++) {
                Nv[i] = first ? Nv[i] + L->dt * FN[i] : Nv[i] + 0.5 * L->dt * (3. * FN[i] - FNold[i]);
                FNold[i] = FN[i];
            }
    }
    if (L->pc.flags2 & CH4G_LAGRANGE_ORIENTATION_RANDOM) {
        /* lagrange.c:3698-3713: p replaced by random_vector (random.c:139-150),
         * dp/dt from the stored p */
        double theta = acos(2.0 * oracle_drand48(o) - 1.0);
        double phi = 2 * 3.14159265359 * oracle_drand48(o);
        P[0] = sin(theta) * sin(phi);
        P[1] = sin(theta) * cos(phi);
        P[2] = cos(theta);
        Fold[0] = (P[0] - p->px) / L->dt;
        Fold[1] = (P[1] - p->py) / L->dt;
        Fold[2] = (P[2] - p->pz) / L->dt;
    }
    double norm = 0.0;
    for (int i = 0; i < 3; i++) norm += P[i] * P[i];
    for (int i = 0; i < 3; i++) P[i] /= sqrt(norm);
    p->px = P[0];
    p->py = P[1];
    p->pz = P[2];
    p->dt_px = Fold[0];
    p->dt_py = Fold[1];
    p->dt_pz = Fold[2];
    if (second) {
        double nn = 0.0;
        for (int i = 0; i < 3; i++) nn += Nv[i] * Nv[i];
        for (int i = 0; i < 3; i++) Nv[i] /= sqrt(nn);
        p->nx = Nv[0];
        p->ny = Nv[1];
        p->nz = Nv[2];
        p->dt_nx = FNold[0];
        p->dt_ny = FNold[1];
        p->dt_nz = FNold[2];
    }
}

static void move_at(oracle *o, int first)
{
    olagr *L = o->lagr;
    const double dt = L->dt;
    const double gx = o->cfg.gravity_x, gy = o->cfg.gravity_y, gz = o->cfg.gravity_z;
    const double SX = o->NX, SY = o->NY, SZ = o->NZ;
    for (int ip = 0; ip < L->npart; ip++) {
        ch4g_particle *p = &L->tr[ip];
        if (HAS(o, CH4G_LB_TEMPERATURE)) {
            if (first) {
                p->t_old = p->t;
            } else {
                p->dt_t = (p->t - p->t_old) / dt;
                p->t_old = p->t;
            }
        }
        if (p->tau_drag < 0.0) { /* Eulerian probe: not moved, v = u (lagrange.c:3024-3033) */
            p->vx = p->ux;
            p->vy = p->uy;
            p->vz = p->uz;
        }
        if (p->tau_drag == 0.0) { /* fluid tracer (lagrange.c:3010-3227) */
            p->vx = p->ux;
            p->vy = p->uy;
            p->vz = p->uz;
            if (LHAS(L, CH4G_LAGRANGE_ORIENTATION) && LHAS(L, CH4G_LAGRANGE_ORIENTATION_ACTIVE)) {
                /* swimming: the #elif chain of lagrange.c:3048-3186 */
                const unsigned f2 = L->pc.flags2;
                if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) {
                    double shear_rate, velocity_amplitude, jump_time_duration;
                    if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE) {
                        shear_rate = p->t;
                    } else {
                        p->shear_rate = shear_rate =
                            sqrt(2.) * sqrt(p->dx_ux * p->dx_ux + p->dy_uy * p->dz_uy + p->dz_uz * p->dz_uz +
                                            0.5 * (p->dy_ux + p->dx_uy) * (p->dy_ux + p->dx_uy) +
                                            0.5 * (p->dz_ux + p->dx_uz) * (p->dz_ux + p->dx_uz) +
                                            0.5 * (p->dy_uz + p->dz_uy) * (p->dy_uz + p->dz_uy));
                    }
                    jump_time_duration = -(p->jump_time) * log(0.01);
                    if (p->time_from_jump > jump_time_duration) {
                        velocity_amplitude = 0.0;
                        int go = (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS) ? shear_rate < p->critical_shear_rate
                                                                                        : shear_rate > p->critical_shear_rate;
                        if (go) {
                            p->time_from_jump = 0.0;
                            velocity_amplitude = p->swim_velocity;
                            p->px_jump = p->px;
                            p->py_jump = p->py;
                            p->pz_jump = p->pz;
                        }
                    } else {
                        velocity_amplitude = p->swim_velocity * exp(-(p->time_from_jump) / (p->jump_time));
                        if (p->jump_time == 0.0) velocity_amplitude = 0.0;
                    }
                    p->vx += velocity_amplitude * p->px_jump;
                    p->vy += velocity_amplitude * p->py_jump;
                    p->vz += velocity_amplitude * p->pz_jump;
                    p->time_from_jump += dt;
                } else if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC) {
                    p->vx = p->swim_velocity * p->px;
                    p->vy = p->swim_velocity * p->py;
                    p->vz = p->swim_velocity * p->pz;
                } else if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE) {
                    double reactivity = p->dt_t > 0.0 ? 1.0 : 0.0;
                    p->vx += reactivity * p->swim_velocity * p->px;
                    p->vy += reactivity * p->swim_velocity * p->py;
                    p->vz += reactivity * p->swim_velocity * p->pz;
                } else if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION) {
                    p->vx += p->swim_velocity * p->nx;
                    p->vy += p->swim_velocity * p->ny;
                    p->vz += p->swim_velocity * p->nz;
                } else if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION) {
                    if (p->aspect_ratio >= 1.0) { /* rods */
                        p->vx += p->swim_velocity * p->px;
                        p->vy += p->swim_velocity * p->py;
                        p->vz += p->swim_velocity * p->pz;
                    } else { /* disks */
                        p->vx += p->swim_velocity * p->nx;
                        p->vy += p->swim_velocity * p->ny;
                        p->vz += p->swim_velocity * p->nz;
                    }
                } else {
                    p->vx += p->swim_velocity * p->px;
                    p->vy += p->swim_velocity * p->py;
                    p->vz += p->swim_velocity * p->pz;
                }
            }
            if (first) {
                p->vx_old = p->vx;
                p->vy_old = p->vy;
                p->vz_old = p->vz;
            }
            p->ax = (p->vx - p->vx_old) / dt;
            p->ay = (p->vy - p->vy_old) / dt;
            p->az = (p->vz - p->vz_old) / dt;
            if (first) {
                p->x += dt * p->vx;
                p->y += dt * p->vy;
                p->z += dt * p->vz;
            } else {
                p->x += dt * 0.5 * (3.0 * p->vx - p->vx_old);
                p->y += dt * 0.5 * (3.0 * p->vy - p->vy_old);
                p->z += dt * 0.5 * (3.0 * p->vz - p->vz_old);
            }
            p->vx_old = p->vx;
            p->vy_old = p->vy;
            p->vz_old = p->vz;
            p->ux_old = p->ux;
            p->uy_old = p->uy;
            p->uz_old = p->uz;
        }
        if (p->tau_drag > 0.0) { /* inertial particle (lagrange.c:3229-3466) */
            double dux = 0.0, duy = 0.0, duz = 0.0; /* Dt_u */
            double invtau = 1.0 / p->tau_drag;
            if (L->pc.flags2 & CH4G_LAGRANGE_ORIENTATION_DRAG) {
                /* Stokes drag of a spheroid (lagrange.c:3267-3310, PRL 119,
                 * 254501): resistance across and along the axis p relative to
                 * the sphere of radius a_perp, time tau_drag * aspect_ratio */
                double alpha = p->aspect_ratio, beta = 0.0, c_perp, c_par;
                invtau = 1.0 / (p->tau_drag * alpha);
                if (alpha == 1) {
                    beta = c_perp = c_par = 1.0;
                } else {
                    if (alpha > 1) beta = log(alpha + sqrt(alpha * alpha - 1.0)) / (alpha * sqrt(alpha * alpha - 1.0));
                    if (alpha < 1) beta = acos(alpha) / (alpha * sqrt(1.0 - alpha * alpha));
                    c_perp = 8.0 * (alpha * alpha - 1.) / (3.0 * alpha * ((2.0 * alpha * alpha - 3.0) * beta + 1.0));
                    c_par = 4.0 * (alpha * alpha - 1.) / (3.0 * alpha * ((2.0 * alpha * alpha - 1.0) * beta - 1.0));
                }
                double P0 = p->px, P1 = p->py, P2 = p->pz, M[3][3];
                M[0][0] = c_perp + (c_par - c_perp) * P0 * P0;
                M[1][1] = c_perp + (c_par - c_perp) * P1 * P1;
                M[2][2] = c_perp + (c_par - c_perp) * P2 * P2;
                M[0][1] = M[1][0] = (c_par - c_perp) * P0 * P1;
                M[0][2] = M[2][0] = (c_par - c_perp) * P0 * P2;
                M[1][2] = M[2][1] = (c_par - c_perp) * P1 * P2;
                double uvx = p->ux - p->vx, uvy = p->uy - p->vy, uvz = p->uz - p->vz;
                p->ax = (M[0][0] * uvx + M[0][1] * uvy + M[0][2] * uvz) * invtau;
                p->ay = (M[1][0] * uvx + M[1][1] * uvy + M[1][2] * uvz) * invtau;
                p->az = (M[2][0] * uvx + M[2][1] * uvy + M[2][2] * uvz) * invtau;
            } else {
                p->ax = (p->ux - p->vx) * invtau;
                p->ay = (p->uy - p->vy) * invtau;
                p->az = (p->uz - p->vz) * invtau;
            }
            if ((L->pc.flags2 & CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG) &&
                !(L->pc.flags2 & CH4G_LAGRANGE_ORIENTATION_DRAG)) { /* lagrange.c:3242-3264 */
                double diameter_p = 2.0 * sqrt(p->tau_drag * 3.0 * p->beta_coeff * o->nu);
                double re_p = sqrt((p->ux - p->vx) * (p->ux - p->vx) + (p->uy - p->vy) * (p->uy - p->vy) +
                                   (p->uz - p->vz) * (p->uz - p->vz)) *
                              diameter_p / o->nu;
                invtau = 0.15 * pow(re_p, 0.687) / p->tau_drag;
                p->ax += (p->ux - p->vx) * invtau;
                p->ay += (p->uy - p->vy) * invtau;
                p->az += (p->uz - p->vz) * invtau;
            }
            const int gvar = LHAS(L, CH4G_LAGRANGE_GRAVITY_VARIABLE); /* lagrange.c:3314-3335 */
            if (LHAS(L, CH4G_LAGRANGE_GRAVITY)) {
                if (gvar) {
                    p->ax -= p->gravity_coeff * gx;
                    p->ay -= p->gravity_coeff * gy;
                    p->az -= p->gravity_coeff * gz;
                } else {
                    p->ax -= gx;
                    p->ay -= gy;
                    p->az -= gz;
                }
            }
            if (LHAS(L, CH4G_LAGRANGE_ADDEDMASS) && p->beta_coeff != 0.0) {
                if (LHAS(L, CH4G_LAGRANGE_GRAVITY)) {
                    if (gvar) {
                        p->ax -= (-p->beta_coeff) * p->gravity_coeff * gx;
                        p->ay -= (-p->beta_coeff) * p->gravity_coeff * gy;
                        p->az -= (-p->beta_coeff) * p->gravity_coeff * gz;
                    } else {
                        p->ax -= (-p->beta_coeff) * gx;
                        p->ay -= (-p->beta_coeff) * gy;
                        p->az -= (-p->beta_coeff) * gz;
                    }
                }
                if (first) {
                    p->ux_old = p->ux;
                    p->uy_old = p->uy;
                    p->uz_old = p->uz;
                }
                dux = (p->ux - p->ux_old) / dt + (p->ux - p->vx) * p->dx_ux + (p->uy - p->vy) * p->dy_ux +
                      (p->uz - p->vz) * p->dz_ux;
                duy = (p->uy - p->uy_old) / dt + (p->ux - p->vx) * p->dx_uy + (p->uy - p->vy) * p->dy_uy +
                      (p->uz - p->vz) * p->dz_uy;
                duz = (p->uz - p->uz_old) / dt + (p->ux - p->vx) * p->dx_uz + (p->uy - p->vy) * p->dy_uz +
                      (p->uz - p->vz) * p->dz_uz;
                p->ax += dux * p->beta_coeff;
                p->ay += duy * p->beta_coeff;
                p->az += duz * p->beta_coeff;
                p->Dt_ux = dux;
                p->Dt_uy = duy;
                p->Dt_uz = duz;
                if (L->pc.flags2 & CH4G_LAGRANGE_ADDEDMASS_LIFT) { /* lagrange.c:3365-3380 */
                    double ox = p->dy_uz - p->dz_uy, oy = p->dz_ux - p->dx_uz, oz = p->dx_uy - p->dy_ux;
                    double lift_coeff = p->beta_coeff / 3.0;
                    p->ax += lift_coeff * ((p->uy - p->vy) * oz - (p->uz - p->vz) * oy);
                    p->ay += lift_coeff * ((p->uz - p->vz) * ox - (p->ux - p->vx) * oz);
                    p->az += lift_coeff * ((p->ux - p->vx) * oy - (p->uy - p->vy) * ox);
                }
            }
            if (first) {
                p->x += dt * p->vx;
                p->y += dt * p->vy;
                p->z += dt * p->vz;
            } else {
                p->x += dt * 0.5 * (3.0 * p->vx - p->vx_old);
                p->y += dt * 0.5 * (3.0 * p->vy - p->vy_old);
                p->z += dt * 0.5 * (3.0 * p->vz - p->vz_old);
            }
            p->vx_old = p->vx;
            p->vy_old = p->vy;
            p->vz_old = p->vz;
            p->ux_old = p->ux;
            p->uy_old = p->uy;
            p->uz_old = p->uz;
            if (first) {
                p->vx += dt * p->ax;
                p->vy += dt * p->ay;
                p->vz += dt * p->az;
            } else {
                p->vx += dt * 0.5 * (3.0 * p->ax - p->ax_old);
                p->vy += dt * 0.5 * (3.0 * p->ay - p->ay_old);
                p->vz += dt * 0.5 * (3.0 * p->az - p->az_old);
            }
            p->ax_old = p->ax;
            p->ay_old = p->ay;
            p->az_old = p->az;
            if (LHAS(L, CH4G_LAGRANGE_SMALLTAUD_BETA) && p->tau_drag < 10.0 && p->tau_drag > 0.0) {
                p->vx = p->ux;
                p->vy = p->uy;
                p->vz = p->uz;
                p->vx -= gx * (1.0 - p->beta_coeff) * p->tau_drag;
                p->vy -= gy * (1.0 - p->beta_coeff) * p->tau_drag;
                p->vz -= gz * (1.0 - p->beta_coeff) * p->tau_drag;
                p->vx -= dux * (1.0 - p->beta_coeff) * p->tau_drag;
                p->vy -= duy * (1.0 - p->beta_coeff) * p->tau_drag;
                p->vz -= duz * (1.0 - p->beta_coeff) * p->tau_drag;
            }
            if (LHAS(L, CH4G_LAGRANGE_SMALLTAUD_FLOATER) && p->tau_drag > 0.0) {
                p->vx = p->ux;
                p->vy = p->uy;
                p->vz = p->uz;
                double f1 = p->tau_drag * (1.0 - p->beta_coeff + 2. / 3. * o->cfg.beta_t * (p->t - 0.4));
                p->vx += f1 * (-gx - p->Dt_ux);
                p->vy += f1 * (-gy - p->Dt_uy);
                p->vz += f1 * (-gz - p->Dt_uz);
            }
        }
        if (LHAS(L, CH4G_LAGRANGE_ORIENTATION)) orient_step(o, L, p, first, gx, gy, gz);
        /* wall rules (lagrange.c:3754-3993), LB_FLUID_BC_Y only */
        if (HAS(o, CH4G_FLUID_BC_Y)) {
            const int inel = LHAS(L, CH4G_LAGRANGE_BC_INELASTIC);
            const int reinj = LHAS(L, CH4G_LAGRANGE_BC_INELASTIC_REINJECT);
            const double fac1 = inel ? 0.0 : -1.0, fac2 = inel ? 0.0 : 1.0;
            /* lagrange.c:3775-3788: 0, or with LAGRANGE_BC_RADIUSandDENSITY the
             * radius from tau = r^2 / (3 beta nu) */
            double radius = 0.0;
            if (L->pc.flags2 & CH4G_LAGRANGE_BC_RADIUSANDDENSITY)
                radius = sqrt(p->tau_drag * o->nu * 3.0 * p->beta_coeff);
            if (p->y < radius) {
                if (!reinj) {
                    p->y = fac1 * (p->y - radius) + radius;
                    p->vx *= fac2;
                    if (p->vy < 0.0) p->vy *= fac1;
                    p->vz *= fac2;
                }
                if (inel && reinj) {
                    if (LHAS(L, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM)) {
                        p->y += SY;
                    } else {
                        p->x = SX * oracle_drand48(o);
                        p->y += SY;
                        p->z = SZ * oracle_drand48(o);
                    }
                    if (LHAS(L, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO)) {
                        p->vx = 0.0;
                        p->vy = 0.0;
                        p->vz = 0.0;
                    } else if (LHAS(L, CH4G_LAGRANGE_GRAVITY_VARIABLE)) { /* lagrange.c:3824-3827, :3905-3908 */
                        p->vx = -gx * p->gravity_coeff * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vy = -gy * p->gravity_coeff * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vz = -gz * p->gravity_coeff * (1.0 - p->beta_coeff) * p->tau_drag;
                    } else {
                        p->vx = -gx * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vy = -gy * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vz = -gz * (1.0 - p->beta_coeff) * p->tau_drag;
                    }
                    if (LHAS(L, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF)) {
                        p->vx /= 2.0;
                        p->vy /= 2.0;
                        p->vz /= 2.0;
                    }
                    /* the particle takes the temperature of the wall it re-enters at */
                    if (LHAS(L, CH4G_LAGRANGE_TEMPERATURE)) p->t_p = o->cfg.T_top;
                }
                /* LAGRANGE_BC_INELASTIC_SEDIMENT, without nucleation: count the
                 * deposition (lagrange.c:3845-3863) */
                if (inel && (L->pc.flags2 & CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT)) p->sediment += 1.0;
                /* LAGRANGE_ORIENTATION_BC (lagrange.c:3866-3870) */
                if (LHAS(L, CH4G_LAGRANGE_ORIENTATION_BC) && p->py < 0.0) p->py *= -1.0;
            }
            if (p->y >= SY - radius) {
                if (!reinj) {
                    p->y = fac1 * (p->y - (SY - radius)) + (SY - radius);
                    p->vx *= fac2;
                    if (p->vy > 0.0) p->vy *= fac1;
                    p->vz *= fac2;
                }
                if (inel && reinj) {
                    if (LHAS(L, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM)) {
                        p->y -= (SY - FLT_EPSILON);
                    } else {
                        p->x = SX * oracle_drand48(o);
                        p->y -= (SY - FLT_EPSILON);
                        p->z = SZ * oracle_drand48(o);
                    }
                    if (LHAS(L, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO)) {
                        p->vx = 0.0;
                        p->vy = 0.0;
                        p->vz = 0.0;
                    } else if (LHAS(L, CH4G_LAGRANGE_GRAVITY_VARIABLE)) { /* lagrange.c:3824-3827, :3905-3908 */
                        p->vx = -gx * p->gravity_coeff * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vy = -gy * p->gravity_coeff * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vz = -gz * p->gravity_coeff * (1.0 - p->beta_coeff) * p->tau_drag;
                    } else {
                        p->vx = -gx * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vy = -gy * (1.0 - p->beta_coeff) * p->tau_drag;
                        p->vz = -gz * (1.0 - p->beta_coeff) * p->tau_drag;
                    }
                    if (LHAS(L, CH4G_LAGRANGE_TEMPERATURE)) p->t_p = o->cfg.T_bot;
                }
                /* LAGRANGE_ORIENTATION_BC (lagrange.c:3942-3946) */
                if (LHAS(L, CH4G_LAGRANGE_ORIENTATION_BC) && p->py > 0.0) p->py *= -1.0;
            }
        }
        /* polymer conformation tensor (polymer.c:150-240), after the wall
         * rules (lagrange.c:3998-4002); time_dt = 1, and the Euler branch of
         * itime == 0 is never reached inside move_particles */
        if (LHAS(L, CH4G_LAGRANGE_GRADIENT) && (L->pc.flags2 & CH4G_LAGRANGE_POLYMER)) {
            double A[3][3] = {{p->dx_ux, p->dy_ux, p->dz_ux}, {p->dx_uy, p->dy_uy, p->dz_uy},
                              {p->dx_uz, p->dy_uz, p->dz_uz}};
            double C[3][3] = {{p->cxx, p->cxy, p->cxz}, {p->cxy, p->cyy, p->cyz}, {p->cxz, p->cyz, p->czz}};
            double Fo[3][3] = {{p->dt_cxx, p->dt_cxy, p->dt_cxz}, {p->dt_cxy, p->dt_cyy, p->dt_cyz},
                               {p->dt_cxz, p->dt_cyz, p->dt_czz}};
            double I3[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
            double F[3][3];
            double invtau = 1. / L->pc.tau_polymer;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    F[i][j] = 0.0;
                    for (int k = 0; k < 3; k++) F[i][j] += A[k][i] * C[k][j] + C[i][k] * A[k][j];
                    F[i][j] += (I3[i][j] - C[i][j]) * invtau;
                }
            for (int j = 0; j < 3; j++)
                for (int i = 0; i <= j; i++) C[i][j] = C[i][j] + 0.5 * 1.0 * (3. * F[i][j] - Fo[i][j]);
            p->cxx = C[0][0];
            p->cyy = C[1][1];
            p->czz = C[2][2];
            p->cxy = C[0][1];
            p->cyz = C[1][2];
            p->cxz = C[0][2];
            p->dt_cxx = F[0][0];
            p->dt_cyy = F[1][1];
            p->dt_czz = F[2][2];
            p->dt_cxy = F[0][1];
            p->dt_cyz = F[1][2];
            p->dt_cxz = F[0][2];
        }
        /* particle temperature (lagrange.c:4007-4020): first-order relaxation
         * to the fluid's t, T_p <- (T_p + fac1 t) exp(-fac1), fac1 the
         * sub-step over the heat response time */
        if (HAS(o, CH4G_LB_TEMPERATURE) && LHAS(L, CH4G_LAGRANGE_TEMPERATURE)) {
            double fac1 = (o->kappa * 2.0) / (p->tau_drag * o->nu * p->cp_p * (3 - p->beta_coeff));
            fac1 *= dt;
            p->t_p = (p->t_p + fac1 * p->t) * exp(-fac1);
        }
    }
    /* sendrecv_particles: wrap, keep the particles inside [0, S) (mesh
     * bounds, lagrange.c:4090-4092); on one rank the others are dropped
     * (they come back from the Allgatherv and fail the same test)          */
    int kept = 0;
    for (int ip = 0; ip < L->npart; ip++) {
        ch4g_particle *p = &L->tr[ip];
        double px = WRAP(p->x, SX), py = WRAP(p->y, SY), pz = WRAP(p->z, SZ);
        if (LHAS(L, CH4G_LAGRANGE_WRAP)) {
            p->x = WRAP(p->x, SX);
            p->z = WRAP(p->z, SZ);
            p->y = WRAP(p->y, SY);
        }
        if (px >= 0.0 && px < SX && py >= 0.0 && py < SY && pz >= 0.0 && pz < SZ) L->tr[kept++] = *p;
    }
    L->npart = kept;
}

static void move(oracle *o) { move_at(o, o->itime == 1); /* itime == 1 && resume == 0 */ }

void olagr_after_step(oracle *o)
{
    olagr *L = o->lagr;
    bc_hydro(o);
    for (int s = 0; s < (int)L->pc.time_dt_lagr_factor; s++) {
        interpolate(o);
        move(o);
    }
}

/* add_particle_feedbacks (lagrange_feedback.c:11-256), trilinear
 * extrapolation, momentum part.  add_sendrecv_borders_vector (:525-836)
 * only ever adds into ghost nodes (it sends interior values and accumulates
 * what it receives into the halo), so the interior force receives the
 * deposits made directly on interior nodes; deposits on ghost nodes do not
 * reach it.                                                                 */
void olagr_feedback(oracle *o)
{
    olagr *L = o->lagr;
    memset(L->ftw, 0, L->ng * sizeof(ovec));
    const double nu = o->nu;
    ovec gv = {0.0, 0.0, 0.0};
    if (LHAS(L, CH4G_LAGRANGE_GRAVITY)) {
        gv.x = o->cfg.gravity_x;
        gv.y = -o->cfg.gravity_y;
        gv.z = o->cfg.gravity_z;
    }
    for (int ip = 0; ip < L->npart; ip++) {
        const ch4g_particle *p = &L->tr[ip];
        ostencil s;
        if (stencil(o, p, &s)) continue;
        double density_ratio = (3.0 - p->beta_coeff) / (2.0 * p->beta_coeff);
        double squared_radius = 3.0 * nu * p->beta_coeff * p->tau_drag;
        double fac = (4. / 3.) * 3.14159265359 * pow(squared_radius, 3. / 2.);
        ovec fp;
        if (LHAS(L, CH4G_LAGRANGE_TWOWAY_SMALLTAUD)) {
            fp.x = (p->Dt_ux - gv.x) * (1 - density_ratio);
            fp.y = (p->Dt_uy - gv.y) * (1 - density_ratio);
            fp.z = (p->Dt_uz - gv.z) * (1 - density_ratio);
        } else {
            fp.x = p->Dt_ux - gv.x + density_ratio * (gv.x - p->ax);
            fp.y = p->Dt_uy - gv.y + density_ratio * (gv.y - p->ay);
            fp.z = p->Dt_uz - gv.z + density_ratio * (gv.z - p->az);
        }
        for (int c = 0; c < 8; c++) L->ftw[s.n[c]].x += fac * fp.x * s.w[c];
        for (int c = 0; c < 8; c++) L->ftw[s.n[c]].y += fac * fp.y * s.w[c];
        for (int c = 0; c < 8; c++) L->ftw[s.n[c]].z += fac * fp.z * s.w[c];
    }
    for (int k = 0; k < o->NZ; k++)
        for (int j = 0; j < o->NY; j++)
            for (int i = 0; i < o->NX; i++) {
                size_t n = IDX(o, i + BRD, j + BRD, k + BRD), m = IDX2(L, i + B2, j + B2, k + B2);
                o->force[n].x += L->ftw[m].x;
                o->force[n].y += L->ftw[m].y;
                o->force[n].z += L->ftw[m].z;
            }
}

/* ------------------------------------------------------------------------ */
/* Public API (ctypes). */
long oracle_particles_count(const oracle *o) { return o->lagr ? o->lagr->npart : 0; }

long oracle_get_particles(const oracle *o, ch4g_particle *out, long cap)
{
    if (!o->lagr) return 0;
    long n = o->lagr->npart < cap ? o->lagr->npart : cap;
    memcpy(out, o->lagr->tr, (size_t)n * sizeof(ch4g_particle));
    return o->lagr->npart;
}

int oracle_set_particles(oracle *o, const ch4g_particle *in, long n)
{
    olagr *L = o->lagr;
    if (!L) return -1;
    ch4g_particle *t = (ch4g_particle *)realloc(L->tr, (size_t)(n > 0 ? n : 1) * sizeof(ch4g_particle));
    if (!t) return -1;
    L->tr = t;
    memcpy(L->tr, in, (size_t)n * sizeof(ch4g_particle));
    L->npart = (int)n;
    return 0;
}

/* Interpolation probe for pinning tests: velocity / temperature fields of
 * the interior given directly ([NZ][NY][NX], u interleaved), halo rules of
 * boundary_conditions_hydro, interpolation at the particles' positions.     */
int oracle_particles_probe(oracle *o, const double *u, const double *t)
{
    if (!o->lagr) return -1;
    size_t m = 0;
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++, m++) {
                size_t n = IDX(o, i, j, k);
                if (u) {
                    o->u[n].x = u[3 * m];
                    o->u[n].y = u[3 * m + 1];
                    o->u[n].z = u[3 * m + 2];
                }
                if (t) o->t[n] = t[m];
            }
    bc_hydro(o);
    interpolate(o);
    return 0;
}

/* Test hook: one particle sub-step (interpolate + move_particles) on given
 * interior fields u, t held fixed, Euler start when `first`. */
int oracle_particles_move(oracle *o, const double *u, const double *t, int first)
{
    if (oracle_particles_probe(o, u, t)) return -1;
    move_at(o, first);
    return 0;
}
