/*
 * ch4_oracle.c — FP64 CPU restatement of the ch4-project streaming
 * lattice-Boltzmann step.  TEST INFRASTRUCTURE ONLY: it is the checker for the
 * HIP library and the CPU baseline leg of bench.py; nothing in the product
 * links, loads or calls it.
 *
 * It restates the reference's algorithm with the reference's structure:
 * AoS `pop` records with a one-node ghost halo, one full sweep per reference
 * function, called in the order of src/main.c:23-48 (init) and
 * src/main.c:92-161 (step), with the same floating-point evaluation order
 * (the reference is compiled -std=c11, i.e. without FP contraction), so that
 * its fields are the reference's fields on the same inputs.
 *
 * Each function cites the reference lines it follows.  How each part is
 * pinned (closed forms, invariants, glibc) is listed in tests/test_oracle_*.py
 * (test_oracle_pins.py: RNG, collision, streaming, walls, forcing lag;
 * test_oracle_forcing.py: HIT, buoyancy and channel forcing; test_oracle_les.py,
 * test_oracle_rulers.py, test_oracle_particles.py).  One boundary is left
 * "parity unpinned": the libm routines (sin, pow, sqrt, exp, log) are
 * glibc's, which the reference links as well; no reference test pins them.
 *
 * oracle_set_plant() switches in one deliberate mistake (PLANT_* below) so
 * that the pinning tests can show that they fail on it; 0 (the default)
 * is the restatement itself.
 *
 * Build: oracle/Makefile (gcc -O2 -std=c11 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "ch4_oracle_internal.h"

/* ------------------------------------------------------------------------ */
/* drand48: POSIX 48-bit LCG, X <- (a X + c) mod 2^48, a = 0x5DEECE66D,
 * c = 0xB; srand48(s) sets X = (s << 16) | 0x330E.  Used by the reference
 * through myrand() (random.c:4-16, SYSTEM_RANDOM48 define.h:14).
 * Pinned against glibc drand48 in tests/test_oracle_pins.py.               */
void oracle_srand48(oracle *o, long seed)
{
    o->x48 = ((((uint64_t)(uint32_t)seed) << 16) | 0x330Eu) & 0xFFFFFFFFFFFFull;
}

double oracle_drand48(oracle *o)
{
    o->x48 = (o->x48 * 0x5DEECE66Dull + 0xBull) & 0xFFFFFFFFFFFFull;
    return ldexp((double)o->x48, -48);
}

/* random_gauss, polar method with rejection (random.c:82-100). */
double oracle_random_gauss(oracle *o, double mu, double sigma)
{
    double U1, U2, W, mult, X1;
    do {
        U1 = -1 + oracle_drand48(o) * 2;
        U2 = -1 + oracle_drand48(o) * 2;
        W = pow(U1, 2) + pow(U2, 2);
    } while (W >= 1 || W == 0);
    mult = sqrt((-2 * log(W)) / W);
    X1 = U1 * mult;
    return (mu + sigma * X1);
}

/* ------------------------------------------------------------------------ */
/* design_lb, D3Q19 block (lb.c:139-230) and sound-speed constants computed
 * with sqrt/pow as the reference does (lb.c:368-377).                       */
static void design_lb(oracle *o)
{
    static const int C[NP][3] = {
        {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
        {1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},
        {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
    for (int q = 0; q < NP; q++) {
        o->cx[q] = C[q][0];
        o->cy[q] = C[q][1];
        o->cz[q] = C[q][2];
        o->wgt[q] = (q == 0) ? 1. / 3. : (q <= 6 ? 1. / 18. : 1. / 36.);
    }
    for (int q = 0; q < NP; q++)
        for (int r = 0; r < NP; r++)
            if (C[r][0] == -C[q][0] && C[r][1] == -C[q][1] && C[r][2] == -C[q][2]) o->inv[q] = r;
    o->cs = 1.0 / sqrt(3.0);
    o->cs2 = pow(o->cs, 2.0);
    o->cs4 = pow(o->cs, 4.0);
    o->invcs = 1.0 / o->cs;
    o->invcs2 = 1.0 / o->cs2;
    o->invcs4 = 1.0 / o->cs4;
    o->twocs2 = 2.0 * o->cs2;
    o->twocs4 = 2.0 * o->cs4;
    o->invtwocs2 = 1.0 / o->twocs2;
    o->invtwocs4 = 1.0 / o->twocs4;
}

/* Moment macros m, mvx, mvy, mvz (typedef.h:286-291): left-to-right sums. */
static inline double mom_m(const opop *a)
{
    const double *p = a->p;
    return p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8] + p[9] + p[10] + p[11] +
           p[12] + p[13] + p[14] + p[15] + p[16] + p[17] + p[18];
}
static inline double mom_x(const opop *a)
{
    const double *p = a->p;
    return p[1] - p[2] + p[7] + p[8] - p[9] - p[10] + p[11] - p[12] + p[13] - p[14];
}
static inline double mom_y(const opop *a)
{
    const double *p = a->p;
    return p[3] - p[4] + p[7] - p[8] + p[9] - p[10] + p[15] + p[16] - p[17] - p[18];
}
static inline double mom_z(const opop *a)
{
    const double *p = a->p;
    return p[5] - p[6] + p[11] + p[12] - p[13] - p[14] + p[15] - p[16] + p[17] - p[18];
}

/* equilibrium (lb.c:382-415): rho from the node's own populations, u from
 * the global fluid velocity field.                                          */
static opop equilibrium(const oracle *o, const opop *f, size_t n)
{
    opop e;
    double rhof = mom_m(&f[n]);
    double ux = o->u[n].x, uy = o->u[n].y, uz = o->u[n].z;
    double u2 = (ux * ux + uy * uy + uz * uz);
    for (int q = 0; q < NP; q++) {
        double cu = (o->cx[q] * ux + o->cy[q] * uy + o->cz[q] * uz);
        e.p[q] = rhof * o->wgt[q] * (1.0 + o->invcs2 * cu + o->invtwocs4 * cu * cu - o->invtwocs2 * u2);
    }
    return e;
}

/* ------------------------------------------------------------------------ */
/* sendrecv_borders_pop on one rank (lb.c:733-1058): every ghost node
 * (faces, edges, corners) receives the periodic image of the interior.     */
static void halo_pop(oracle *o, opop *f)
{
    const int GX = o->GX, GY = o->GY, GZ = o->GZ;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < GZ; k++)
        for (int j = 0; j < GY; j++)
            for (int i = 0; i < GX; i++) {
                int gi = (i == 0 || i == GX - 1), gj = (j == 0 || j == GY - 1), gk = (k == 0 || k == GZ - 1);
                if (!(gi || gj || gk)) continue;
                int si = i == 0 ? o->NX : (i == GX - 1 ? 1 : i);
                int sj = j == 0 ? o->NY : (j == GY - 1 ? 1 : j);
                int sk = k == 0 ? o->NZ : (k == GZ - 1 ? 1 : k);
                f[IDX(o, i, j, k)] = f[IDX(o, si, sj, sk)];
            }
}

/* sendrecv_borders_vector / _scalar on one rank (grid.c:609,893). */
static void halo_vec(oracle *o, ovec *f)
{
    const int GX = o->GX, GY = o->GY, GZ = o->GZ;
    for (int k = 0; k < GZ; k++)
        for (int j = 0; j < GY; j++)
            for (int i = 0; i < GX; i++) {
                int si = i == 0 ? o->NX : (i == GX - 1 ? 1 : i);
                int sj = j == 0 ? o->NY : (j == GY - 1 ? 1 : j);
                int sk = k == 0 ? o->NZ : (k == GZ - 1 ? 1 : k);
                if (si != i || sj != j || sk != k) f[IDX(o, i, j, k)] = f[IDX(o, si, sj, sk)];
            }
}
static void halo_scal(oracle *o, double *f)
{
    const int GX = o->GX, GY = o->GY, GZ = o->GZ;
    for (int k = 0; k < GZ; k++)
        for (int j = 0; j < GY; j++)
            for (int i = 0; i < GX; i++) {
                int si = i == 0 ? o->NX : (i == GX - 1 ? 1 : i);
                int sj = j == 0 ? o->NY : (j == GY - 1 ? 1 : j);
                int sk = k == 0 ? o->NZ : (k == GZ - 1 ? 1 : k);
                if (si != i || sj != j || sk != k) f[IDX(o, i, j, k)] = f[IDX(o, si, sj, sk)];
            }
}

/* ------------------------------------------------------------------------ */
/* hydro_fields (lb.c:420-540). 'p': rho = m(p), u = mv/rho + dt/2 F under
 * METHOD_FORCING_GUO with LB_FLUID_FORCING (:464-466).  'g': t = m(g) (:486). */
/* LB_FLUID_FORCING: any force term, or the particle feedback (which only
 * enters through build_forcing's force field, main.c:118-131)              */
static int has_fluid_forcing(const oracle *o)
{
    return HAS(o, CH4G_FLUID_FORCING_POISEUILLE | CH4G_FLUID_FORCING_CHANNEL | CH4G_FLUID_FORCING_HIT |
                      CH4G_TEMPERATURE_BUOYANCY) ||
           olagr_twoway(o);
}

static void hydro_fields(oracle *o, char which)
{
    const int forced = has_fluid_forcing(o);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++) {
                size_t n = IDX(o, i, j, k);
                if (which == 'p') {
                    o->dens[n] = mom_m(&o->p[n]);
                    if (forced) {
                        o->u[n].x = mom_x(&o->p[n]) / o->dens[n] + 0.5 * 1.0 * o->force[n].x;
                        o->u[n].y = mom_y(&o->p[n]) / o->dens[n] + 0.5 * 1.0 * o->force[n].y;
                        o->u[n].z = mom_z(&o->p[n]) / o->dens[n] + 0.5 * 1.0 * o->force[n].z;
                    } else {
                        o->u[n].x = mom_x(&o->p[n]) / o->dens[n];
                        o->u[n].y = mom_y(&o->p[n]) / o->dens[n];
                        o->u[n].z = mom_z(&o->p[n]) / o->dens[n];
                    }
                } else {
                    o->t[n] = mom_m(&o->g[n]);
                }
            }
}

/* ------------------------------------------------------------------------ */
/* Planted mistakes for the pinning tests (test infrastructure only). */
enum {
    PLANT_NONE = 0,
    PLANT_HIT_AX_IN_FX = 1,       /* sin(k_x x) in F_x instead of sin(k_y y) */
    PLANT_HIT_NO_REVERSION = 2,   /* phase walk without the -phi fac1 term */
    PLANT_HIT_EXPONENT = 3,       /* pow(k^2, -2/3) instead of pow(k^2, -1/3) */
    PLANT_BUOY_T_BOT = 4,         /* theta = t - T_bot instead of t - (T_bot + T_top)/2 */
    PLANT_BUOY_SIGN = 5,          /* -beta_t theta g */
    PLANT_CHANNEL_LX = 6,         /* channel F_x = 2 Amp_x^2 / LX */
    PLANT_HIT_NOISE = 7,          /* phase walk noise sqrt(time_dt phi_diffusivity) (factor 2 dropped) */
};
static int plant = PLANT_NONE;
void oracle_set_plant(int which) { plant = which; }

/* build_forcing (forcing.c:5-1246), branches used on the streaming step:
 * HIT phase Ornstein-Uhlenbeck walk on the root (:40-69), per node force
 * reset (:361-364), Poiseuille (:366-371), channel (:374-379), HIT default
 * large-scale forcing (:576-590), buoyancy with the T0 = T_mean default or
 * T_ref branch (:620-667), NOZEROMODE mean removal (:1123-1152).             */
static void build_forcing(oracle *o)
{
    const double LX = o->NX, LY = o->NY, LZ = o->NZ;
    const double nu = o->nu;
    if (HAS(o, CH4G_FLUID_FORCING_HIT)) {
        double t_turnover = (LX / 0.1);
        double phi_diffusivity = 0.01 / t_turnover;
        double fac1 = 1.0 / t_turnover;
        double fac2 = sqrt(2.0 * 1.0 * phi_diffusivity);
        if (plant == PLANT_HIT_NO_REVERSION) fac1 = 0.0;
        if (plant == PLANT_HIT_NOISE) fac2 = sqrt(1.0 * phi_diffusivity);
        for (int ii = 0; ii < o->nk; ii++) {
            double val = oracle_random_gauss(o, 0.0, 1.0);
            o->phi[ii].x += -o->phi[ii].x * fac1 + val * fac2;
            val = oracle_random_gauss(o, 0.0, 1.0);
            o->phi[ii].y += -o->phi[ii].y * fac1 + val * fac2;
            val = oracle_random_gauss(o, 0.0, 1.0);
            o->phi[ii].z += -o->phi[ii].z * fac1 + val * fac2;
        }
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++) {
                size_t n = IDX(o, i, j, k);
                /* cell centres (grid.c:183-185, :385-387): i - BRD + 1/2 */
                double x = (double)(i - BRD) + 0.5, y = (double)(j - BRD) + 0.5, z = (double)(k - BRD) + 0.5;
                ovec F = {0.0, 0.0, 0.0};
                if (HAS(o, CH4G_FLUID_FORCING_POISEUILLE)) {
                    F.x += 2.0 * o->cfg.Amp_x * ((4. * nu) * pow(LY, -2.0));
                    F.y += 2.0 * o->cfg.Amp_y * ((4. * nu) * pow(LX, -2.0));
                    F.z += 2.0 * o->cfg.Amp_z * ((4. * nu) * pow(LY, -2.0));
                }
                if (HAS(o, CH4G_FLUID_FORCING_CHANNEL)) {
                    F.x += 2.0 * pow(o->cfg.Amp_x, 2.0) / (plant == PLANT_CHANNEL_LX ? LX : LY);
                    F.y += 2.0 * pow(o->cfg.Amp_y, 2.0) / LX;
                    F.z += 2.0 * pow(o->cfg.Amp_z, 2.0) / LY;
                }
                if (HAS(o, CH4G_FLUID_FORCING_HIT)) {
                    for (int ii = 0; ii < o->nk; ii++) {
                        double fac = pow(o->vk2[ii], plant == PLANT_HIT_EXPONENT ? -4. / 6. : -2. / 6.);
                        /* two_pi is `2*3.14159265359` (typedef.h:367) */
                        double ax = sin(2 * 3.14159265359 * (o->vk[ii].x * x / LX + o->phi[ii].x));
                        double ay = sin(2 * 3.14159265359 * (o->vk[ii].y * y / LY + o->phi[ii].y));
                        double az = sin(2 * 3.14159265359 * (o->vk[ii].z * z / LZ + o->phi[ii].z));
                        F.x += fac * o->cfg.Amp_x * ((plant == PLANT_HIT_AX_IN_FX ? ax : ay) + az);
                        F.y += fac * o->cfg.Amp_y * (ax + az);
                        F.z += fac * o->cfg.Amp_z * (ay + ax);
                    }
                }
                if (HAS(o, CH4G_TEMPERATURE_BUOYANCY)) {
                    double temp;
                    if (HAS(o, CH4G_TEMPERATURE_BUOYANCY_T0_REF))
                        temp = (o->t[n] - o->cfg.T_ref);
                    else if (plant == PLANT_BUOY_T_BOT)
                        temp = o->t[n] - o->cfg.T_bot;
                    else
                        temp = o->t[n] - 0.5 * (o->cfg.T_bot + o->cfg.T_top);
                    double fac = (plant == PLANT_BUOY_SIGN ? -o->cfg.beta_t : o->cfg.beta_t) * temp;
                    if (o->cfg.beta2_t != 0.0) fac += o->cfg.beta2_t * temp * temp;
                    F.x += fac * o->cfg.gravity_x;
                    F.y += fac * o->cfg.gravity_y;
                    F.z += fac * o->cfg.gravity_z;
                }
                o->force[n] = F;
            }
    if (HAS(o, CH4G_FLUID_FORCING_NOZEROMODE)) {
        ovec f0 = {0.0, 0.0, 0.0};
        for (int k = BRD; k < o->NZ + BRD; k++)
            for (int j = BRD; j < o->NY + BRD; j++)
                for (int i = BRD; i < o->NX + BRD; i++) {
                    size_t n = IDX(o, i, j, k);
                    f0.x += o->force[n].x;
                    f0.y += o->force[n].y;
                    f0.z += o->force[n].z;
                }
        double norm = 1.0 / (LX * LY * LZ);
        f0.x *= norm;
        f0.y *= norm;
        f0.z *= norm;
        for (int k = BRD; k < o->NZ + BRD; k++)
            for (int j = BRD; j < o->NY + BRD; j++)
                for (int i = BRD; i < o->NX + BRD; i++) {
                    size_t n = IDX(o, i, j, k);
                    o->force[n].x -= f0.x;
                    o->force[n].y -= f0.y;
                    o->force[n].z -= f0.z;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* stress_tensor (gradients.c:42-69) including its S.zz = S.xz, S.zx = 0
 * assignment, used by the local-strain Smagorinsky closure.                */
static oten stress_tensor(const oracle *o, const opop *f, size_t n)
{
    oten S;
    memset(&S, 0, sizeof S);
    opop e = equilibrium(o, f, n);
    for (int q = 0; q < NP; q++) {
        double d = f[n].p[q] - e.p[q];
        S.xx += o->cx[q] * o->cx[q] * d;
        S.yy += o->cy[q] * o->cy[q] * d;
        S.zz += o->cz[q] * o->cz[q] * d;
        S.xy += o->cx[q] * o->cy[q] * d;
        S.xz += o->cx[q] * o->cz[q] * d;
        S.yz += o->cy[q] * o->cz[q] * d;
    }
    S.yx = S.xy;
    S.zz = S.xz;
    S.zy = S.yz;
    return S;
}

/* tau_u_les, LB_FLUID_LES_LOCAL_STRAIN branch (les.c:20-30, :144-145,
 * :160-180 nu_les = |S| (C vol)^2, :253 tau = 3(nu+nu_t)+1/2, :266).  The
 * Smagorinsky constant is cfg.les_cs (the reference hard-codes 0.18).      */
static double tau_u_les(oracle *o, size_t n)
{
    oten QQ = stress_tensor(o, o->p, n);
    if (o->itime == 1) o->tau_les[n] = o->cfg.tau_u;
    double fac = -1.0 / (2 * o->cs2 * o->dens[n] * o->tau_les[n] * 1.0);
    double s_norm = fabs(fac) * sqrt(2.0 * (QQ.xx * QQ.xx + QQ.xy * QQ.xy + QQ.xz * QQ.xz + QQ.yx * QQ.yx +
                                            QQ.yy * QQ.yy + QQ.yz * QQ.yz + QQ.zx * QQ.zx + QQ.zy * QQ.zy +
                                            QQ.zz * QQ.zz));
    double vol = 1.0;
    double nu_les = s_norm * pow(o->cfg.les_cs * vol, 2.0);
    double tau = 3.0 * (nu_les + o->nu) + 0.5;
    o->tau_les[n] = tau;
    return tau;
}

static oten gradient_vector(const oracle *o, const ovec *t, int i, int j, int k);

/* tau_u_les, default finite-difference branch (les.c:31-43): |S| from the
 * gradient_vector of u (centred, one-sided second order at y walls), then
 * the same eddy viscosity and relaxation time as above.                    */
static double tau_u_les_fd(oracle *o, int i, int j, int k)
{
    oten g = gradient_vector(o, o->u, i, j, k);
    double s_norm = sqrt(0.5 * ((g.xx + g.xx) * (g.xx + g.xx) + (g.xy + g.yx) * (g.xy + g.yx) +
                                (g.xz + g.zx) * (g.xz + g.zx) + (g.yx + g.xy) * (g.yx + g.xy) +
                                (g.yy + g.yy) * (g.yy + g.yy) + (g.yz + g.zy) * (g.yz + g.zy) +
                                (g.zx + g.xz) * (g.zx + g.xz) + (g.zy + g.yz) * (g.zy + g.yz) +
                                (g.zz + g.zz) * (g.zz + g.zz)));
    double vol = 1.0;
    double nu_les = s_norm * pow(o->cfg.les_cs * vol, 2.0);
    double tau = 3.0 * (nu_les + o->nu) + 0.5;
    o->tau_les[IDX(o, i, j, k)] = tau;
    return tau;
}

/* tau_t_les (les.c:273-297, METHOD_STREAMING branch, prandtl forced to 1):
 * the temperature relaxation time from this step's tau_u_les_total.       */
static double tau_t_les(const oracle *o, size_t n)
{
    double prandtl = o->nu / o->kappa;
    prandtl = 1.0;
    double tau_les = o->tau_les[n];
    return 3.0 * (o->kappa + (tau_les - 0.5 - 3 * o->nu) / (3 * prandtl)) + 0.5;
}

/* compute_advection, streaming branch with dt = 1 (fluid.c:267-284):
 * rhs <- f, followed by add_collision SRT (fluid.c:1033-1225):
 * rhs += -(1/tau)(f - f_eq), with the per-node LES relaxation times of
 * fluid.c:1070-1088 (u halo first, fluid.c:1056-1058).                     */
static void advection_collision(oracle *o, const opop *f, opop *rhs, double tau, char which)
{
    const int les = (which == 'p') && HAS(o, CH4G_FLUID_LES);
    const int les_fd = les && !HAS(o, CH4G_FLUID_LES_LOCAL_STRAIN);
    const int les_t = (which == 'g') && HAS(o, CH4G_FLUID_LES) && HAS(o, CH4G_TEMPERATURE_LES);
    if (HAS(o, CH4G_FLUID_LES)) halo_vec(o, o->u);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++) {
                size_t n = IDX(o, i, j, k);
                rhs[n] = f[n];
                double invtau = 1.0 / tau;
                if (les) invtau = 1.0 / (les_fd ? tau_u_les_fd(o, i, j, k) : tau_u_les(o, n));
                if (les_t) invtau = 1.0 / tau_t_les(o, n);
                opop e = equilibrium(o, f, n);
                for (int q = 0; q < NP; q++) {
                    double fcoll = -invtau * (f[n].p[q] - e.p[q]);
                    rhs[n].p[q] += fcoll;
                }
            }
}

/* add_forcing, METHOD_FORCING_GUO (forcing.c:1250-1529, Guo term :1424-1435)
 * with the molecular tau_u in the prefactor (the LES override at :1414 is
 * never compiled).  No temperature source term on these configurations.  */
static void add_forcing(oracle *o)
{
    const double invtau = 1.0 / o->cfg.tau_u;
    const double fac = (1.0 - 0.5 * 1.0 * invtau);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++) {
                size_t n = IDX(o, i, j, k);
                for (int q = 0; q < NP; q++) {
                    double rho = o->dens[n];
                    double ux = o->u[n].x, uy = o->u[n].y, uz = o->u[n].z;
                    double cu = (o->cx[q] * ux + o->cy[q] * uy + o->cz[q] * uz);
                    double dx = (o->cx[q] - ux) * o->invcs2 + o->cx[q] * cu * o->invcs4;
                    double dy = (o->cy[q] - uy) * o->invcs2 + o->cy[q] * cu * o->invcs4;
                    double dz = (o->cz[q] - uz) * o->invcs2 + o->cz[q] * cu * o->invcs4;
                    o->rhs_p[n].p[q] += fac * o->wgt[q] * rho * o->force[n].x * dx;
                    o->rhs_p[n].p[q] += fac * o->wgt[q] * rho * o->force[n].y * dy;
                    o->rhs_p[n].p[q] += fac * o->wgt[q] * rho * o->force[n].z * dz;
                }
            }
}

/* boundary_and_pbc_conditions_for_streaming (boundary_conditions.c:419-1524):
 * periodic ghosts (:434-445), fluid no-slip at y walls for every i,k
 * including ghost columns (:714-717, :775, :798), temperature fixed at y
 * walls on interior i,k using the node's own same-direction population
 * (:966-1078, :1026, :1072).                                                */
static void boundary_conditions_streaming(oracle *o)
{
    const int GX = o->GX, GZ = o->GZ;
    if (HAS(o, CH4G_LB_FLUID)) halo_pop(o, o->rhs_p);
    if (HAS(o, CH4G_LB_TEMPERATURE)) halo_pop(o, o->rhs_g);

    if (HAS(o, CH4G_FLUID_BC_Y)) {
        for (int i = 0; i < GX; i++)
            for (int k = 0; k < GZ; k++)
                for (int q = 0; q < NP; q++) {
                    int ii = i + (int)o->cx[q];
                    int kk = k + (int)o->cz[q];
                    int j = o->NY + BRD - 1;
                    if (ii >= 0 && ii < GX && kk >= 0 && kk < GZ)
                        o->rhs_p[IDX(o, ii, j + 1, kk)].p[o->inv[q]] = o->rhs_p[IDX(o, i, j, k)].p[q];
                    j = BRD;
                    if (ii >= 0 && ii < GX && kk >= 0 && kk < GZ)
                        o->rhs_p[IDX(o, ii, j - 1, kk)].p[o->inv[q]] = o->rhs_p[IDX(o, i, j, k)].p[q];
                }
    }
    if (HAS(o, CH4G_TEMPERATURE_BC_Y)) {
        for (int i = BRD; i < o->NX + BRD; i++)
            for (int k = BRD; k < o->NZ + BRD; k++) {
                int j = o->NY + BRD - 1;
                double T_wall = o->cfg.T_top;
                double fac = 2.0 * ((T_wall - o->cfg.T_ref) - o->t[IDX(o, i, j, k)]);
                for (int q = 0; q < NP; q++) {
                    if (o->cy[q] > 0) {
                        int ii = i + (int)o->cx[q], kk = k + (int)o->cz[q];
                        int iq = o->inv[q];
                        o->rhs_g[IDX(o, ii, j + 1, kk)].p[iq] = o->rhs_g[IDX(o, i, j, k)].p[iq] + o->wgt[iq] * fac;
                    }
                }
                j = BRD;
                T_wall = o->cfg.T_bot;
                fac = 2.0 * ((T_wall - o->cfg.T_ref) - o->t[IDX(o, i, j, k)]);
                for (int q = 0; q < NP; q++) {
                    if (o->cy[q] < 0) {
                        int ii = i + (int)o->cx[q], kk = k + (int)o->cz[q];
                        int iq = o->inv[q];
                        o->rhs_g[IDX(o, ii, j - 1, kk)].p[iq] = o->rhs_g[IDX(o, i, j, k)].p[iq] + o->wgt[iq] * fac;
                    }
                }
            }
    }
}

/* time_stepping + streaming, pull with dt = 1 (lb.c:544-583, :588-695). */
static void time_stepping(oracle *o, opop *f, const opop *rhs)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++) {
                size_t n = IDX(o, i, j, k);
                for (int q = 0; q < NP; q++) {
                    int ii = i - (int)o->cx[q], jj = j - (int)o->cy[q], kk = k - (int)o->cz[q];
                    f[n].p[q] = rhs[IDX(o, ii, jj, kk)].p[q];
                }
            }
}

/* ------------------------------------------------------------------------ */
/* One time step of main.c:92-161. */
static void one_step(oracle *o)
{
    o->itime++;
    if (HAS(o, CH4G_LB_FLUID)) advection_collision(o, o->p, o->rhs_p, o->cfg.tau_u, 'p');
    if (HAS(o, CH4G_LB_TEMPERATURE)) advection_collision(o, o->g, o->rhs_g, o->cfg.tau_t, 'g');
    if (has_fluid_forcing(o)) {
        build_forcing(o);
        /* main.c:125-130 without restart: from the second step on */
        if (olagr_twoway(o) && o->itime > 1) olagr_feedback(o);
        add_forcing(o);
    }
    boundary_conditions_streaming(o);
    if (HAS(o, CH4G_LB_FLUID)) time_stepping(o, o->p, o->rhs_p);
    if (HAS(o, CH4G_LB_TEMPERATURE)) time_stepping(o, o->g, o->rhs_g);
    if (HAS(o, CH4G_LB_FLUID)) hydro_fields(o, 'p');
    if (HAS(o, CH4G_LB_TEMPERATURE)) hydro_fields(o, 'g');
    /* dump_averages (main.c:162) reads only the fields; then main.c:163-183 */
    if (o->lagr) olagr_after_step(o);
}

/* First reconstruction of main.c:26-48. */
static void first_hydro(oracle *o)
{
    if (HAS(o, CH4G_LB_FLUID)) hydro_fields(o, 'p');
    if (HAS(o, CH4G_LB_TEMPERATURE)) hydro_fields(o, 'g');
    if (has_fluid_forcing(o)) {
        build_forcing(o);
        if (HAS(o, CH4G_LB_FLUID)) hydro_fields(o, 'p');
        if (HAS(o, CH4G_LB_TEMPERATURE)) hydro_fields(o, 'g');
    }
}

/* ------------------------------------------------------------------------ */
/* Public API (ctypes, tests and bench.py CPU leg only). */
oracle *oracle_create(const ch4g_config *cfg)
{
    if (!cfg || cfg->NX < 1 || cfg->NY < 1 || cfg->NZ < 1) return NULL;
    oracle *o = (oracle *)calloc(1, sizeof(oracle));
    if (!o) return NULL;
    o->cfg = *cfg;
    o->NX = cfg->NX;
    o->NY = cfg->NY;
    o->NZ = cfg->NZ;
    o->GX = o->NX + 2 * BRD;
    o->GY = o->NY + 2 * BRD;
    o->GZ = o->NZ + 2 * BRD;
    o->ng = (size_t)o->GX * o->GY * o->GZ;
    o->p = (opop *)calloc(o->ng, sizeof(opop));
    o->rhs_p = (opop *)calloc(o->ng, sizeof(opop));
    o->g = (opop *)calloc(o->ng, sizeof(opop));
    o->rhs_g = (opop *)calloc(o->ng, sizeof(opop));
    o->u = (ovec *)calloc(o->ng, sizeof(ovec));
    o->force = (ovec *)calloc(o->ng, sizeof(ovec));
    o->dens = (double *)calloc(o->ng, sizeof(double));
    o->t = (double *)calloc(o->ng, sizeof(double));
    o->tau_les = (double *)calloc(o->ng, sizeof(double));
    if (!o->p || !o->rhs_p || !o->g || !o->rhs_g || !o->u || !o->force || !o->dens || !o->t || !o->tau_les) {
        free(o);
        return NULL;
    }
    design_lb(o);
    /* parameters.c:179 and :285 under METHOD_STREAMING */
    o->nu = (cfg->tau_u - 0.5 * 1.0) / 3.0;
    o->kappa = (cfg->tau_t - 0.5 * 1.0) / 3.0;
    o->deltaT = cfg->T_bot - cfg->T_top;
    oracle_srand48(o, cfg->seed);
    /* HIT modes |k|^2 <= 2 (initial_conditions.c:838-846) */
    o->nk = CH4G_HIT_NK;
    const int VK[CH4G_HIT_NK][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
    for (int ii = 0; ii < o->nk; ii++) {
        o->vk[ii].x = VK[ii][0];
        o->vk[ii].y = VK[ii][1];
        o->vk[ii].z = VK[ii][2];
        o->vk2[ii] = o->vk[ii].x * o->vk[ii].x + o->vk[ii].y * o->vk[ii].y + o->vk[ii].z * o->vk[ii].z;
    }
    return o;
}

void oracle_destroy(oracle *o)
{
    if (!o) return;
    free(o->p);
    free(o->rhs_p);
    free(o->g);
    free(o->rhs_g);
    free(o->u);
    free(o->force);
    free(o->dens);
    free(o->t);
    free(o->tau_les);
    olagr_free(o);
    free(o);
}

/* initial_conditions (initial_conditions.c:3-492) for the switches of
 * ch4g.h, then initialization_forcing (:822-875), then main.c:26-48.       */
int oracle_initial_conditions(oracle *o)
{
    const double LY = o->NY;
    if (HAS(o, CH4G_LB_FLUID)) {
        for (int k = BRD; k < o->NZ + BRD; k++)
            for (int j = BRD; j < o->NY + BRD; j++)
                for (int i = BRD; i < o->NX + BRD; i++) {
                    size_t n = IDX(o, i, j, k);
                    for (int q = 0; q < NP; q++) o->p[n].p[q] = o->wgt[q];
                    if (HAS(o, CH4G_TEMPERATURE_BUOYANCY) && HAS(o, CH4G_INITIAL_BAROMETRIC)) {
                        double L = LY, y = (double)(j - BRD) + 0.5;
                        for (int q = 0; q < NP; q++)
                            o->p[n].p[q] = o->wgt[q] * (exp(o->cfg.beta_t * o->cfg.gravity_y * y *
                                                            (0.5 * o->deltaT - 0.5 * (o->deltaT / L) * y) / o->cs2));
                    }
                }
        if (HAS(o, CH4G_FLUID_INITIAL_UNIT_DENSITY)) {
            double norm = 0.0;
            for (int k = BRD; k < o->NZ + BRD; k++)
                for (int j = BRD; j < o->NY + BRD; j++)
                    for (int i = BRD; i < o->NX + BRD; i++)
                        for (int q = 0; q < NP; q++) norm += o->p[IDX(o, i, j, k)].p[q];
            norm /= (double)o->NX * (double)o->NY * (double)o->NZ;
            for (int k = BRD; k < o->NZ + BRD; k++)
                for (int j = BRD; j < o->NY + BRD; j++)
                    for (int i = BRD; i < o->NX + BRD; i++)
                        for (int q = 0; q < NP; q++) o->p[IDX(o, i, j, k)].p[q] /= norm;
        }
    }
    if (HAS(o, CH4G_LB_TEMPERATURE)) {
        for (int k = BRD; k < o->NZ + BRD; k++)
            for (int j = BRD; j < o->NY + BRD; j++)
                for (int i = BRD; i < o->NX + BRD; i++) {
                    size_t n = IDX(o, i, j, k);
                    o->t[n] = 0.0;
                    if (HAS(o, CH4G_TEMPERATURE_INITIAL_LINEAR)) {
                        double L = LY, y = (double)(j - BRD) + 0.5;
                        o->t[n] = ((o->cfg.T_bot - o->cfg.T_ref) - (o->deltaT / L) * y);
                    }
                    if (HAS(o, CH4G_TEMPERATURE_INITIAL_ADD_NOISE)) {
                        double val = o->cfg.noise_t;
                        o->t[n] += val * (2.0 * oracle_drand48(o) - 1.0);
                    }
                    for (int q = 0; q < NP; q++) o->g[n].p[q] = o->wgt[q] * o->t[n];
                }
    }
    if (HAS(o, CH4G_FLUID_FORCING_HIT)) {
        for (int ii = 0; ii < o->nk; ii++) {
            o->phi[ii].x = oracle_drand48(o);
            o->phi[ii].y = oracle_drand48(o);
            o->phi[ii].z = oracle_drand48(o);
        }
    }
    o->itime = 0;
    first_hydro(o);
    return 0;
}

int oracle_set_populations(oracle *o, const double *p, const double *g)
{
    size_t m = 0;
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++, m++) {
                size_t n = IDX(o, i, j, k);
                if (p) memcpy(o->p[n].p, p + m * NP, sizeof(double) * NP);
                if (g) memcpy(o->g[n].p, g + m * NP, sizeof(double) * NP);
            }
    first_hydro(o);
    return 0;
}

int oracle_get_populations(const oracle *o, double *p, double *g)
{
    size_t m = 0;
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++, m++) {
                size_t n = IDX(o, i, j, k);
                if (p) memcpy(p + m * NP, o->p[n].p, sizeof(double) * NP);
                if (g) memcpy(g + m * NP, o->g[n].p, sizeof(double) * NP);
            }
    return 0;
}

int oracle_step(oracle *o, int nsteps)
{
    for (int s = 0; s < nsteps; s++) one_step(o);
    return 0;
}

long oracle_itime(const oracle *o) { return o->itime; }

int oracle_hydro_fields(const oracle *o, double *dens, double *u, double *t, double *force)
{
    size_t m = 0;
    for (int k = BRD; k < o->NZ + BRD; k++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int i = BRD; i < o->NX + BRD; i++, m++) {
                size_t n = IDX(o, i, j, k);
                if (dens) dens[m] = o->dens[n];
                if (u) {
                    u[3 * m] = o->u[n].x;
                    u[3 * m + 1] = o->u[n].y;
                    u[3 * m + 2] = o->u[n].z;
                }
                if (t) t[m] = o->t[n];
                if (force) {
                    force[3 * m] = o->force[n].x;
                    force[3 * m + 1] = o->force[n].y;
                    force[3 * m + 2] = o->force[n].z;
                }
            }
    return 0;
}

int oracle_get_hit_phases(const oracle *o, double *phi)
{
    for (int ii = 0; ii < o->nk; ii++) {
        phi[3 * ii] = o->phi[ii].x;
        phi[3 * ii + 1] = o->phi[ii].y;
        phi[3 * ii + 2] = o->phi[ii].z;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* gradient_vector / gradient_scalar (gradients.c:78-499, :505-~840), no
 * LAGRANGE: centred differences on the cell-centre spacing, one-sided
 * second order at y walls (SECOND_ORDER_GRAD, gradients.c:74), components
 * along a size-one axis zeroed for the vector (:494-496).                  */
static oten gradient_vector(const oracle *o, const ovec *t, int i, int j, int k)
{
    oten g;
    size_t xp = IDX(o, i + 1, j, k), xm = IDX(o, i - 1, j, k);
    size_t yp = IDX(o, i, j + 1, k), ym = IDX(o, i, j - 1, k);
    size_t zp = IDX(o, i, j, k + 1), zm = IDX(o, i, j, k - 1);
    const double h2 = 2.0; /* centre(i+1) - centre(i-1) on the unit grid */
    g.xx = (t[xp].x - t[xm].x) / h2;
    g.yx = (t[xp].y - t[xm].y) / h2;
    g.zx = (t[xp].z - t[xm].z) / h2;
    g.xy = (t[yp].x - t[ym].x) / h2;
    g.yy = (t[yp].y - t[ym].y) / h2;
    g.zy = (t[yp].z - t[ym].z) / h2;
    g.xz = (t[zp].x - t[zm].x) / h2;
    g.yz = (t[zp].y - t[zm].y) / h2;
    g.zz = (t[zp].z - t[zm].z) / h2;
    if (HAS(o, CH4G_FLUID_BC_Y)) {
        /* h1 = h2 = 1: a0 = -3/2, a1 = 2, a2 = -1/2 */
        const double h1 = 1.0, hh = 1.0;
        if (j == BRD) {
            double a0 = -(2.0 * h1 + hh) / (h1 * (h1 + hh)), a1 = (h1 + hh) / (h1 * hh), a2 = -(h1) / (hh * (h1 + hh));
            size_t n0 = IDX(o, i, j, k), n1 = IDX(o, i, j + 1, k), n2 = IDX(o, i, j + 2, k);
            g.xy = a0 * t[n0].x + a1 * t[n1].x + a2 * t[n2].x;
            g.yy = a0 * t[n0].y + a1 * t[n1].y + a2 * t[n2].y;
            g.zy = a0 * t[n0].z + a1 * t[n1].z + a2 * t[n2].z;
        }
        if (j == o->NY + BRD - 1) {
            double a0 = (2.0 * h1 + hh) / (h1 * (h1 + hh)), a1 = -(h1 + hh) / (h1 * hh), a2 = (h1) / (hh * (h1 + hh));
            size_t n0 = IDX(o, i, j, k), n1 = IDX(o, i, j - 1, k), n2 = IDX(o, i, j - 2, k);
            g.xy = a0 * t[n0].x + a1 * t[n1].x + a2 * t[n2].x;
            g.yy = a0 * t[n0].y + a1 * t[n1].y + a2 * t[n2].y;
            g.zy = a0 * t[n0].z + a1 * t[n1].z + a2 * t[n2].z;
        }
    }
    if (o->NX == 1) g.xx = g.yx = g.zx = 0.0;
    if (o->NY == 1) g.xy = g.yy = g.zy = 0.0;
    if (o->NZ == 1) g.xz = g.yz = g.zz = 0.0;
    return g;
}

static ovec gradient_scalar(const oracle *o, const double *t, int i, int j, int k)
{
    ovec g;
    const double h2 = 2.0;
    g.x = (t[IDX(o, i + 1, j, k)] - t[IDX(o, i - 1, j, k)]) / h2;
    g.y = (t[IDX(o, i, j + 1, k)] - t[IDX(o, i, j - 1, k)]) / h2;
    g.z = (t[IDX(o, i, j, k + 1)] - t[IDX(o, i, j, k - 1)]) / h2;
    if (HAS(o, CH4G_TEMPERATURE_BC_Y)) {
        const double h1 = 1.0, hh = 1.0;
        if (j == BRD) {
            double a0 = -(2.0 * h1 + hh) / (h1 * (h1 + hh)), a1 = (h1 + hh) / (h1 * hh), a2 = -(h1) / (hh * (h1 + hh));
            g.y = a0 * t[IDX(o, i, j, k)] + a1 * t[IDX(o, i, j + 1, k)] + a2 * t[IDX(o, i, j + 2, k)];
        }
        if (j == o->NY + BRD - 1) {
            double a0 = (2.0 * h1 + hh) / (h1 * (h1 + hh)), a1 = -(h1 + hh) / (h1 * hh), a2 = (h1) / (hh * (h1 + hh));
            g.y = a0 * t[IDX(o, i, j, k)] + a1 * t[IDX(o, i, j - 1, k)] + a2 * t[IDX(o, i, j - 2, k)];
        }
    }
    return g;
}

/* Per-node terms of dump_averages (output.c:183-491) at interior node
 * (i,j,k), vol = 1: coordinates, u, u^2, rho, rho^2, kinetic energy,
 * dissipation from gradient_vector, temperature, its gradient, u T and the
 * Nusselt integrands -kappa dT + u T.  Fields must have their halos.      */
static void node_terms(const oracle *o, int i, int j, int k, ch4g_output *v)
{
    memset(v, 0, sizeof *v);
    const int fluid = HAS(o, CH4G_LB_FLUID), temp = HAS(o, CH4G_LB_TEMPERATURE);
    size_t n = IDX(o, i, j, k);
    double vol = 1.0;
    if (fluid) {
        v->x = (double)(i - BRD) + 0.5;
        v->y = (double)(j - BRD) + 0.5;
        v->z = (double)(k - BRD) + 0.5;
        v->ux = o->u[n].x * vol;
        v->uy = o->u[n].y * vol;
        v->uz = o->u[n].z * vol;
        v->ux2 = o->u[n].x * o->u[n].x * vol;
        v->uy2 = o->u[n].y * o->u[n].y * vol;
        v->uz2 = o->u[n].z * o->u[n].z * vol;
        v->rho = o->dens[n] * vol;
        v->rho2 = o->dens[n] * o->dens[n] * vol;
        v->ene = 0.5 * (o->u[n].x * o->u[n].x + o->u[n].y * o->u[n].y + o->u[n].z * o->u[n].z) * vol;
        oten G = gradient_vector(o, o->u, i, j, k);
        v->eps = ((G.xx + G.xx) * (G.xx + G.xx) + (G.xy + G.yx) * (G.xy + G.yx) + (G.xz + G.zx) * (G.xz + G.zx) +
                  (G.yx + G.xy) * (G.yx + G.xy) + (G.yy + G.yy) * (G.yy + G.yy) + (G.yz + G.zy) * (G.yz + G.zy) +
                  (G.zx + G.xz) * (G.zx + G.xz) + (G.zy + G.yz) * (G.zy + G.yz) + (G.zz + G.zz) * (G.zz + G.zz)) *
                 0.5 * o->nu * vol;
    }
    if (temp) {
        ovec gt = gradient_scalar(o, o->t, i, j, k);
        v->dxt = gt.x * vol;
        v->dyt = gt.y * vol;
        v->dzt = gt.z * vol;
        v->uxt = o->u[n].x * o->t[n] * vol;
        v->uyt = o->u[n].y * o->t[n] * vol;
        v->uzt = o->u[n].z * o->t[n] * vol;
        v->nux = (-o->kappa * v->dxt + v->uxt);
        v->nuy = (-o->kappa * v->dyt + v->uyt);
        v->nuz = (-o->kappa * v->dzt + v->uzt);
        v->t = o->t[n] * vol;
        v->t2 = o->t[n] * o->t[n] * vol;
        v->epst = o->kappa * (gt.x * gt.x + gt.y * gt.y + gt.z * gt.z) * vol;
    }
}

/* r += v, member by member (the reference's += per field) */
static void add_terms(ch4g_output *r, const ch4g_output *v)
{
    double *a = (double *)r;
    const double *b = (const double *)v;
    for (size_t m = 0; m < sizeof *r / sizeof(double); m++) a[m] += b[m];
}

/* normalisation of output.c:499-690: fluid members by fnorm, temperature
 * members by tnorm, Nusselt numbers by tnorm / (kappa deltaT / SY), lb from
 * the normalised eps and epst with buoyancy.                               */
static void normalise(const oracle *o, ch4g_output *a, double fnorm, double tnorm)
{
    if (HAS(o, CH4G_LB_FLUID)) {
        a->ene *= fnorm;
        a->rho *= fnorm;
        a->ux *= fnorm;
        a->uy *= fnorm;
        a->uz *= fnorm;
        a->ux2 *= fnorm;
        a->uy2 *= fnorm;
        a->uz2 *= fnorm;
        a->eps *= fnorm;
        a->rho2 *= fnorm;
    }
    if (HAS(o, CH4G_LB_TEMPERATURE)) {
        double nn = tnorm / (o->kappa * o->deltaT / (double)o->NY);
        a->t *= tnorm;
        a->t2 *= tnorm;
        a->epst *= tnorm;
        a->dxt *= tnorm;
        a->dyt *= tnorm;
        a->dzt *= tnorm;
        a->uxt *= tnorm;
        a->uyt *= tnorm;
        a->uzt *= tnorm;
        a->nux *= nn;
        a->nuy *= nn;
        a->nuz *= nn;
        a->lb = 0.0;
        if (HAS(o, CH4G_TEMPERATURE_BUOYANCY))
            a->lb = pow(a->eps, 5. / 4.) * pow(a->epst, -3. / 4.) * pow(o->cfg.gravity_y * o->cfg.beta_t, -3. / 2.);
    }
}

/* dump_averages (output.c:63-690): FP64 sums over interior nodes in i,j,k
 * loop order of the fields of the current step; global averages normalised
 * by SX*SY*SZ, Nusselt numbers by kappa*deltaT/SY (:577-579); the x, y, z
 * rulers (plane sums, :212-491) when rx/ry/rz are given, normalised by the
 * plane sizes with the reference's factors: fluid 1/(SY SZ), 1/(SX SZ),
 * 1/(SX SY); temperature 1/(SY SZ), 1/(SX SZ) and, for z, 1/(SY SZ)
 * (:585-650); the ruler's own coordinate by 1/(NY NZ) etc. and the two
 * other coordinates left 0.                                               */
int oracle_dump_rulers(oracle *o, ch4g_output *out, ch4g_output *rx, ch4g_output *ry, ch4g_output *rz)
{
    ch4g_output a, v;
    memset(&a, 0, sizeof a);
    const int fluid = HAS(o, CH4G_LB_FLUID), temp = HAS(o, CH4G_LB_TEMPERATURE);
    if (rx) memset(rx, 0, sizeof(ch4g_output) * (size_t)o->NX);
    if (ry) memset(ry, 0, sizeof(ch4g_output) * (size_t)o->NY);
    if (rz) memset(rz, 0, sizeof(ch4g_output) * (size_t)o->NZ);
    if (fluid) halo_vec(o, o->u);
    if (temp) halo_scal(o, o->t);
    for (int i = BRD; i < o->NX + BRD; i++)
        for (int j = BRD; j < o->NY + BRD; j++)
            for (int k = BRD; k < o->NZ + BRD; k++) {
                node_terms(o, i, j, k, &v);
                add_terms(&a, &v);
                if (rx) add_terms(&rx[i - BRD], &v);
                if (ry) add_terms(&ry[j - BRD], &v);
                if (rz) add_terms(&rz[k - BRD], &v);
            }
    const double SX = o->NX, SY = o->NY, SZ = o->NZ;
    normalise(o, &a, 1.0 / (SX * SY * SZ), 1.0 / (SX * SY * SZ));
    if (!temp) a.lb = 0.0;
    if (out) *out = a;
    for (int i = 0; rx && i < o->NX; i++) {
        double c = rx[i].x * (1.0 / (double)(o->NY * o->NZ));
        normalise(o, &rx[i], 1.0 / (SY * SZ), 1.0 / (SY * SZ));
        rx[i].x = fluid ? c : 0.0;
        rx[i].y = rx[i].z = 0.0;
    }
    for (int j = 0; ry && j < o->NY; j++) {
        double c = ry[j].y * (1.0 / (double)(o->NX * o->NZ));
        normalise(o, &ry[j], 1.0 / (SX * SZ), 1.0 / (SX * SZ));
        ry[j].y = fluid ? c : 0.0;
        ry[j].x = ry[j].z = 0.0;
    }
    for (int k = 0; rz && k < o->NZ; k++) {
        double c = rz[k].z * (1.0 / (double)(o->NX * o->NY));
        normalise(o, &rz[k], 1.0 / (SX * SY), 1.0 / (SY * SZ));
        rz[k].z = fluid ? c : 0.0;
        rz[k].x = rz[k].y = 0.0;
    }
    return 0;
}

int oracle_dump_averages(oracle *o, ch4g_output *out) { return oracle_dump_rulers(o, out, NULL, NULL, NULL); }
