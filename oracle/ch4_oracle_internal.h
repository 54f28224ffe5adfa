/*
 * Shared declarations of the FP64 CPU oracle (test infrastructure only):
 * the reference-structured lattice record and the hooks between the lattice
 * step (ch4_oracle.c) and the particle restatement (ch4_oracle_lagrange.c).
 */
#ifndef CH4_ORACLE_INTERNAL_H
#define CH4_ORACLE_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "../include/ch4g.h"

#define NP 19
#define BRD 1

typedef struct { double p[NP]; } opop;
typedef struct { double x, y, z; } ovec;
typedef struct { double xx, xy, xz, yx, yy, yz, zx, zy, zz; } oten;

typedef struct oracle {
    ch4g_config cfg;
    int NX, NY, NZ;         /* interior sizes */
    int GX, GY, GZ;         /* with ghosts */
    size_t ng;              /* ghost-extended node count */
    opop *p, *rhs_p, *g, *rhs_g;
    ovec *u, *force;
    double *dens, *t, *tau_les;
    /* lattice, lb.c:139-230 and :368-377 */
    double wgt[NP], cx[NP], cy[NP], cz[NP];
    int inv[NP];
    double cs, cs2, cs4, invcs, invcs2, invcs4, twocs2, twocs4, invtwocs2, invtwocs4;
    double nu, kappa, deltaT;
    /* HIT forcing, initial_conditions.c:822-875 */
    int nk;
    ovec vk[CH4G_HIT_NK];
    double vk2[CH4G_HIT_NK];
    ovec phi[CH4G_HIT_NK];
    /* drand48 state (48-bit) */
    uint64_t x48;
    long itime;
    /* Lagrangian particles (ch4_oracle_lagrange.c), NULL without them */
    struct olagr *lagr;
} oracle;

#define IDX(o, i, j, k) ((((size_t)(k)) * (size_t)(o)->GY + (size_t)(j)) * (size_t)(o)->GX + (size_t)(i))
#define HAS(o, f) (((o)->cfg.flags & (f)) != 0)

double oracle_drand48(oracle *o);

/* particle hooks: add_particle_feedbacks (main.c:125-130), the Lagrangian
 * block after dump_averages (main.c:163-183), release                      */
int olagr_twoway(const oracle *o);
void olagr_feedback(oracle *o);
void olagr_after_step(oracle *o);
void olagr_free(oracle *o);

#endif
