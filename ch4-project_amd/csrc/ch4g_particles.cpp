// C-ABI of the Lagrangian particles (include/ch4g.h, ch4g_particles_*):
// allocation and initial conditions on the host with the lattice's drand48
// stream and glibc libm (so that the per-particle constants are the
// reference's bits), and the hooks the step sequence calls around the fused
// lattice kernel: the feedback deposit before it, the hydro fields and the
// particle sub-steps after it.
//
// Compiled with -ffp-contract=off (host arithmetic included).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "ch4g_lattice.h"

using namespace ch4g;

namespace {

bool lhas(const ch4g_particle_config &pc, uint32_t f) { return (pc.flags & f) != 0; }

bool random_reinjection(const ch4g_lattice *L, const ch4g_particle_config &pc)
{
    return L->base.wall_f && lhas(pc, CH4G_LAGRANGE_BC_INELASTIC) && lhas(pc, CH4G_LAGRANGE_BC_INELASTIC_REINJECT) &&
           !lhas(pc, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM);
}

// The sub-steps draw from the lattice's drand48 stream (random re-injection,
// random reorientation): the draws are taken on the device in particle order
// and the rows stay in reference order.
bool device_draws(const ch4g_lattice *L, const ch4g_particle_config &pc)
{
    return random_reinjection(L, pc) || (pc.flags2 & CH4G_LAGRANGE_ORIENTATION_RANDOM) != 0;
}

// Switch combinations whose reference result is defined and restated.
int validate_particles(const ch4g_lattice *L, const ch4g_particle_config *pc)
{
    if (!pc) return fail(-1, "null particle configuration");
    if (L->itime != 0) return fail(-1, "particles are initialised before the first step (itime 0)");
    if (pc->particle_number < 0 || (int64_t)pc->particle_number * 8 > (int64_t)1 << 31)
        return fail(-1, "particle_number out of range");
    if (!L->base.fluid) return fail(-1, "particles need LB_FLUID");
    // parameters.c:538, :551 (radius and density) or :566, :581 (tau_drag, beta_coeff), :167
    const bool rad = (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY) != 0;
    const bool am = lhas(*pc, CH4G_LAGRANGE_ADDEDMASS);
    if (rad) {
        if (pc->particle_radius_types < 1 || pc->particle_radius_max < pc->particle_radius_min)
            return fail(-1, "Error in particle_radius parameters");
        if (pc->particle_density_types < 1 || pc->particle_density_max < pc->particle_density_min)
            return fail(-1, "Error in particle_density parameters");
    } else {
        if (pc->tau_drag_types < 1 || pc->tau_drag_max < pc->tau_drag_min)
            return fail(-1, "Error in tau_drag parameters");
        if (am && (pc->beta_coeff_types < 1 || pc->beta_coeff_max < pc->beta_coeff_min))
            return fail(-1, "Error in beta_coeff parameters");
    }
    if (!rad && (pc->flags2 & (CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS |
                               CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY | CH4G_LAGRANGE_BC_RADIUSANDDENSITY)))
        return fail(-1, "LAGRANGE_RADIUSandDENSITY_* / LAGRANGE_BC_RADIUSandDENSITY need LAGRANGE_RADIUSandDENSITY");
    if ((pc->flags2 & CH4G_LAGRANGE_BC_RADIUSANDDENSITY) && !am)
        return fail(-1, "LAGRANGE_BC_RADIUSandDENSITY needs LAGRANGE_ADDEDMASS (the radius is sqrt(3 nu beta tau))");
    if (!(pc->time_dt_lagr_factor >= 1.0)) return fail(-1, "time_dt_lagr > time_dt (time_dt_lagr_factor < 1)");
    // the reference reads beta_coeff and the gradient only when both switches
    // are on (lagrange.c:239-262 allocates beta under GRADIENT && ADDEDMASS)
    if (am && !lhas(*pc, CH4G_LAGRANGE_GRADIENT))
        return fail(-1, "LAGRANGE_ADDEDMASS needs LAGRANGE_GRADIENT (beta_coeff is set only with both)");
    const bool uses_dtu = lhas(*pc, CH4G_LAGRANGE_SMALLTAUD_BETA) || lhas(*pc, CH4G_LAGRANGE_SMALLTAUD_FLOATER) ||
                          lhas(*pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM);
    // these read Du/Dt, set only for beta != 0 with added mass; otherwise the
    // reference reads a value left over from another particle
    if (uses_dtu && (!am || pc->beta_coeff_min <= 0.0))
        return fail(-1, "LAGRANGE_SMALLTAUD_* and LAGRANGE_TWOWAY_MOMENTUM need LAGRANGE_ADDEDMASS with beta_coeff_min > 0");
    // orientation (lagrange.c:389-521 builds its ladders under LAGRANGE_GRADIENT;
    // Jeffery's equation reads the interpolated velocity gradient)
    const bool orient = lhas(*pc, CH4G_LAGRANGE_ORIENTATION), jeff = lhas(*pc, CH4G_LAGRANGE_ORIENTATION_JEFFREY);
    if (orient && !lhas(*pc, CH4G_LAGRANGE_GRADIENT)) return fail(-1, "LAGRANGE_ORIENTATION needs LAGRANGE_GRADIENT");
    if (!orient && (jeff || lhas(*pc, CH4G_LAGRANGE_ORIENTATION_ACTIVE) || lhas(*pc, CH4G_LAGRANGE_ORIENTATION_BC) ||
                    lhas(*pc, CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION)))
        return fail(-1, "LAGRANGE_ORIENTATION_JEFFREY / _ACTIVE / _BC / _SECONDORIENTATION need LAGRANGE_ORIENTATION");
    if (!jeff && (lhas(*pc, CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS) ||
                  lhas(*pc, CH4G_LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG)))
        return fail(-1, "LAGRANGE_ORIENTATION_JEFFREY_* need LAGRANGE_ORIENTATION_JEFFREY");
    // parameters.c:632, :646, :692
    if (jeff && (pc->aspect_ratio_types < 1 || pc->aspect_ratio_max < pc->aspect_ratio_min))
        return fail(-1, "Error in aspect_ratio parameters");
    if (lhas(*pc, CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS) &&
        (pc->gyrotaxis_velocity_types < 1 || pc->gyrotaxis_velocity_max < pc->gyrotaxis_velocity_min))
        return fail(-1, "Error in gyrotaxis_velocity parameters");
    if (lhas(*pc, CH4G_LAGRANGE_ORIENTATION_ACTIVE) &&
        (pc->swim_velocity_types < 1 || pc->swim_velocity_max < pc->swim_velocity_min))
        return fail(-1, "Error in swim_velocity parameters");
    // particle temperature: t_p exists with LB_TEMPERATURE only (typedef.h:467-470);
    // the log ladder and the fluid-temperature start sit inside
    // #ifdef LAGRANGE_TEMPERATURE (lagrange.c:587, :1005); parameters.c:616
    const bool tpart = lhas(*pc, CH4G_LAGRANGE_TEMPERATURE);
    if (tpart && !L->base.thermal) return fail(-1, "LAGRANGE_TEMPERATURE needs LB_TEMPERATURE");
    // the relaxation rate reads beta_coeff, a member only with LAGRANGE_GRADIENT
    // and LAGRANGE_ADDEDMASS (typedef.h:408-413)
    if (tpart && !am) return fail(-1, "LAGRANGE_TEMPERATURE needs LAGRANGE_GRADIENT and LAGRANGE_ADDEDMASS");
    if (!tpart && (lhas(*pc, CH4G_LAGRANGE_TEMPERATURE_INCREMENT_LOG) ||
                   lhas(*pc, CH4G_LAGRANGE_INITIAL_TEMPERATURE_FLUID)))
        return fail(-1, "LAGRANGE_TEMPERATURE_INCREMENT_LOG / LAGRANGE_INITIAL_TEMPERATURE_FLUID need LAGRANGE_TEMPERATURE");
    if (tpart && (pc->cp_coeff_types < 1 || pc->cp_coeff_max < pc->cp_coeff_min))
        return fail(-1, "Error in cp_coeff parameters");
    // the Eulerian probe type (parameters.c:744-750); its feedback volume
    // would be pow of a negative number (lagrange_feedback.c:87-90)
    if (lhas(*pc, CH4G_LAGRANGE_ADD_TRACER_STATIC) && !lhas(*pc, CH4G_LAGRANGE_ADD_TRACER))
        return fail(-1, "LAGRANGE_ADD_TRACER_STATIC needs LAGRANGE_ADD_TRACER");
    if (lhas(*pc, CH4G_LAGRANGE_ADD_TRACER_STATIC) && lhas(*pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM))
        return fail(-1, "LAGRANGE_ADD_TRACER_STATIC with LAGRANGE_TWOWAY_MOMENTUM: the probes' feedback is NaN");
    const uint32_t known2 = CH4G_LAGRANGE_INITIAL_GRID | CH4G_LAGRANGE_INITIAL_PAIRS | CH4G_LAGRANGE_ADDEDMASS_LIFT |
                            CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG | CH4G_LAGRANGE_ORIENTATION_DRAG |
                            CH4G_LAGRANGE_RADIUSANDDENSITY | CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS |
                            CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY | CH4G_LAGRANGE_BC_RADIUSANDDENSITY |
                            CH4G_LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC | CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE |
                            CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION |
                            CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION | CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT |
                            CH4G_LAGRANGE_ORIENTATION_RANDOM | CH4G_LAGRANGE_POLYMER |
                            CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP | CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS |
                            CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE;
    if ((pc->flags2 & ~known2) || pc->reserved) return fail(-1, "unknown bits in flags2 / reserved");
    if ((pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) &&
        (pc->critical_shear_rate_types < 1 || pc->critical_shear_rate_max < pc->critical_shear_rate_min ||
         pc->jump_time_types < 1 || pc->jump_time_max < pc->jump_time_min))
        return fail(-1, "Error in critical_shear_rate / jump_time parameters");
    if ((pc->flags2 & (CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS | CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE)) &&
        !(pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP))
        return fail(-1, "LAGRANGE_ORIENTATION_ACTIVE_JUMP_* need LAGRANGE_ORIENTATION_ACTIVE_JUMP");
    // the jump reads the tracer's velocity gradient (lagrange.c:3056-3066),
    // which the kernel carries only in its LAGRANGE_GRADIENT variant
    if ((pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) && !lhas(*pc, CH4G_LAGRANGE_GRADIENT))
        return fail(-1, "LAGRANGE_ORIENTATION_ACTIVE_JUMP needs LAGRANGE_GRADIENT");
    if ((pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE) && !L->base.thermal)
        return fail(-1, "LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE needs LB_TEMPERATURE");
    if ((pc->flags2 & CH4G_LAGRANGE_POLYMER) && !lhas(*pc, CH4G_LAGRANGE_GRADIENT))
        return fail(-1, "LAGRANGE_POLYMER needs LAGRANGE_GRADIENT");
    if ((pc->flags2 & CH4G_LAGRANGE_POLYMER) && !(pc->tau_polymer > 0.0)) return fail(-1, "tau_polymer must be > 0");
    if ((pc->flags2 & CH4G_LAGRANGE_ORIENTATION_RANDOM) && !lhas(*pc, CH4G_LAGRANGE_ORIENTATION))
        return fail(-1, "LAGRANGE_ORIENTATION_RANDOM needs LAGRANGE_ORIENTATION");
    if ((pc->flags2 & CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT) && !lhas(*pc, CH4G_LAGRANGE_BC_INELASTIC))
        return fail(-1, "LAGRANGE_BC_INELASTIC_SEDIMENT needs LAGRANGE_BC_INELASTIC");
    // swimming modes: members they read (typedef.h:416-430, :462-466)
    {
        const uint32_t f2 = pc->flags2;
        const bool active = lhas(*pc, CH4G_LAGRANGE_ORIENTATION_ACTIVE);
        const bool second = lhas(*pc, CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION);
        if ((f2 & (CH4G_LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC | CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE |
                   CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION |
                   CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION)) && !active)
            return fail(-1, "LAGRANGE_ORIENTATION_ACTIVE_* modes need LAGRANGE_ORIENTATION_ACTIVE");
        if ((f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE) && !L->base.thermal)
            return fail(-1, "LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE needs LB_TEMPERATURE");
        if ((f2 & (CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION |
                   CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION)) && !second)
            return fail(-1, "swimming along n needs LAGRANGE_ORIENTATION_SECONDORIENTATION");
        if ((f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION) && !jeff)
            return fail(-1, "LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION needs LAGRANGE_ORIENTATION_JEFFREY");
    }
    if ((pc->flags2 & CH4G_LAGRANGE_ORIENTATION_DRAG) && !jeff)
        return fail(-1, "LAGRANGE_ORIENTATION_DRAG needs LAGRANGE_ORIENTATION_JEFFREY (it reads aspect_ratio)");
    if ((pc->flags2 & CH4G_LAGRANGE_ORIENTATION_DRAG) && (pc->flags2 & CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG))
        return fail(-1, "LAGRANGE_ORIENTATION_DRAG replaces the Stokes drag and its wake correction");
    if ((pc->flags2 & (CH4G_LAGRANGE_ADDEDMASS_LIFT | CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG)) && !am)
        return fail(-1, "LAGRANGE_ADDEDMASS_LIFT / _WAKEDRAG need LAGRANGE_GRADIENT and LAGRANGE_ADDEDMASS");
    if ((pc->flags2 & CH4G_LAGRANGE_INITIAL_GRID) && pc->particle_number < 1)
        return fail(-1, "LAGRANGE_INITIAL_GRID needs particle_number >= 1");
    // variable gravity (parameters.c:589-600, lagrange.c:347-376)
    const bool gvar = lhas(*pc, CH4G_LAGRANGE_GRAVITY_VARIABLE);
    if (gvar && !lhas(*pc, CH4G_LAGRANGE_GRAVITY)) return fail(-1, "LAGRANGE_GRAVITY_VARIABLE needs LAGRANGE_GRAVITY");
    if (!gvar && lhas(*pc, CH4G_LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG))
        return fail(-1, "LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG needs LAGRANGE_GRAVITY_VARIABLE");
    if (gvar && (pc->gravity_coeff_types < 1 || pc->gravity_coeff_max < pc->gravity_coeff_min))
        return fail(-1, "Error in gravity_coeff parameters");
    // the Stokes start reads beta_coeff (lagrange.c:981-991)
    if (lhas(*pc, CH4G_LAGRANGE_INITIAL_VELOCITY_STOKES) && !am)
        return fail(-1, "LAGRANGE_INITIAL_VELOCITY_STOKES needs LAGRANGE_GRADIENT and LAGRANGE_ADDEDMASS");
    if (lhas(*pc, CH4G_LAGRANGE_TWOWAY_SMALLTAUD) && !lhas(*pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM))
        return fail(-1, "LAGRANGE_TWOWAY_SMALLTAUD needs LAGRANGE_TWOWAY_MOMENTUM");
    // the feedback sorts 32-bit stencil-cell keys with a sentinel above every cell
    if (lhas(*pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM) &&
        (uint64_t)(L->cfg.NX + 1) * (uint64_t)(L->cfg.NY + 1) * (uint64_t)(L->cfg.NZ + 1) >= ((uint64_t)1 << 32))
        return fail(-1, "two-way coupling needs (NX+1)(NY+1)(NZ+1) < 2^32");
    return 0;
}

// One property ladder (lagrange.c:264-298, same loop for beta :305-336):
// the values of the `types` steps repeat with period `repetitions` over the
// particle types after the fluid tracers.
void ladder(std::vector<double> &tab, int ntr, double trval, double types, double vmin, double vmax, bool logs,
            double repetitions)
{
    const int nt = (int)tab.size();
    double cycles = (nt - ntr) / types;
    double step;
    if (logs)
        step = types > 1 ? pow(vmax / vmin, 1.0 / (types - 1.0)) : 0;
    else
        step = types > 1 ? (vmax - vmin) / (types - 1.0) : 0;
    for (int n = 0; n < ntr; n++) tab[n] = trval;
    const int reps = (int)repetitions, ntyp = (int)types;
    for (int k = 0; k < (int)(cycles / repetitions); k++)
        for (int i = 0; i < ntyp; i++)
            for (int j = 0; j < reps; j++) {
                const int at = ntr + j + i * reps + k * reps * ntyp;
                if (at < nt) tab[at] = logs ? vmin * pow(step, (double)i) : vmin + i * step;
            }
}

// time_dt_lagr / tau_heat of the particle temperature (lagrange.c:4011-4013):
// 2 kappa / (tau_drag nu cp_p (3 - beta)) times the sub-step.  It holds only
// per-particle constants, so it is formed once here and the kernel reads it
// with its exponential (glibc exp: the bits the reference's call produces).
// Tracers (tau_drag 0, cp_p 0) get +inf and, as in the reference, t_p NaN.
void heat_constants(double kappa, double nu, double dt, double tau, double beta, double cp, double *fac,
                    double *efac)
{
    double fac1 = (kappa * 2.0) / (tau * nu * cp * (3 - beta));
    fac1 *= dt;
    *fac = fac1;
    *efac = exp(-fac1);
}

// Drag coefficients of a spheroid of aspect ratio alpha = a_par / a_perp
// across and along its axis, relative to the sphere of radius a_perp
// (lagrange.c:3276-3288); per-particle constants, formed here with glibc.
void spheroid_drag_constants(double alpha, double *c_perp, double *c_par)
{
    double beta = 0.0;
    if (alpha == 1) {
        *c_perp = *c_par = 1.0;
        return;
    }
    if (alpha > 1) beta = log(alpha + sqrt(alpha * alpha - 1.0)) / (alpha * sqrt(alpha * alpha - 1.0));
    if (alpha < 1) beta = acos(alpha) / (alpha * sqrt(1.0 - alpha * alpha));
    *c_perp = 8.0 * (alpha * alpha - 1.) / (3.0 * alpha * ((2.0 * alpha * alpha - 3.0) * beta + 1.0));
    *c_par = 4.0 * (alpha * alpha - 1.) / (3.0 * alpha * ((2.0 * alpha * alpha - 1.0) * beta - 1.0));
}

// particle volume and density ratio of the feedback (lagrange_feedback.c:84-90)
void feedback_constants(double nu, double tau, double beta, double *vol, double *drat)
{
    double squared_radius = 3.0 * nu * beta * tau;
    *vol = (4. / 3.) * 3.14159265359 * pow(squared_radius, 3. / 2.);
    *drat = (3.0 - beta) / (2.0 * beta);
}

uint64_t xs_get(const unsigned short xs[3])
{
    return (uint64_t)xs[0] | ((uint64_t)xs[1] << 16) | ((uint64_t)xs[2] << 32);
}

void xs_set(unsigned short xs[3], uint64_t v)
{
    xs[0] = (unsigned short)(v & 0xFFFF);
    xs[1] = (unsigned short)((v >> 16) & 0xFFFF);
    xs[2] = (unsigned short)((v >> 32) & 0xFFFF);
}

// stencil cells of the feedback sort: (NX+1)(NY+1)(NZ+1), corners from -1
uint64_t fb_cells(const ch4g_lattice *L)
{
    return (uint64_t)(L->cfg.NX + 1) * (uint64_t)(L->cfg.NY + 1) * (uint64_t)(L->cfg.NZ + 1);
}

// bytes of one set of feedback planes [3][N]
size_t fb_bytes(const ch4g_lattice *L) { return (size_t)3 * L->N * (L->dev32 ? sizeof(float) : sizeof(double)); }

void free_state(Particles *P)
{
    void *ptrs[] = {P->d_state, P->d_alive, P->d_ndraw, P->d_bsum, P->d_kin, P->d_vin, P->d_vout, P->d_sort_tmp,
                    P->d_keys[0], P->d_keys[1], P->d_rec, P->d_dep, P->d_ord, P->d_head, P->d_orig, P->d_orig2, P->d_alive2,
                    P->d_state2, P->d_ck, P->d_ck2, P->d_cv, P->d_cv2, P->d_ro_tmp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    P->d_state = nullptr;
    P->d_alive = P->d_ndraw = P->d_bsum = nullptr;
    P->d_kin = P->d_vin = P->d_vout = nullptr;
    P->d_rec = P->d_dep = nullptr;
    P->d_ord = nullptr;
    P->d_head = nullptr;
    P->d_sort_tmp = nullptr;
    P->d_orig = P->d_orig2 = P->d_alive2 = nullptr;
    P->d_state2 = nullptr;
    P->d_ck = P->d_ck2 = P->d_cv = P->d_cv2 = nullptr;
    P->d_ro_tmp = nullptr;
    P->ro_tmp_bytes = 0;
    P->d_keys[0] = P->d_keys[1] = nullptr;
    P->nkeys[0] = P->nkeys[1] = 0;
    P->cap = 0;
}

// (Re)allocate every array sized by the particle capacity.
int alloc_state(ch4g_lattice *L, int64_t cap)
{
    Particles *P = L->part;
    free_state(P);
    const int64_t c = cap > 0 ? cap : 1;
    int rc;
    if ((rc = dalloc(L, (void **)&P->d_state, (size_t)PF_COUNT * c * sizeof(double)))) return rc;
    if ((rc = dalloc(L, (void **)&P->d_alive, (size_t)c * sizeof(int)))) return rc;
    if ((rc = dalloc(L, (void **)&P->d_ndraw, (size_t)c * sizeof(int)))) return rc;
    if ((rc = dalloc(L, (void **)&P->d_orig, (size_t)c * sizeof(int)))) return rc;
    if (device_draws(L, P->pc)) {
        if ((rc = dalloc(L, (void **)&P->d_bsum, (size_t)((c + 255) / 256) * sizeof(int)))) return rc;
    } else {
        // rows reordered by cell (not with random re-injection, whose draws
        // follow the particles' order)
        if ((rc = dalloc(L, (void **)&P->d_state2, (size_t)PF_COUNT * c * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&P->d_alive2, (size_t)c * sizeof(int)))) return rc;
        if ((rc = dalloc(L, (void **)&P->d_orig2, (size_t)c * sizeof(int)))) return rc;
        uint32_t **cb[] = {&P->d_ck, &P->d_ck2, &P->d_cv, &P->d_cv2};
        for (uint32_t **b : cb)
            if ((rc = dalloc(L, (void **)b, (size_t)c * sizeof(uint32_t)))) return rc;
        P->cell_bits = 1;
        while (P->cell_bits < 32 && ((uint64_t)1 << P->cell_bits) <= (uint64_t)L->N) P->cell_bits++;
        P->ro_tmp_bytes = feedback_sort_bytes(c, P->cell_bits);
        if ((rc = dalloc(L, &P->d_ro_tmp, P->ro_tmp_bytes > 0 ? P->ro_tmp_bytes : 1))) return rc;
    }
    if (lhas(P->pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM)) {
        // the stencil-cell sort of the feedback and the cell heads
        const size_t M = (size_t)c;
        if ((rc = dalloc(L, (void **)&P->d_kin, M * sizeof(uint32_t)))) return rc;
        if ((rc = dalloc(L, (void **)&P->d_vin, M * sizeof(uint32_t)))) return rc;
        if ((rc = dalloc(L, (void **)&P->d_vout, M * sizeof(uint32_t)))) return rc;
        if ((rc = dalloc(L, (void **)&P->d_rec, 8 * M * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&P->d_dep, 24 * M * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&P->d_ord, M * sizeof(uint32_t)))) return rc;
        for (int b = 0; b < 2; b++)
            if ((rc = dalloc(L, (void **)&P->d_keys[b], M * sizeof(uint32_t)))) return rc;
        const uint64_t cells = fb_cells(L);
        if ((rc = dalloc(L, (void **)&P->d_head, cells * sizeof(int)))) return rc;
        if (hipMemsetAsync(P->d_head, 0xFF, cells * sizeof(int), L->stream) != hipSuccess)
            return fail(-2, "hipMemsetAsync failed");
        // the sentinel 2^key_bits - 1 lies beyond every cell and sorts last
        P->key_bits = 1;
        while (P->key_bits < 32 && ((uint64_t)1 << P->key_bits) <= cells) P->key_bits++;
        P->sort_tmp_bytes = feedback_sort_bytes((int64_t)M, P->key_bits);
        if ((rc = dalloc(L, &P->d_sort_tmp, P->sort_tmp_bytes > 0 ? P->sort_tmp_bytes : 1))) return rc;
    }
    P->cap = c;
    return 0;
}

PartParams part_params(ch4g_lattice *L)
{
    Particles *Q = L->part;
    PartParams P;
    memset(&P, 0, sizeof P);
    P.NX = L->cfg.NX;
    P.NY = L->cfg.NY;
    P.NZ = L->cfg.NZ;
    P.N = L->N;
    P.np = Q->np;
    P.ld = Q->cap;
    P.s = Q->d_state;
    P.alive = Q->d_alive;
    P.orig = Q->d_orig;
    P.ndraw = Q->d_ndraw;
    P.bsum = Q->d_bsum;
    P.rng = Q->d_rng;
    P.u = L->d_u;
    P.t = L->d_t;
    P.flags = Q->pc.flags;
    P.flags2 = Q->pc.flags2;
    P.nu = L->nu;
    P.fluid = L->base.fluid;
    P.thermal = L->base.thermal;
    P.wall_f = L->base.wall_f;
    P.wall_t = L->base.wall_t;
    P.Tw_bot = L->base.Tw_bot;
    P.Tw_top = L->base.Tw_top;
    P.T_bot = L->cfg.T_bot;
    P.T_top = L->cfg.T_top;
    P.inv_tau_polymer = 1. / Q->pc.tau_polymer;
    P.log_001 = log(0.01);
    P.first = L->itime == 1;
    P.dt = Q->dt;
    P.grav[0] = L->cfg.gravity_x;
    P.grav[1] = L->cfg.gravity_y;
    P.grav[2] = L->cfg.gravity_z;
    P.beta_t = L->cfg.beta_t;
    return P;
}

// Host records -> device rows (constants of the feedback recomputed).
int upload(ch4g_lattice *L, const ch4g_particle *in, int64_t n)
{
    Particles *Q = L->part;
    const int64_t cap = Q->cap;
    std::vector<double> rows((size_t)PF_COUNT * cap, 0.0);
    std::vector<int> alive((size_t)cap, 0);
    for (int64_t p = 0; p < n; p++) {
        const ch4g_particle &r = in[p];
        auto S = [&](int f) -> double & { return rows[(size_t)f * cap + p]; };
        S(PX) = r.x;
        S(PY) = r.y;
        S(PZ) = r.z;
        S(PVX) = r.vx;
        S(PVY) = r.vy;
        S(PVZ) = r.vz;
        S(PVXO) = r.vx_old;
        S(PVYO) = r.vy_old;
        S(PVZO) = r.vz_old;
        S(PAX) = r.ax;
        S(PAY) = r.ay;
        S(PAZ) = r.az;
        S(PAXO) = r.ax_old;
        S(PAYO) = r.ay_old;
        S(PAZO) = r.az_old;
        S(PUX) = r.ux;
        S(PUY) = r.uy;
        S(PUZ) = r.uz;
        S(PUXO) = r.ux_old;
        S(PUYO) = r.uy_old;
        S(PUZO) = r.uz_old;
        const double g[9] = {r.dx_ux, r.dy_ux, r.dz_ux, r.dx_uy, r.dy_uy, r.dz_uy, r.dx_uz, r.dy_uz, r.dz_uz};
        for (int m = 0; m < 9; m++) S(PG + m) = g[m];
        S(PDTU) = r.Dt_ux;
        S(PDTU + 1) = r.Dt_uy;
        S(PDTU + 2) = r.Dt_uz;
        S(PT) = r.t;
        S(PTOLD) = r.t_old;
        S(PDTT) = r.dt_t;
        S(PGT) = r.dx_t;
        S(PGT + 1) = r.dy_t;
        S(PGT + 2) = r.dz_t;
        S(PTAU) = r.tau_drag;
        S(PBETA) = r.beta_coeff;
        double vol = 0.0, drat = 0.0;
        if (lhas(Q->pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM)) feedback_constants(L->nu, r.tau_drag, r.beta_coeff, &vol, &drat);
        S(PVOL) = vol;
        S(PDRAT) = drat;
        S(PNAME) = r.name;
        S(PPO) = r.px;
        S(PPO + 1) = r.py;
        S(PPO + 2) = r.pz;
        S(PDTP) = r.dt_px;
        S(PDTP + 1) = r.dt_py;
        S(PDTP + 2) = r.dt_pz;
        S(PASP) = r.aspect_ratio;
        S(PGYV) = r.gyrotaxis_velocity;
        S(PSWIM) = r.swim_velocity;
        const double nv[6] = {r.nx, r.ny, r.nz, r.dt_nx, r.dt_ny, r.dt_nz};
        for (int m = 0; m < 6; m++) S(PNO + m) = nv[m];
        S(PTP) = r.t_p;
        S(PCP) = r.cp_p;
        double hf = 0.0, he = 0.0;
        if (lhas(Q->pc, CH4G_LAGRANGE_TEMPERATURE))
            heat_constants(L->kappa, L->nu, Q->dt, r.tau_drag, r.beta_coeff, r.cp_p, &hf, &he);
        S(PTFAC) = hf;
        S(PTEXP) = he;
        S(PGC) = r.gravity_coeff;
        double cpe = 0.0, cpa = 0.0;
        if (Q->pc.flags2 & CH4G_LAGRANGE_ORIENTATION_DRAG) spheroid_drag_constants(r.aspect_ratio, &cpe, &cpa);
        S(PCPE) = cpe;
        S(PCPA) = cpa;
        S(PSED) = r.sediment;
        const double ct[12] = {r.cxx, r.cyy, r.czz, r.cxy, r.cyz, r.cxz,
                               r.dt_cxx, r.dt_cyy, r.dt_czz, r.dt_cxy, r.dt_cyz, r.dt_cxz};
        for (int m = 0; m < 12; m++) S(PCT + m) = ct[m];
        const double jr[7] = {r.critical_shear_rate, r.shear_rate, r.jump_time, r.time_from_jump,
                              r.px_jump, r.py_jump, r.pz_jump};
        for (int m = 0; m < 7; m++) S(PCSR + m) = jr[m];
        alive[(size_t)p] = 1;
    }
    HIPCHK(hipMemcpyAsync(Q->d_state, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
    if (Q->d_state2) {
        // the rows the reordering leaves alone: uniform and never written by
        // this configuration's kernels; both buffers hold them
        HIPCHK(hipMemcpyAsync(Q->d_state2, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice,
                              L->stream));
        bool written[PF_COUNT] = {};
        const uint32_t fl = Q->pc.flags;
        auto mark = [&](int f, int k) {
            for (int m = 0; m < k; m++) written[f + m] = true;
        };
        mark(PX, PUZO - PX + 1);  // positions, velocities, accelerations, u and u_old
        if (fl & CH4G_LAGRANGE_GRADIENT) mark(PG, 9);
        if (fl & CH4G_LAGRANGE_ADDEDMASS) mark(PDTU, 3);
        if (L->base.thermal) {
            mark(PT, 3);
            mark(PGT, 3);
        }
        if (fl & CH4G_LAGRANGE_ORIENTATION) mark(PPO, 6);
        if (fl & CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION) mark(PNO, 6);
        if (fl & CH4G_LAGRANGE_TEMPERATURE) mark(PTP, 1);
        if (Q->pc.flags2 & CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT) mark(PSED, 1);
        if (Q->pc.flags2 & CH4G_LAGRANGE_POLYMER) mark(PCT, 12);
        if (Q->pc.flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) {
            mark(PSR, 1);
            mark(PTFJ, 4);
        }
        PermRows pr{};
        for (int f = 0; f < PF_COUNT; f++) {
            bool uniform = true;
            const double *row = rows.data() + (size_t)f * cap;
            for (int64_t p = 1; p < n && uniform; p++) uniform = memcmp(&row[p], &row[0], sizeof(double)) == 0;
            if (written[f] || !uniform) pr.r[pr.n++] = (uint8_t)f;
        }
        Q->perm = pr;
    }
    HIPCHK(hipMemcpyAsync(Q->d_alive, alive.data(), alive.size() * sizeof(int), hipMemcpyHostToDevice, L->stream));
    std::vector<int> orig((size_t)cap);
    for (int64_t p = 0; p < cap; p++) orig[(size_t)p] = (int)p;
    HIPCHK(hipMemcpyAsync(Q->d_orig, orig.data(), orig.size() * sizeof(int), hipMemcpyHostToDevice, L->stream));
    HIPCHK(hipMemsetAsync(Q->d_ndraw, 0, (size_t)cap * sizeof(int), L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    Q->np = n;
    Q->last_reorder = L->itime - REORDER_STEPS;  // reordered at the next step
    return 0;
}

}  // namespace

namespace ch4g {

int rng_to_host(ch4g_lattice *L)
{
    uint64_t v = 0;
    HIPCHK(hipStreamSynchronize(L->stream));
    HIPCHK(hipMemcpy(&v, L->part->d_rng, sizeof v, hipMemcpyDeviceToHost));
    xs_set(L->xsubi, v);
    return 0;
}

int rng_to_device(ch4g_lattice *L)
{
    uint64_t v = xs_get(L->xsubi);
    HIPCHK(hipStreamSynchronize(L->stream));
    HIPCHK(hipMemcpy(L->part->d_rng, &v, sizeof v, hipMemcpyHostToDevice));
    return 0;
}

void particles_free(ch4g_lattice *L)
{
    Particles *P = L->part;
    if (!P) return;
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    if (P->rng_on_device && P->d_rng) {
        uint64_t v = 0;
        if (hipMemcpy(&v, P->d_rng, sizeof v, hipMemcpyDeviceToHost) == hipSuccess) xs_set(L->xsubi, v);
    }
    free_state(P);
    for (int b = 0; b < 2; b++)
        if (P->d_fb[b]) (void)hipFree(P->d_fb[b]);
    if (P->d_rng) (void)hipFree(P->d_rng);
    delete P;
    L->part = nullptr;
    L->base.twoway = 0;
    L->base.part = 0;
    L->base.fb_prev = L->base.fb_now = nullptr;
}

// add_particle_feedbacks (main.c:125-130): from the second step on, the
// deposits of the particles as moved at the end of the previous step.
int particles_before_step(ch4g_lattice *L, int itime)
{
    Particles *Q = L->part;
    if (Q->d_state2 && itime - Q->last_reorder >= REORDER_STEPS) {
        // rows into cell order (particles move slowly: every REORDER_STEPS
        // steps keeps neighbouring threads on neighbouring cells)
        PartParams P = part_params(L);
        ReorderBufs B;
        B.kin = Q->d_ck;
        B.kout = Q->d_ck2;
        B.vin = Q->d_cv;
        B.vout = Q->d_cv2;
        B.tmp = Q->d_ro_tmp;
        B.tmp_bytes = Q->ro_tmp_bytes;
        B.key_bits = Q->cell_bits;
        B.s_out = Q->d_state2;
        B.alive_out = Q->d_alive2;
        B.orig_out = Q->d_orig2;
        B.orig_in = Q->d_orig;
        B.rows = Q->perm;
        if (launch_reorder(P, B, L->stream))
            return fail(-2, "particle reorder failed: %s", hipGetErrorString(hipGetLastError()));
        std::swap(Q->d_state, Q->d_state2);
        std::swap(Q->d_alive, Q->d_alive2);
        std::swap(Q->d_orig, Q->d_orig2);
        Q->last_reorder = itime;
    }
    if (!lhas(Q->pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM) || itime < 2) return 0;
    const int par = itime & 1;
    PartParams P = part_params(L);
    double gv[3] = {0.0, 0.0, 0.0};
    if (lhas(Q->pc, CH4G_LAGRANGE_GRAVITY)) {
        gv[0] = L->cfg.gravity_x;
        gv[1] = -L->cfg.gravity_y;
        gv[2] = L->cfg.gravity_z;
    }
    FeedbackBufs B;
    memset(&B, 0, sizeof B);
    B.kin = Q->d_kin;
    B.vin = Q->d_vin;
    B.kout = Q->d_keys[par];
    B.vout = Q->d_vout;
    // the heads of the previous call's cells go back to -1 first
    B.old_keys = Q->d_keys[par ^ 1];
    B.old_n = Q->nkeys[par ^ 1];
    B.head = Q->d_head;
    B.rec = Q->d_rec;
    B.dep = Q->d_dep;
    B.ord = Q->d_ord;
    B.tmp = Q->d_sort_tmp;
    B.tmp_bytes = Q->sort_tmp_bytes;
    B.key_bits = Q->key_bits;
    B.fb = Q->d_fb[par];
    B.fb_float = L->dev32;
    for (int c = 0; c < 3; c++) B.gv[c] = gv[c];
    if (launch_feedback(P, B, L->stream)) return fail(-2, "feedback launch failed: %s", hipGetErrorString(hipGetLastError()));
    Q->nkeys[par ^ 1] = 0;
    Q->nkeys[par] = Q->np;
    return 0;
}

// Float-storage lattices interpolate float fields, 16 bytes per node (u_x,
// u_y, u_z, T): half the bytes of the field pass, the 268 MB of 256^3 about
// the size of the MALL, and one aligned 16-byte load per corner (u and T of a
// node in one request).
int ensure_float_fields(ch4g_lattice *L)
{
    const size_t N = (size_t)L->N;
    int rc;
    if (!L->d_uT32 && (rc = dalloc(L, (void **)&L->d_uT32, N * sizeof(float4)))) return rc;
    return 0;
}

// Whether the step kernel writes the fields the particles interpolate: float
// storage, no two-way feedback (the next step needs the moved particles
// first) and no device random draws (their order with the host's HIT phase
// walk is part of the stream).  The particle update of step s then runs
// after the kernel of step s + 1, which has gathered the fields of step s:
// the same inputs, without a separate field pass.
bool fields_fused(const ch4g_lattice *L)
{
    return L->part && L->dev32 && !L->base.twoway && !L->part->rng_on_device;
}

// time_dt_lagr_factor particle sub-steps of step `step` (main.c:169-182) from
// the fields of that step already in place (float or FP64 planes).
int move_particles_on_fields(ch4g_lattice *L, int64_t step)
{
    int rc = ensure_scratch(L);
    if (rc) return rc;
    PartParams P = part_params(L);
    P.first = step == 1;
    if (L->dev32) {
        P.uT = L->d_uT32;
    }
    for (int s = 0; s < L->part->substeps; s++)
        if (launch_particles(P, L->stream)) return fail(-2, "particle launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

// main.c:163-183: hydro fields of the step (with their halo rules applied in
// the particle kernel), then time_dt_lagr_factor sub-steps.
int particles_after_step(ch4g_lattice *L)
{
    int rc = ensure_scratch(L);  // the FP64 field planes the gather writes
    if (rc) return rc;
    if (L->dev32) {
        if ((rc = ensure_float_fields(L))) return rc;
        KParams K = params_for(L, L->first, (int)L->itime);
        launch_fields_f32(K, (const float *)L->f[L->cur], (const float *)L->g[L->cur], (const float *)L->tlag,
                          L->d_uT32, L->stream);
    } else {
        gather(L, nullptr, nullptr, nullptr, L->base.fluid ? L->d_u : nullptr, L->base.thermal ? L->d_t : nullptr);
    }
    HIPCHK(hipGetLastError());
    return move_particles_on_fields(L, L->itime);
}

// The particle update still owed by a fused sequence (end of ch4g_step).
int particles_flush(ch4g_lattice *L)
{
    if (!L->part || !L->part->pending) return 0;
    L->part->pending = 0;
    return particles_after_step(L);
}

}  // namespace ch4g

extern "C" {

int ch4g_particles_init(ch4g_lattice *L, const ch4g_particle_config *pc)
{
    if (!L) return fail(-1, "null lattice");
    int rc = validate_particles(L, pc);
    if (rc) return rc;
    particles_free(L);
    destroy_graphs(L);
    L->part = new (std::nothrow) Particles();
    if (!L->part) return fail(-3, "out of host memory");
    Particles *Q = L->part;
    Q->pc = *pc;
    Q->dt = 1.0 / pc->time_dt_lagr_factor; // parameters.c:164 with time_dt = 1
    Q->substeps = (int)pc->time_dt_lagr_factor;

    // particle types (parameters.c:559-584, :621-695, :743-755) and their
    // ladders in the reference's order (lagrange.c:264-336, :389-521)
    const bool am = lhas(*pc, CH4G_LAGRANGE_GRADIENT) && lhas(*pc, CH4G_LAGRANGE_ADDEDMASS);
    const bool orient = lhas(*pc, CH4G_LAGRANGE_ORIENTATION);
    const bool jeff = orient && lhas(*pc, CH4G_LAGRANGE_ORIENTATION_JEFFREY);
    const bool gyro = jeff && lhas(*pc, CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS);
    const bool swim = orient && lhas(*pc, CH4G_LAGRANGE_ORIENTATION_ACTIVE);
    const bool tpart = lhas(*pc, CH4G_LAGRANGE_TEMPERATURE);
    const bool gvar = lhas(*pc, CH4G_LAGRANGE_GRAVITY_VARIABLE);
    const bool jump = swim && (pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP);
    const bool rad = (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY) != 0;
    double types = pc->tau_drag_types;
    if (am) types *= pc->beta_coeff_types;
    if (rad) types = pc->particle_radius_types * pc->particle_density_types;  // parameters.c:541, :554
    if (gvar) types *= pc->gravity_coeff_types;  // parameters.c:600
    if (tpart) types *= pc->cp_coeff_types;  // parameters.c:615
    if (jeff) types *= pc->aspect_ratio_types;
    if (gyro) types *= pc->gyrotaxis_velocity_types;
    if (swim) types *= pc->swim_velocity_types;
    if (jump) types *= pc->critical_shear_rate_types * pc->jump_time_types;  // parameters.c:709, :720
    // fluid_tracers (parameters.c:743-755): the tracer, and with
    // LAGRANGE_ADD_TRACER_STATIC an Eulerian probe type before it
    const int ntr = lhas(*pc, CH4G_LAGRANGE_ADD_TRACER) ? (lhas(*pc, CH4G_LAGRANGE_ADD_TRACER_STATIC) ? 2 : 1) : 0;
    types += ntr;
    const int nt = (int)types;
    std::vector<double> tau((size_t)nt, 0.0), beta((size_t)nt, 0.0), asp((size_t)nt, 0.0), gyv((size_t)nt, 0.0),
        swv((size_t)nt, 0.0), cpv((size_t)nt, 0.0), gcv((size_t)nt, 0.0), csr((size_t)nt, 0.0), jtv((size_t)nt, 0.0);
    double rep = 1.0;
    std::vector<double> rv, dv;  // radius and density ladders (LAGRANGE_RADIUSandDENSITY)
    if (rad) {
        // lagrange.c:150-260: radius and density ladders (tracers 0 and 1),
        // tau = (2/9) rho_p r^2 / nu, or with added mass beta = 3 / (1 + 2 rho_p)
        // and tau = r^2 / (3 beta nu) (rho_f = 1)
        rv.assign((size_t)nt, 0.0);
        dv.assign((size_t)nt, 0.0);
        ladder(rv, ntr, 0.0, pc->particle_radius_types, pc->particle_radius_min, pc->particle_radius_max,
               (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS) != 0, rep);
        rep *= pc->particle_radius_types;
        ladder(dv, ntr, 1.0, pc->particle_density_types, pc->particle_density_min, pc->particle_density_max,
               (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY) != 0, rep);
        rep *= pc->particle_density_types;
        for (int i = 0; i < nt; i++) tau[(size_t)i] = (2. / 9.) * dv[(size_t)i] * pow(rv[(size_t)i], 2.0) / (L->nu);
        if (am)
            for (int i = 0; i < nt; i++) {
                beta[(size_t)i] = 3.0 / (1.0 + 2.0 * dv[(size_t)i]);
                tau[(size_t)i] = pow(rv[(size_t)i], 2.0) / (3.0 * beta[(size_t)i] * L->nu);
            }
    } else {
        ladder(tau, ntr, 0.0, pc->tau_drag_types, pc->tau_drag_min, pc->tau_drag_max,
               lhas(*pc, CH4G_LAGRANGE_TAUDRAG_INCREMENT_LOG), rep);
        if (lhas(*pc, CH4G_LAGRANGE_ADD_TRACER_STATIC)) tau[0] = -1.0;  // lagrange.c:280-282
        rep *= pc->tau_drag_types;
        if (am) {
            ladder(beta, ntr, 1.0, pc->beta_coeff_types, pc->beta_coeff_min, pc->beta_coeff_max,
                   lhas(*pc, CH4G_LAGRANGE_ADDEDMASS_INCREMENT_LOG), rep);
            rep *= pc->beta_coeff_types;
        }
    }
    if (gvar) {  // after beta, before the orientation ladders (lagrange.c:347-384)
        ladder(gcv, ntr, 0.0, pc->gravity_coeff_types, pc->gravity_coeff_min, pc->gravity_coeff_max,
               lhas(*pc, CH4G_LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG), rep);
        rep *= pc->gravity_coeff_types;
    }
    if (jeff) {
        ladder(asp, ntr, 1.0, pc->aspect_ratio_types, pc->aspect_ratio_min, pc->aspect_ratio_max,
               lhas(*pc, CH4G_LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG), rep);
        rep *= pc->aspect_ratio_types;
    }
    if (gyro) { // geometric steps only (lagrange.c:448-466)
        ladder(gyv, ntr, 0.0, pc->gyrotaxis_velocity_types, pc->gyrotaxis_velocity_min, pc->gyrotaxis_velocity_max,
               true, rep);
        rep *= pc->gyrotaxis_velocity_types;
    }
    if (swim) { // linear steps only (lagrange.c:503-518)
        ladder(swv, ntr, 0.0, pc->swim_velocity_types, pc->swim_velocity_min, pc->swim_velocity_max, false, rep);
        rep *= pc->swim_velocity_types;
    }
    if (jump) {  // lagrange.c:526-570: linear steps; tracers 1e10 and 0
        ladder(csr, ntr, 1.e+10, pc->critical_shear_rate_types, pc->critical_shear_rate_min,
               pc->critical_shear_rate_max, false, rep);
        rep *= pc->critical_shear_rate_types;
        ladder(jtv, ntr, 0.0, pc->jump_time_types, pc->jump_time_min, pc->jump_time_max, false, rep);
        rep *= pc->jump_time_types;
    }
    if (tpart) { // after the orientation ladders (lagrange.c:579-617)
        ladder(cpv, ntr, 0.0, pc->cp_coeff_types, pc->cp_coeff_min, pc->cp_coeff_max,
               lhas(*pc, CH4G_LAGRANGE_TEMPERATURE_INCREMENT_LOG), rep);
        rep *= pc->cp_coeff_types;
    }

    // the per-type table of particle_properties.dat (lagrange.c:644-700)
    Q->types.assign((size_t)nt, ch4g_particle_type{});
    for (int t = 0; t < nt; t++) {
        ch4g_particle_type &y = Q->types[(size_t)t];
        if (rad) {
            y.particle_radius = rv[(size_t)t];
            y.particle_density = dv[(size_t)t];
        }
        y.tau_drag = tau[(size_t)t];
        y.beta_coeff = beta[(size_t)t];
        y.gravity_coeff = gcv[(size_t)t];
        y.cp_coeff = cpv[(size_t)t];
        y.aspect_ratio = asp[(size_t)t];
        y.gyrotaxis_velocity = gyv[(size_t)t];
        y.swim_velocity = swv[(size_t)t];
        y.jump_time = jtv[(size_t)t];
        y.critical_shear_rate = csr[(size_t)t];
    }
    for (int64_t i = 0; i < pc->particle_number; i++) Q->types[(size_t)(i % nt)].count += 1.0;

    // names, properties and positions (lagrange.c:634-641, :711-771): three
    // draws per particle from the lattice's stream, then with orientation
    // random_vector's two (random.c:139-150, lagrange.c:891), zero velocity
    const int64_t np = pc->particle_number;
    std::vector<ch4g_particle> h((size_t)(np > 0 ? np : 1));
    memset(h.data(), 0, h.size() * sizeof(ch4g_particle));
    const double LNX = L->cfg.NX, LNY = L->cfg.NY, LNZ = L->cfg.NZ;
    double gd = 1.0;
    int gnx = 1, gny = 1, gnz = 1;
    for (int64_t i = 0; i < np; i++) {
        ch4g_particle &r = h[(size_t)i];
        r.name = (double)i;
        const int type = ((int)r.name) % nt;
        r.tau_drag = tau[(size_t)type];
        if (am) r.beta_coeff = beta[(size_t)type];
        if (gvar) r.gravity_coeff = gcv[(size_t)type];  // lagrange.c:719-722
        r.x = 0 + erand48(L->xsubi) * LNX;
        r.y = 0 + erand48(L->xsubi) * LNY;
        r.z = 0 + erand48(L->xsubi) * LNZ;
        if (pc->flags2 & CH4G_LAGRANGE_INITIAL_GRID) {
            // lagrange.c:774-795 (the draws above are still consumed): the
            // spacing from an integer quotient, rows x-fastest
            if (i == 0) {
                const int64_t q = ((int64_t)L->cfg.NX * L->cfg.NY * L->cfg.NZ) / np;
                gd = pow((double)q, 1. / 3.);
                gnx = (int)floor(L->cfg.NX / gd);
                gny = (int)floor(L->cfg.NY / gd);
                gnz = (int)floor(L->cfg.NZ / gd);
                if (gnx == 0) gnx = 1;
                if (gny == 0) gny = 1;
                if (gnz == 0) gnz = 1;
            }
            const int ip = (int)(i % gnx);
            const int jp = (int)(i / gnx) % gny;
            const int kp = (int)(i / ((int64_t)gnx * gny)) % gnz;
            r.x = 0 + gd / 2. + ip * gd;
            r.y = 0 + gd / 2. + jp * gd;
            r.z = 0 + gd / 2. + kp * gd;
        }
        if ((pc->flags2 & CH4G_LAGRANGE_INITIAL_PAIRS) && i % 2 == 0 && i > 0) {
            // lagrange.c:798-814: particle i-1 at distance 1 from particle i
            // along a random_vector (two draws), mirrored per axis if outside
            const double theta = acos(2.0 * erand48(L->xsubi) - 1.0);
            const double phi = 2 * 3.14159265359 * erand48(L->xsubi);
            double st, ct, sp, cp;
            sincos(theta, &st, &ct);
            sincos(phi, &sp, &cp);
            const double d[3] = {st * sp, st * cp, ct};
            const double dist = 1.0;
            ch4g_particle &q = h[(size_t)(i - 1)];
            q.x = r.x + d[0] * dist;
            q.y = r.y + d[1] * dist;
            q.z = r.z + d[2] * dist;
            if (q.x < 0 || q.x >= 0 + LNX) q.x = r.x - d[0] * dist;
            if (q.y < 0 || q.y >= 0 + LNY) q.y = r.y - d[1] * dist;
            if (q.z < 0 || q.z >= 0 + LNZ) q.z = r.z - d[2] * dist;
        }
        if (jeff) r.aspect_ratio = asp[(size_t)type];
        if (gyro) r.gyrotaxis_velocity = gyv[(size_t)type];
        if (swim) r.swim_velocity = swv[(size_t)type];
        if (jump) {
            r.critical_shear_rate = csr[(size_t)type];
            r.jump_time = jtv[(size_t)type];
        }
        if (tpart) {
            r.cp_p = cpv[(size_t)type];  // lagrange.c:730-732
            r.t_p = L->cfg.T_bot;        // lagrange.c:862-864
        }
        if (lhas(*pc, CH4G_LAGRANGE_INITIAL_VELOCITY_STOKES)) {
            // the settling velocity (lagrange.c:981-991); it replaces the
            // fluid velocity of LAGRANGE_INITIAL_VELOCITY_FLUID, which the
            // reference sets first
            r.vx = -L->cfg.gravity_x * (1.0 - r.beta_coeff) * r.tau_drag;
            r.vy = -L->cfg.gravity_y * (1.0 - r.beta_coeff) * r.tau_drag;
            r.vz = -L->cfg.gravity_z * (1.0 - r.beta_coeff) * r.tau_drag;
        }
        if (orient) {
            // random_vector with glibc sincos for each angle: the call gcc
            // makes of sin(a) and cos(a) together; separate sin and cos
            // calls differ from it in the last bit for some angles
            const double theta = acos(2.0 * erand48(L->xsubi) - 1.0);
            const double phi = 2 * 3.14159265359 * erand48(L->xsubi); // two_pi of typedef.h:367
            double st, ct, sp, cp;
            sincos(theta, &st, &ct);
            sincos(phi, &sp, &cp);
            r.px = st * sp;
            r.py = st * cp;
            r.pz = ct;
            if (lhas(*pc, CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION)) {
                // n = p x v / |p x v| for a second random_vector v (lagrange.c:904-917)
                const double t2 = acos(2.0 * erand48(L->xsubi) - 1.0);
                const double f2 = 2 * 3.14159265359 * erand48(L->xsubi);
                double s2, c2, sf, cf;
                sincos(t2, &s2, &c2);
                sincos(f2, &sf, &cf);
                const double vx = s2 * sf, vy = s2 * cf, vz = c2;
                r.nx = r.py * vz - r.pz * vy;
                r.ny = -r.px * vz + r.pz * vx;
                r.nz = r.px * vy - r.py * vx;
                const double val = sqrt(r.nx * r.nx + r.ny * r.ny + r.nz * r.nz);
                r.nx /= val;
                r.ny /= val;
                r.nz /= val;
            }
        }
        if (jump) {
            // lagrange.c:927-934: time from the last jump in [0, 10 jump_time),
            // one more draw; the jump direction starts as p
            r.time_from_jump = 10. * erand48(L->xsubi) * r.jump_time;
            r.px_jump = r.px;
            r.py_jump = r.py;
            r.pz_jump = r.pz;
        }
        if (pc->flags2 & CH4G_LAGRANGE_POLYMER) {
            // lagrange.c:937-955: the diagonal of the conformation tensor from
            // a random_vector (two more draws), the rest 0
            const double theta = acos(2.0 * erand48(L->xsubi) - 1.0);
            const double phi = 2 * 3.14159265359 * erand48(L->xsubi);
            double st, ct, sp, cp;
            sincos(theta, &st, &ct);
            sincos(phi, &sp, &cp);
            r.cxx = st * sp;
            r.cyy = st * cp;
            r.czz = ct;
        }
    }
    if ((rc = alloc_state(L, np))) goto err;
    if (lhas(*pc, CH4G_LAGRANGE_TWOWAY_MOMENTUM)) {
        for (int b = 0; b < 2; b++) {
            if ((rc = dalloc(L, (void **)&Q->d_fb[b], fb_bytes(L)))) goto err;
            if (hipMemsetAsync(Q->d_fb[b], 0, fb_bytes(L), L->stream) != hipSuccess) {
                rc = fail(-2, "hipMemsetAsync failed");
                goto err;
            }
        }
        L->base.twoway = 1;
        L->base.forced = 1; // the feedback enters through the force field
    }
    if (device_draws(L, *pc)) {
        if ((rc = dalloc(L, (void **)&Q->d_rng, 2 * sizeof(uint64_t)))) goto err;
        Q->rng_on_device = 1;
        if ((rc = rng_to_device(L))) goto err;
    }
    if ((rc = upload(L, h.data(), np))) goto err;
    if (lhas(*pc, CH4G_LAGRANGE_INITIAL_VELOCITY_FLUID) || lhas(*pc, CH4G_LAGRANGE_INITIAL_TEMPERATURE_FLUID)) {
        // fields of the initial state (FP64 planes on every storage), then
        // lagrange.c:968-980 and :1004-1015 in that order
        if ((rc = ensure_scratch(L))) goto err;
        gather(L, nullptr, nullptr, nullptr, L->d_u, L->d_t);
        const PartParams P = part_params(L);
        const bool vel = lhas(*pc, CH4G_LAGRANGE_INITIAL_VELOCITY_FLUID);
        const bool keep_v = lhas(*pc, CH4G_LAGRANGE_INITIAL_VELOCITY_STOKES);
        if ((vel && launch_particle_finit(P, true, !keep_v, L->stream)) ||
            (lhas(*pc, CH4G_LAGRANGE_INITIAL_TEMPERATURE_FLUID) && launch_particle_finit(P, false, false, L->stream)) ||
            hipStreamSynchronize(L->stream) != hipSuccess) {
            rc = fail(-2, "particle field initialisation failed: %s", hipGetErrorString(hipGetLastError()));
            goto err;
        }
    }
    L->base.part = fields_fused(L) ? 1 : 0;
    return 0;
err:
    particles_free(L);
    return rc;
}

int64_t ch4g_particle_size(void) { return (int64_t)sizeof(ch4g_particle); }

int ch4g_particle_types(ch4g_lattice *L, ch4g_particle_type *out, int cap, int *ntypes)
{
    if (!L || !ntypes || (cap > 0 && !out)) return fail(-1, "null argument");
    if (!L->part) return fail(-1, "ch4g_particles_init first");
    const int nt = (int)L->part->types.size();
    for (int t = 0; t < nt && t < cap; t++) out[t] = L->part->types[(size_t)t];
    *ntypes = nt;
    return 0;
}

int ch4g_particles_count(ch4g_lattice *L, int64_t *n)
{
    if (!L || !n) return fail(-1, "null argument");
    *n = 0;
    Particles *Q = L->part;
    if (!Q || Q->np == 0) return 0;
    std::vector<int> alive((size_t)Q->np);
    HIPCHK(hipMemcpyAsync(alive.data(), Q->d_alive, alive.size() * sizeof(int), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    int64_t c = 0;
    for (int a : alive) c += a != 0;
    *n = c;
    return 0;
}

int ch4g_get_particles(ch4g_lattice *L, ch4g_particle *out, int64_t cap, int64_t *n)
{
    if (!L || !n || (cap > 0 && !out)) return fail(-1, "null argument");
    *n = 0;
    Particles *Q = L->part;
    if (!Q || Q->np == 0) return 0;
    const int64_t ld = Q->cap, np = Q->np;
    std::vector<double> rows((size_t)PF_COUNT * ld);
    std::vector<int> alive((size_t)np), orig((size_t)np);
    HIPCHK(hipMemcpyAsync(rows.data(), Q->d_state, rows.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipMemcpyAsync(alive.data(), Q->d_alive, alive.size() * sizeof(int), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipMemcpyAsync(orig.data(), Q->d_orig, orig.size() * sizeof(int), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    // records in the reference's particle order (rows may be in cell order)
    std::vector<int64_t> row_of((size_t)np);
    for (int64_t p = 0; p < np; p++) row_of[(size_t)orig[(size_t)p]] = p;
    int64_t k = 0;
    for (int64_t o = 0; o < np; o++) {
        const int64_t p = row_of[(size_t)o];
        if (!alive[(size_t)p]) continue;
        if (k < cap) {
            auto S = [&](int f) { return rows[(size_t)f * ld + p]; };
            ch4g_particle &r = out[k];
            r.name = S(PNAME);
            r.x = S(PX);
            r.y = S(PY);
            r.z = S(PZ);
            r.vx = S(PVX);
            r.vy = S(PVY);
            r.vz = S(PVZ);
            r.vx_old = S(PVXO);
            r.vy_old = S(PVYO);
            r.vz_old = S(PVZO);
            r.ax = S(PAX);
            r.ay = S(PAY);
            r.az = S(PAZ);
            r.ax_old = S(PAXO);
            r.ay_old = S(PAYO);
            r.az_old = S(PAZO);
            r.ux = S(PUX);
            r.uy = S(PUY);
            r.uz = S(PUZ);
            r.ux_old = S(PUXO);
            r.uy_old = S(PUYO);
            r.uz_old = S(PUZO);
            double *g = &r.dx_ux;
            for (int m = 0; m < 9; m++) g[m] = S(PG + m);
            r.Dt_ux = S(PDTU);
            r.Dt_uy = S(PDTU + 1);
            r.Dt_uz = S(PDTU + 2);
            r.t = S(PT);
            r.t_old = S(PTOLD);
            r.dt_t = S(PDTT);
            r.dx_t = S(PGT);
            r.dy_t = S(PGT + 1);
            r.dz_t = S(PGT + 2);
            r.tau_drag = S(PTAU);
            r.beta_coeff = S(PBETA);
            r.px = S(PPO);
            r.py = S(PPO + 1);
            r.pz = S(PPO + 2);
            r.dt_px = S(PDTP);
            r.dt_py = S(PDTP + 1);
            r.dt_pz = S(PDTP + 2);
            r.aspect_ratio = S(PASP);
            r.gyrotaxis_velocity = S(PGYV);
            r.swim_velocity = S(PSWIM);
            double *nv = &r.nx;
            for (int m = 0; m < 6; m++) nv[m] = S(PNO + m);
            r.t_p = S(PTP);
            r.cp_p = S(PCP);
            r.gravity_coeff = S(PGC);
            r.sediment = S(PSED);
            double *ct = &r.cxx;
            for (int m = 0; m < 12; m++) ct[m] = S(PCT + m);
            double *jr = &r.critical_shear_rate;
            for (int m = 0; m < 7; m++) jr[m] = S(PCSR + m);
        }
        k++;
    }
    *n = k;
    return 0;
}

int ch4g_set_particles(ch4g_lattice *L, const ch4g_particle *in, int64_t n)
{
    if (!L || (n > 0 && !in)) return fail(-1, "null argument");
    if (!L->part) return fail(-1, "ch4g_particles_init first");
    if (n < 0 || n * 8 > (int64_t)1 << 31) return fail(-1, "particle count out of range");
    int rc;
    if (n > L->part->cap) {
        HIPCHK(hipStreamSynchronize(L->stream));
        if ((rc = alloc_state(L, n))) return rc;
    }
    return upload(L, in, n);
}

}  // extern "C"
