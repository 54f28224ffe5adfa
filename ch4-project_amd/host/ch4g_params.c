/*
 * Reader of a reference run description: the active #define names of a
 * define.h (src/define.h, examples/<case>/define.h) and the name/value pairs
 * of param.in (src/parameters.c:3-39 read_parameter, :41-800
 * assign_parameters).  Switches that change the streaming step in ways the
 * device lattice does not implement are rejected by name.
 */
#define _POSIX_C_SOURCE 200809L
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/ch4g_dropin.h"

/* messages go to stderr, as read_parameter (parameters.c:15-34) reports */
static char g_msg[512];
static int perr(const char *fmt, const char *a)
{
    snprintf(g_msg, sizeof g_msg, fmt, a);
    fprintf(stderr, "Error message -> %s\n", g_msg);
    return -1;
}

/* define.h switches mapped onto ch4g flags */
static const struct { const char *name; unsigned flag; } FLAGS[] = {
    {"LB_FLUID", CH4G_LB_FLUID},
    {"LB_TEMPERATURE", CH4G_LB_TEMPERATURE},
    {"LB_FLUID_FORCING_POISEUILLE", CH4G_FLUID_FORCING_POISEUILLE},
    {"LB_FLUID_FORCING_CHANNEL", CH4G_FLUID_FORCING_CHANNEL},
    {"LB_FLUID_FORCING_HIT", CH4G_FLUID_FORCING_HIT},
    {"LB_TEMPERATURE_BUOYANCY", CH4G_TEMPERATURE_BUOYANCY},
    {"LB_FLUID_FORCING_NOZEROMODE", CH4G_FLUID_FORCING_NOZEROMODE},
    {"LB_FLUID_BC_Y", CH4G_FLUID_BC_Y},
    {"LB_TEMPERATURE_BC_Y", CH4G_TEMPERATURE_BC_Y},
    {"LB_FLUID_LES", CH4G_FLUID_LES},
    {"LB_FLUID_LES_LOCAL_STRAIN", CH4G_FLUID_LES_LOCAL_STRAIN},
    {"LB_TEMPERATURE_BUOYANCY_T0_REF", CH4G_TEMPERATURE_BUOYANCY_T0_REF},
    {"LB_TEMPERATURE_INITIAL_LINEAR", CH4G_TEMPERATURE_INITIAL_LINEAR},
    {"LB_TEMPERATURE_INITIAL_ADD_NOISE", CH4G_TEMPERATURE_INITIAL_ADD_NOISE},
    {"LB_FLUID_INITIAL_UNIT_DENSITY", CH4G_FLUID_INITIAL_UNIT_DENSITY},
    {"LB_INITIAL_BAROMETRIC", CH4G_INITIAL_BAROMETRIC},
    {"LB_TEMPERATURE_LES", CH4G_TEMPERATURE_LES},
};

/* particle switches (define.h:316-382) onto ch4g_lagrange_flag; umbrella
 * switches (LAGRANGE_SMALLTAUD, LAGRANGE_TWOWAY) are resolved below      */
static const struct { const char *name; unsigned flag; } LFLAGS[] = {
    {"LAGRANGE_ADD_TRACER", CH4G_LAGRANGE_ADD_TRACER},
    {"LAGRANGE_WRAP", CH4G_LAGRANGE_WRAP},
    {"LAGRANGE_GRADIENT", CH4G_LAGRANGE_GRADIENT},
    {"LAGRANGE_ADDEDMASS", CH4G_LAGRANGE_ADDEDMASS},
    {"LAGRANGE_GRAVITY", CH4G_LAGRANGE_GRAVITY},
    {"LAGRANGE_BC_INELASTIC", CH4G_LAGRANGE_BC_INELASTIC},
    {"LAGRANGE_BC_INELASTIC_REINJECT", CH4G_LAGRANGE_BC_INELASTIC_REINJECT},
    {"LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM", CH4G_LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM},
    {"LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO", CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO},
    {"LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF", CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF},
    {"LAGRANGE_SMALLTAUD_BETA", CH4G_LAGRANGE_SMALLTAUD_BETA},
    {"LAGRANGE_SMALLTAUD_FLOATER", CH4G_LAGRANGE_SMALLTAUD_FLOATER},
    {"LAGRANGE_TWOWAY_MOMENTUM", CH4G_LAGRANGE_TWOWAY_MOMENTUM},
    {"LAGRANGE_TWOWAY_SMALLTAUD", CH4G_LAGRANGE_TWOWAY_SMALLTAUD},
    {"LAGRANGE_TAUDRAG_INCREMENT_LOG", CH4G_LAGRANGE_TAUDRAG_INCREMENT_LOG},
    {"LAGRANGE_ADDEDMASS_INCREMENT_LOG", CH4G_LAGRANGE_ADDEDMASS_INCREMENT_LOG},
    {"LAGRANGE_ORIENTATION", CH4G_LAGRANGE_ORIENTATION},
    {"LAGRANGE_ORIENTATION_JEFFREY", CH4G_LAGRANGE_ORIENTATION_JEFFREY},
    {"LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG", CH4G_LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG},
    {"LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS", CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS},
    {"LAGRANGE_ORIENTATION_ACTIVE", CH4G_LAGRANGE_ORIENTATION_ACTIVE},
    {"LAGRANGE_ORIENTATION_BC", CH4G_LAGRANGE_ORIENTATION_BC},
    {"LAGRANGE_ORIENTATION_SECONDORIENTATION", CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION},
    {"LAGRANGE_TEMPERATURE", CH4G_LAGRANGE_TEMPERATURE},
    {"LAGRANGE_TEMPERATURE_INCREMENT_LOG", CH4G_LAGRANGE_TEMPERATURE_INCREMENT_LOG},
    {"LAGRANGE_INITIAL_TEMPERATURE_FLUID", CH4G_LAGRANGE_INITIAL_TEMPERATURE_FLUID},
    {"LAGRANGE_INITIAL_VELOCITY_FLUID", CH4G_LAGRANGE_INITIAL_VELOCITY_FLUID},
    {"LAGRANGE_INITIAL_VELOCITY_STOKES", CH4G_LAGRANGE_INITIAL_VELOCITY_STOKES},
    {"LAGRANGE_GRAVITY_VARIABLE", CH4G_LAGRANGE_GRAVITY_VARIABLE},
    {"LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG", CH4G_LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG},
    {"LAGRANGE_OUTPUT_FLUID_AVERAGES", CH4G_LAGRANGE_OUTPUT_FLUID_AVERAGES},
    {"LAGRANGE_ADD_TRACER_STATIC", CH4G_LAGRANGE_ADD_TRACER_STATIC},
};

/* the second word: initial positions */
static const struct { const char *name; unsigned flag; } LFLAGS2[] = {
    {"LAGRANGE_INITIAL_GRID", CH4G_LAGRANGE_INITIAL_GRID},
    {"LAGRANGE_INITIAL_PAIRS", CH4G_LAGRANGE_INITIAL_PAIRS},
    {"LAGRANGE_ADDEDMASS_LIFT", CH4G_LAGRANGE_ADDEDMASS_LIFT},
    {"LAGRANGE_ADDEDMASS_WAKEDRAG", CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG},
    {"LAGRANGE_ORIENTATION_DRAG", CH4G_LAGRANGE_ORIENTATION_DRAG},
    {"LAGRANGE_RADIUSandDENSITY", CH4G_LAGRANGE_RADIUSANDDENSITY},
    {"LAGRANGE_RADIUSandDENSITY_INCREMENT_LOG_RADIUS", CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS},
    {"LAGRANGE_RADIUSandDENSITY_INCREMENT_LOG_DENSITY", CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY},
    {"LAGRANGE_BC_RADIUSandDENSITY", CH4G_LAGRANGE_BC_RADIUSANDDENSITY},
    {"LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC", CH4G_LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC},
    {"LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE", CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE},
    {"LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION", CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION},
    {"LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION", CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION},
    {"LAGRANGE_BC_INELASTIC_SEDIMENT", CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT},
    {"LAGRANGE_ORIENTATION_RANDOM", CH4G_LAGRANGE_ORIENTATION_RANDOM},
    {"LAGRANGE_POLYMER", CH4G_LAGRANGE_POLYMER},
    {"LAGRANGE_ORIENTATION_ACTIVE_JUMP", CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP},
    {"LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS", CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS},
    {"LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE", CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE},
};

/* switches with no effect on the streaming step of these configurations
 * (build, I/O, timing, umbrella names, choices implied by METHOD_STREAMING) */
static const char *NEUTRAL[] = {
    "NEW_SENDRECV", "OUTPUT_ASCII", "OUTPUT_H5", "OUTPUT_H5_GRID", "TIMING", "SYSTEM_TIMING", "DEBUG",
    "GRID_REGULAR", "LB", "LB_FLUID_FORCING", "LB_FLUID_BC", "LB_TEMPERATURE_BC", "LB_FLUID_FORCING_GRAVITY",
    "METHOD_EDGES_AND_CORNERS", "METHOD_FORCING_GUO", "METHOD_STEPPING_EULER", "METHOD_MYQUICK",
    "METHOD_MYQUICK_CARTESIAN", "SYSTEM_RANDOM48", "PRECISION_DOUBLE", "OUTPUT_AT_START", "OUTPUT_DIAGN_APPEND",
    "METHOD_STREAMING", "GRID_POP_D3Q19", "NO_XMLHEADERS",
    /* particles: umbrella switches, debug output, and the two-node halo the
       particle fields always use here */
    "LAGRANGE", "LAGRANGE_SMALLTAUD", "LAGRANGE_TWOWAY", "LAGRANGE_GRADIENT_BRD2", "LAGRANGE_DEBUG",
    "LAGRANGE_OUTPUT_DEBUG", "LAGRANGE_OUTPUT_FLUID",
};

static int in_list(const char *n, const char *const *l, size_t k)
{
    for (size_t i = 0; i < k; i++)
        if (!strcmp(n, l[i])) return 1;
    return 0;
}

/* active #define names (block and line comments skipped) */
static int read_defines(const char *path, char names[][64], int maxn)
{
    FILE *f = fopen(path, "r");
    if (!f) return perr("%s file is missing", path);
    int n = 0, in_block = 0;
    char line[1024];
    while (fgets(line, sizeof line, f)) {
        char buf[1024];
        size_t o = 0;
        for (size_t i = 0; line[i] && o + 1 < sizeof buf; i++) {
            if (in_block) {
                if (line[i] == '*' && line[i + 1] == '/') {
                    in_block = 0;
                    i++;
                }
                continue;
            }
            if (line[i] == '/' && line[i + 1] == '*') {
                in_block = 1;
                i++;
                continue;
            }
            if (line[i] == '/' && line[i + 1] == '/') break;
            buf[o++] = line[i];
        }
        buf[o] = 0;
        char *p = buf;
        while (isspace((unsigned char)*p)) p++;
        if (strncmp(p, "#define", 7) || !isspace((unsigned char)p[7])) continue;
        p += 7;
        while (isspace((unsigned char)*p)) p++;
        size_t k = 0;
        while ((isalnum((unsigned char)p[k]) || p[k] == '_') && k < 63) k++;
        if (!k || n >= maxn) continue;
        memcpy(names[n], p, k);
        names[n][k] = 0;
        n++;
    }
    fclose(f);
    return n;
}

/* param.in tokens, scanned like read_parameter's fscanf("%s %lf") loop */
typedef struct { char name[256]; double val; } kv;

static int read_param_file(const char *path, kv *out, int maxn)
{
    FILE *f = fopen(path, "r");
    if (!f) return perr("%s file is missing", path);
    int n = 0;
    char tok[256];
    char pending[256] = "";
    int have_pending = 0;
    for (;;) {
        if (have_pending) {
            strcpy(tok, pending);
            have_pending = 0;
        } else if (fscanf(f, "%255s", tok) != 1) {
            break;
        }
        double v = 0.0;
        char nxt[256];
        if (fscanf(f, "%255s", nxt) == 1) {
            char *end;
            errno = 0;
            v = strtod(nxt, &end);
            if (end == nxt || *end) { /* not a number: it is the next name */
                v = 0.0;
                strcpy(pending, nxt);
                have_pending = 1;
            }
        }
        if (n < maxn) {
            snprintf(out[n].name, sizeof out[n].name, "%s", tok);
            out[n].val = v;
            n++;
        }
    }
    fclose(f);
    return n;
}

static int getp(const kv *t, int n, const char *name, double *v)
{
    for (int i = 0; i < n; i++)
        if (!strcmp(t[i].name, name)) {
            *v = t[i].val;
            return 1;
        }
    return 0;
}

static int need(const kv *t, int n, const char *name, double *v)
{
    if (!getp(t, n, name, v)) return perr("%s not found. End of File", name);
    return 0;
}

int ch4g_read_run_params(const char *define_h, const char *param_in, ch4g_run_params *out)
{
    static char names[512][64];
    int nd = read_defines(define_h, names, 512);
    if (nd < 0) return -1;
    memset(out, 0, sizeof *out);
    ch4g_config *c = &out->cfg;
    int streaming = 0, d3q19 = 0, forcing_umbrella = 0, smalltaud = 0, twoway = 0;
    for (int i = 0; i < nd; i++) {
        const char *nm = names[i];
        int known = 0;
        for (size_t k = 0; k < sizeof FLAGS / sizeof FLAGS[0]; k++)
            if (!strcmp(nm, FLAGS[k].name)) {
                c->flags |= FLAGS[k].flag;
                known = 1;
            }
        for (size_t k = 0; k < sizeof LFLAGS / sizeof LFLAGS[0]; k++)
            if (!strcmp(nm, LFLAGS[k].name)) {
                out->pcfg.flags |= LFLAGS[k].flag;
                known = 1;
            }
        for (size_t k = 0; k < sizeof LFLAGS2 / sizeof LFLAGS2[0]; k++)
            if (!strcmp(nm, LFLAGS2[k].name)) {
                out->pcfg.flags2 |= LFLAGS2[k].flag;
                known = 1;
            }
        if (!strcmp(nm, "LAGRANGE")) out->lagrange = 1;
        if (!strcmp(nm, "LAGRANGE_SMALLTAUD")) smalltaud = 1;
        if (!strcmp(nm, "LAGRANGE_TWOWAY")) twoway = 1;
        if (!strcmp(nm, "LB_FLUID_FORCING")) forcing_umbrella = 1;
        if (!strcmp(nm, "METHOD_STREAMING")) streaming = 1;
        if (!strcmp(nm, "GRID_POP_D3Q19")) d3q19 = 1;
        if (!strcmp(nm, "OUTPUT_AT_START")) out->output_at_start = 1;
        if (!strcmp(nm, "OUTPUT_DIAGN_APPEND")) out->diagn_append = 1;
        if (!strcmp(nm, "OUTPUT_H5")) out->output_h5 = 1;
        if (!strcmp(nm, "LAGRANGE_OUTPUT_FLUID")) out->lagr_output_fluid = 1;
        if (!strncmp(nm, "GRID_POP_", 9) && strcmp(nm, "GRID_POP_D3Q19"))
            return perr("%.63s: only GRID_POP_D3Q19 is implemented on the device lattice", nm);
        if (!known && !in_list(nm, NEUTRAL, sizeof NEUTRAL / sizeof NEUTRAL[0]))
            return perr("define.h switch %.63s changes the step in a way this lattice does not implement", nm);
    }
    if (!streaming) return perr("%s", "METHOD_STREAMING is required");
    if (!d3q19) return perr("%s", "GRID_POP_D3Q19 is required");
    /* the small-tau variants act inside #ifdef LAGRANGE_SMALLTAUD
       (lagrange.c:3424), the feedback inside #ifdef LAGRANGE_TWOWAY of
       lagrange_feedback.c and only when build_forcing runs (main.c:118-131) */
    if (!smalltaud) out->pcfg.flags &= ~(unsigned)(CH4G_LAGRANGE_SMALLTAUD_BETA | CH4G_LAGRANGE_SMALLTAUD_FLOATER);
    if (!twoway || !forcing_umbrella)
        out->pcfg.flags &= ~(unsigned)(CH4G_LAGRANGE_TWOWAY_MOMENTUM | CH4G_LAGRANGE_TWOWAY_SMALLTAUD);
    /* the orientation variants sit inside #ifdef LAGRANGE_ORIENTATION, the
       aspect-ratio ones inside LAGRANGE_ORIENTATION_JEFFREY (lagrange.c:3470-3752) */
    if (!(out->pcfg.flags & CH4G_LAGRANGE_ORIENTATION))
        out->pcfg.flags &= ~(unsigned)(CH4G_LAGRANGE_ORIENTATION_JEFFREY | CH4G_LAGRANGE_ORIENTATION_ACTIVE |
                                       CH4G_LAGRANGE_ORIENTATION_BC |
                                       CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION);
    if (!(out->pcfg.flags & CH4G_LAGRANGE_ORIENTATION_JEFFREY))
        out->pcfg.flags &= ~(unsigned)(CH4G_LAGRANGE_ORIENTATION_JEFFREY_INCREMENT_LOG |
                                       CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS);
    /* lift and wake drag sit inside #ifdef LAGRANGE_ADDEDMASS (lagrange.c:3241, :3325) */
    if (!((out->pcfg.flags & CH4G_LAGRANGE_GRADIENT) && (out->pcfg.flags & CH4G_LAGRANGE_ADDEDMASS)))
        out->pcfg.flags2 &= ~(unsigned)(CH4G_LAGRANGE_ADDEDMASS_LIFT | CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG);
    /* the spheroid drag reads aspect_ratio (LAGRANGE_ORIENTATION_JEFFREY, typedef.h:423-424)
       and replaces the Stokes drag with its wake correction (lagrange.c:3234-3268) */
    if (out->pcfg.flags2 & CH4G_LAGRANGE_ORIENTATION_DRAG) {
        if (!(out->pcfg.flags & CH4G_LAGRANGE_ORIENTATION_JEFFREY))
            return perr("%s", "LAGRANGE_ORIENTATION_DRAG needs LAGRANGE_ORIENTATION_JEFFREY");
        out->pcfg.flags2 &= ~(unsigned)CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG;
    }
    /* the log ladders act inside LAGRANGE_RADIUSandDENSITY (lagrange.c:160, :199); the
       wall radius needs it and the added mass (lagrange.c:3777-3786) */
    if (!(out->pcfg.flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY))
        out->pcfg.flags2 &= ~(unsigned)(CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_RADIUS |
                                        CH4G_LAGRANGE_RADIUSANDDENSITY_INCREMENT_LOG_DENSITY);
    if (!((out->pcfg.flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY) && (out->pcfg.flags & CH4G_LAGRANGE_GRADIENT) &&
          (out->pcfg.flags & CH4G_LAGRANGE_ADDEDMASS)))
        out->pcfg.flags2 &= ~(unsigned)CH4G_LAGRANGE_BC_RADIUSANDDENSITY;
    /* the swimming modes sit inside #ifdef LAGRANGE_ORIENTATION_ACTIVE (lagrange.c:3134-3178) */
    {
        const unsigned modes = CH4G_LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC | CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE |
                               CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION |
                               CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION |
                               CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP | CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS |
                               CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE;
        if (!((out->pcfg.flags & CH4G_LAGRANGE_ORIENTATION_ACTIVE) && (out->pcfg.flags & CH4G_LAGRANGE_GRADIENT)))
            out->pcfg.flags2 &= ~modes;
        /* the jump's variants sit inside #ifdef LAGRANGE_ORIENTATION_ACTIVE_JUMP (lagrange.c:3052-3087) */
        if (!(out->pcfg.flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP))
            out->pcfg.flags2 &= ~(unsigned)(CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS |
                                            CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE);
    }
    /* random reorientation sits inside #ifdef LAGRANGE_ORIENTATION (lagrange.c:3698) */
    if (!((out->pcfg.flags & CH4G_LAGRANGE_ORIENTATION) && (out->pcfg.flags & CH4G_LAGRANGE_GRADIENT)))
        out->pcfg.flags2 &= ~(unsigned)CH4G_LAGRANGE_ORIENTATION_RANDOM;
    /* the conformation tensor sits inside #ifdef LAGRANGE_GRADIENT (lagrange.c:3998-4002) */
    if (!(out->pcfg.flags & CH4G_LAGRANGE_GRADIENT)) out->pcfg.flags2 &= ~(unsigned)CH4G_LAGRANGE_POLYMER;
    /* the deposition count sits inside #ifdef LAGRANGE_BC_INELASTIC (typedef.h:453-457) */
    if (!(out->pcfg.flags & CH4G_LAGRANGE_BC_INELASTIC)) out->pcfg.flags2 &= ~(unsigned)CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT;
    /* the Eulerian probe type sits inside #ifdef LAGRANGE_ADD_TRACER (parameters.c:744-755) */
    if (!(out->pcfg.flags & CH4G_LAGRANGE_ADD_TRACER)) out->pcfg.flags &= ~(unsigned)CH4G_LAGRANGE_ADD_TRACER_STATIC;
    /* variable gravity sits inside #ifdef LAGRANGE_GRAVITY, its log ladder
       inside LAGRANGE_GRAVITY_VARIABLE (parameters.c:589-590, lagrange.c:347-376) */
    if (!(out->pcfg.flags & CH4G_LAGRANGE_GRAVITY))
        out->pcfg.flags &= ~(unsigned)CH4G_LAGRANGE_GRAVITY_VARIABLE;
    if (!(out->pcfg.flags & CH4G_LAGRANGE_GRAVITY_VARIABLE))
        out->pcfg.flags &= ~(unsigned)CH4G_LAGRANGE_GRAVITY_VARIABLE_INCREMENT_LOG;
    /* the cp ladder's log steps and the fluid-temperature start sit inside
       #ifdef LAGRANGE_TEMPERATURE (lagrange.c:587, :1005) */
    if (!(out->pcfg.flags & CH4G_LAGRANGE_TEMPERATURE))
        out->pcfg.flags &= ~(unsigned)(CH4G_LAGRANGE_TEMPERATURE_INCREMENT_LOG |
                                       CH4G_LAGRANGE_INITIAL_TEMPERATURE_FLUID);
    /* point_particle has t_p only with LB_TEMPERATURE (typedef.h:467-470) */
    if ((out->pcfg.flags & CH4G_LAGRANGE_TEMPERATURE) && !(c->flags & CH4G_LB_TEMPERATURE))
        return perr("%s", "LAGRANGE_TEMPERATURE needs LB_TEMPERATURE");
    /* its relaxation rate reads beta_coeff (typedef.h:408-413) */
    if ((out->pcfg.flags & CH4G_LAGRANGE_TEMPERATURE) &&
        (~out->pcfg.flags & (CH4G_LAGRANGE_GRADIENT | CH4G_LAGRANGE_ADDEDMASS)))
        return perr("%s", "LAGRANGE_TEMPERATURE needs LAGRANGE_GRADIENT and LAGRANGE_ADDEDMASS");
    /* the Stokes start reads beta_coeff too (lagrange.c:981-991) */
    if ((out->pcfg.flags & CH4G_LAGRANGE_INITIAL_VELOCITY_STOKES) &&
        (~out->pcfg.flags & (CH4G_LAGRANGE_GRADIENT | CH4G_LAGRANGE_ADDEDMASS)))
        return perr("%s", "LAGRANGE_INITIAL_VELOCITY_STOKES needs LAGRANGE_GRADIENT and LAGRANGE_ADDEDMASS");
    if (!out->lagrange && (out->pcfg.flags || out->pcfg.flags2))
        return perr("%s", "LAGRANGE_* switches without LAGRANGE");

    static kv t[1024];
    int np = read_param_file(param_in, t, 1024);
    if (np < 0) return -1;
    double v, SX, SY, SZ;
    /* parameters.c:78-88 */
    if (need(t, np, "resume", &v)) return -1;
    out->resume = (int)v;
    if (c->flags & CH4G_LB_FLUID) {
        if (need(t, np, "resume_u", &v)) return -1;
        out->resume_u = (int)v;
    }
    if (c->flags & CH4G_LB_TEMPERATURE) {
        if (need(t, np, "resume_t", &v)) return -1;
        out->resume_t = (int)v;
    }
    if (need(t, np, "NX", &v)) return -1;
    c->NX = (int)v;
    if (need(t, np, "NY", &v)) return -1;
    c->NY = (int)v;
    if (need(t, np, "NZ", &v)) return -1;
    c->NZ = (int)v;
    if (need(t, np, "SX", &SX) || need(t, np, "SY", &SY) || need(t, np, "SZ", &SZ)) return -1;
    if (need(t, np, "time_dt", &out->time_dt) || need(t, np, "time_start", &out->time_start) ||
        need(t, np, "time_end", &out->time_end) || need(t, np, "time_dump_field", &out->time_dump_field) ||
        need(t, np, "time_dump_diagn", &out->time_dump_diagn))
        return -1;
    /* parameters.c:135: under streaming the time step equals the spacing */
    if (out->time_dt != SX / c->NX || SX / c->NX != 1.0 || SY / c->NY != 1.0 || SZ / c->NZ != 1.0)
        return perr("%s", "time_dt != Delta X or Y or Z (streaming needs unit spacing and time step)");
    if (c->flags & CH4G_LB_FLUID) {
        if (need(t, np, "tau_u", &c->tau_u)) return -1;
        if (c->flags & (CH4G_FLUID_FORCING_POISEUILLE | CH4G_FLUID_FORCING_CHANNEL | CH4G_FLUID_FORCING_HIT |
                        CH4G_TEMPERATURE_BUOYANCY)) {
            if (need(t, np, "Amp_x", &c->Amp_x) || need(t, np, "Amp_y", &c->Amp_y) || need(t, np, "Amp_z", &c->Amp_z))
                return -1;
        }
    }
    if (c->flags & CH4G_LB_TEMPERATURE) {
        if (need(t, np, "tau_t", &c->tau_t) || need(t, np, "T_bot", &c->T_bot) || need(t, np, "T_top", &c->T_top) ||
            need(t, np, "T_ref", &c->T_ref))
            return -1;
        if (c->flags & CH4G_TEMPERATURE_BUOYANCY) {
            if (need(t, np, "beta_t", &c->beta_t) || need(t, np, "beta2_t", &c->beta2_t) ||
                need(t, np, "gravity_x", &c->gravity_x) || need(t, np, "gravity_y", &c->gravity_y) ||
                need(t, np, "gravity_z", &c->gravity_z))
                return -1;
        }
        if (c->flags & CH4G_TEMPERATURE_INITIAL_ADD_NOISE)
            if (need(t, np, "noise_t", &c->noise_t)) return -1;
    }
    if (out->lagrange) {
        ch4g_particle_config *pc = &out->pcfg;
        double v2;
        /* parameters.c:155-167, :524-584 */
        if (need(t, np, "time_dump_lagr", &out->time_dump_lagr) ||
            need(t, np, "time_dt_lagr_factor", &pc->time_dt_lagr_factor) || need(t, np, "particle_number", &v2))
            return -1;
        pc->particle_number = (int32_t)v2;
        if (pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY) { /* parameters.c:530-556 */
            if (need(t, np, "particle_radius_types", &pc->particle_radius_types) ||
                need(t, np, "particle_radius_min", &pc->particle_radius_min) ||
                need(t, np, "particle_radius_max", &pc->particle_radius_max) ||
                need(t, np, "particle_density_types", &pc->particle_density_types) ||
                need(t, np, "particle_density_min", &pc->particle_density_min) ||
                need(t, np, "particle_density_max", &pc->particle_density_max))
                return -1;
        } else if (need(t, np, "tau_drag_types", &pc->tau_drag_types) ||
                   need(t, np, "tau_drag_min", &pc->tau_drag_min) || need(t, np, "tau_drag_max", &pc->tau_drag_max)) {
            return -1;
        }
        if (!(pc->flags2 & CH4G_LAGRANGE_RADIUSANDDENSITY) && (pc->flags & CH4G_LAGRANGE_GRADIENT) &&
            (pc->flags & CH4G_LAGRANGE_ADDEDMASS)) {
            if (need(t, np, "beta_coeff_types", &pc->beta_coeff_types) ||
                need(t, np, "beta_coeff_min", &pc->beta_coeff_min) || need(t, np, "beta_coeff_max", &pc->beta_coeff_max))
                return -1;
        }
        /* parameters.c:621-695 */
        const unsigned fl = pc->flags;
        /* parameters.c:696-722 */
        if ((pc->flags2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) &&
            (need(t, np, "critical_shear_rate_types", &pc->critical_shear_rate_types) ||
             need(t, np, "critical_shear_rate_min", &pc->critical_shear_rate_min) ||
             need(t, np, "critical_shear_rate_max", &pc->critical_shear_rate_max) ||
             need(t, np, "jump_time_types", &pc->jump_time_types) || need(t, np, "jump_time_min", &pc->jump_time_min) ||
             need(t, np, "jump_time_max", &pc->jump_time_max)))
            return -1;
        /* parameters.c:729-733 */
        if ((pc->flags2 & CH4G_LAGRANGE_POLYMER) && need(t, np, "tau_polymer", &pc->tau_polymer)) return -1;
        /* parameters.c:589-600 */
        if ((fl & CH4G_LAGRANGE_GRAVITY_VARIABLE) &&
            (need(t, np, "gravity_coeff_types", &pc->gravity_coeff_types) ||
             need(t, np, "gravity_coeff_min", &pc->gravity_coeff_min) ||
             need(t, np, "gravity_coeff_max", &pc->gravity_coeff_max)))
            return -1;
        if ((fl & CH4G_LAGRANGE_GRADIENT) && (fl & CH4G_LAGRANGE_ORIENTATION)) {
            if ((fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY) &&
                (need(t, np, "aspect_ratio_types", &pc->aspect_ratio_types) ||
                 need(t, np, "aspect_ratio_min", &pc->aspect_ratio_min) ||
                 need(t, np, "aspect_ratio_max", &pc->aspect_ratio_max)))
                return -1;
            if ((fl & CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS) &&
                (need(t, np, "gyrotaxis_velocity_types", &pc->gyrotaxis_velocity_types) ||
                 need(t, np, "gyrotaxis_velocity_min", &pc->gyrotaxis_velocity_min) ||
                 need(t, np, "gyrotaxis_velocity_max", &pc->gyrotaxis_velocity_max)))
                return -1;
            if ((fl & CH4G_LAGRANGE_ORIENTATION_ACTIVE) &&
                (need(t, np, "swim_velocity_types", &pc->swim_velocity_types) ||
                 need(t, np, "swim_velocity_min", &pc->swim_velocity_min) ||
                 need(t, np, "swim_velocity_max", &pc->swim_velocity_max)))
                return -1;
        }
        /* parameters.c:606-616 */
        if ((fl & CH4G_LAGRANGE_TEMPERATURE) &&
            (need(t, np, "cp_coeff_types", &pc->cp_coeff_types) || need(t, np, "cp_coeff_min", &pc->cp_coeff_min) ||
             need(t, np, "cp_coeff_max", &pc->cp_coeff_max)))
            return -1;
    }
    /* keys of this implementation (optional) */
    c->precision = CH4G_FP64;
    if (getp(t, np, "precision", &v)) c->precision = (int)v;
    c->les_cs = 0.18; /* les.c:10 */
    if (getp(t, np, "les_cs", &v)) c->les_cs = v;
    c->rho0 = 1.0;
    c->T0 = 0.5 * (c->T_bot + c->T_top);
    /* parallel.c:238-240 seeds srand48 with rank + time(NULL); `seed` fixes it */
    c->seed = getp(t, np, "seed", &v) ? (int)v : (int)time(NULL);
    return 0;
}
