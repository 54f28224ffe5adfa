// C-ABI of include/ch4g.h: device-resident lattice lifecycle, initial
// conditions, the per-step launch sequence and the host-side pieces of the
// reference step (sound-speed constants, forcing constants, HIT phase walk on
// the drand48 stream), all computed with the same libm calls as the reference
// so that the FP64 device path stays bit-identical to it.
//
// Compiled with -ffp-contract=off (host arithmetic included).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "ch4g_lattice.h"

using namespace ch4g;

namespace ch4g {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace ch4g

namespace ch4g {

int dalloc(ch4g_lattice *L, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(-3, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    L->bytes += (int64_t)bytes;
    return 0;
}

}  // namespace ch4g

namespace {

bool has(const ch4g_config &c, uint32_t f) { return (c.flags & f) != 0; }

// Smagorinsky closure from finite differences of u (the reference's default
// LB_FLUID_LES form, les.c:31-43)
bool les_fd(const ch4g_config &c)
{
    return has(c, CH4G_FLUID_LES) && !has(c, CH4G_FLUID_LES_LOCAL_STRAIN) && has(c, CH4G_LB_FLUID);
}

bool forced(const ch4g_config &c)
{
    return has(c, CH4G_FLUID_FORCING_POISEUILLE | CH4G_FLUID_FORCING_CHANNEL | CH4G_FLUID_FORCING_HIT |
                      CH4G_TEMPERATURE_BUOYANCY);
}

// random_gauss (random.c:82-100) on the lattice's drand48 stream.
double random_gauss(unsigned short xs[3], double mu, double sigma)
{
    double U1, U2, W, mult, X1;
    do {
        U1 = -1 + erand48(xs) * 2;
        U2 = -1 + erand48(xs) * 2;
        W = pow(U1, 2) + pow(U2, 2);
    } while (W >= 1 || W == 0);
    mult = sqrt((-2 * log(W)) / W);
    X1 = U1 * mult;
    return (mu + sigma * X1);
}

// HIT phase Ornstein-Uhlenbeck step of build_forcing (forcing.c:40-69).
void hit_walk(ch4g_lattice *L)
{
    const double LX = L->cfg.NX;
    double t_turnover = (LX / 0.1);
    double phi_diffusivity = 0.01 / t_turnover;
    double fac1 = 1.0 / t_turnover;
    double fac2 = sqrt(2.0 * 1.0 * phi_diffusivity);
    for (int ii = 0; ii < CH4G_HIT_NK; ii++)
        for (int a = 0; a < 3; a++) {
            double val = random_gauss(L->xsubi, 0.0, 1.0);
            L->phi[ii][a] += -L->phi[ii][a] * fac1 + val * fac2;
        }
}

// sin(2 pi (k x / L + phi)) tables of the current phases (forcing.c:582-584).
void hit_tables(const ch4g_lattice *L, double *h)
{
    const int NX = L->cfg.NX, NY = L->cfg.NY, NZ = L->cfg.NZ;
    const double LX = NX, LY = NY, LZ = NZ;
    double *sx = h, *sy = h + CH4G_HIT_NK * NX, *sz = sy + CH4G_HIT_NK * NY;
    for (int ii = 0; ii < CH4G_HIT_NK; ii++) {
        for (int i = 0; i < NX; i++) {
            double x = (double)i + 0.5;
            sx[ii * NX + i] = sin(2 * 3.14159265359 * (L->vk[ii][0] * x / LX + L->phi[ii][0]));
        }
        for (int j = 0; j < NY; j++) {
            double y = (double)j + 0.5;
            sy[ii * NY + j] = sin(2 * 3.14159265359 * (L->vk[ii][1] * y / LY + L->phi[ii][1]));
        }
        for (int k = 0; k < NZ; k++) {
            double z = (double)k + 0.5;
            sz[ii * NZ + k] = sin(2 * 3.14159265359 * (L->vk[ii][2] * z / LZ + L->phi[ii][2]));
        }
    }
}

// Upload the tables of the current phases into device slot `slot`.
int hit_upload(ch4g_lattice *L, int slot)
{
    int r = L->ring;
    L->ring = (L->ring + 1) % HIT_RING;
    HIPCHK(hipEventSynchronize(L->ev_ring[r]));
    hit_tables(L, L->h_ring[r]);
    HIPCHK(hipMemcpyAsync(L->d_hit[slot], L->h_ring[r], L->hit_len * sizeof(double), hipMemcpyHostToDevice,
                          L->stream));
    HIPCHK(hipEventRecord(L->ev_ring[r], L->stream));
    return 0;
}

// build_forcing at initialisation (main.c:38): one phase step, tables of
// Phi_1 into slot 1 (the force of step 0).
int first_build_forcing(ch4g_lattice *L)
{
    if (!has(L->cfg, CH4G_FLUID_FORCING_HIT)) return 0;
    hit_walk(L);
    return hit_upload(L, (int)((L->itime + 1) & 1));
}

}  // namespace

namespace ch4g {

KParams params_for(const ch4g_lattice *L, int first, int itime)
{
    KParams P = L->base;
    P.first = first;
    P.itime = itime;
    if (P.twoway) {
        // feedback of step s lives in plane set s & 1 (zero at s = 0, 1):
        // double planes on FP64 lattices, float planes on float storage
        if (L->dev32) {
            P.fbf_prev = (const float *)L->part->d_fb[(itime + 1) & 1];
            P.fbf_now = (const float *)L->part->d_fb[itime & 1];
        } else {
            P.fb_prev = (const double *)L->part->d_fb[(itime + 1) & 1];
            P.fb_now = (const double *)L->part->d_fb[itime & 1];
        }
    }
    if (P.hit) {
        const int NX = L->cfg.NX, NY = L->cfg.NY;
        int sp = itime & 1, sn = (itime + 1) & 1;
        P.hit_prev.sx = L->d_hit[sp];
        P.hit_prev.sy = L->d_hit[sp] + CH4G_HIT_NK * NX;
        P.hit_prev.sz = L->d_hit[sp] + CH4G_HIT_NK * (NX + NY);
        P.hit_now.sx = L->d_hit[sn];
        P.hit_now.sy = L->d_hit[sn] + CH4G_HIT_NK * NX;
        P.hit_now.sz = L->d_hit[sn] + CH4G_HIT_NK * (NX + NY);
    }
    return P;
}

int ensure_scratch(ch4g_lattice *L)
{
    const size_t N = (size_t)L->N;
    int rc;
    if (!L->d_dens) {
        if ((rc = dalloc(L, (void **)&L->d_dens, N * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&L->d_u, 3 * N * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&L->d_t, N * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&L->d_part, (size_t)reduce_blocks(L->base) * NOUT * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&L->d_res, NOUT * sizeof(double)))) return rc;
    }
    return 0;
}

int ensure_stage(ch4g_lattice *L)
{
    const size_t N = (size_t)L->N;
    int rc;
    if (!L->d_fs) {
        if ((rc = dalloc(L, (void **)&L->d_fs, (size_t)Q * N * sizeof(double)))) return rc;
        if ((rc = dalloc(L, (void **)&L->d_gs, (size_t)Q * N * sizeof(double)))) return rc;
    }
    return 0;
}

// Gather of the current state into FP64 scratch (any output may be null).
void gather(ch4g_lattice *L, double *fpost, double *gpost, double *dens, double *u, double *t)
{
    KParams P = params_for(L, L->first, (int)L->itime);
    if (L->d_f0) {
        // the mean force of the last step (its hydro fields carry its half force)
        if (L->dev32)
            launch_force_mean32(P, false, nullptr, (const float *)L->tlag, L->d_f0 + 6, L->d_f0, L->stream);
        else
            launch_force_mean64(P, false, nullptr, (const double *)L->tlag, L->d_f0 + 6, L->d_f0, L->stream);
        P.f0n = L->d_f0 + 3;
    }
    if (L->dev32)
        launch_gather_fp32(P, (const float *)L->f[L->cur], (const float *)L->g[L->cur], (const float *)L->tlag, fpost,
                           gpost, dens, u, t, L->stream);
    else
        launch_gather_fp64(P, (const double *)L->f[L->cur], (const double *)L->g[L->cur],
                           (const double *)L->tlag, fpost, gpost, dens, u, t, L->stream);
}

// Graphs capture kernel arguments: drop them when the step sequence changes.
void destroy_graphs(ch4g_lattice *L)
{
    for (int b = 0; b < 4; b++)
        if (L->graph[b]) {
            (void)hipGraphExecDestroy(L->graph[b]);
            L->graph[b] = nullptr;
        }
}

}  // namespace ch4g

namespace {

// One step: host forcing tables (HIT), then the fused kernel.
int launch_one_step(ch4g_lattice *L)
{
    int itime = (int)L->itime + 1;
    if (L->base.hit) {
        // build_forcing of this step: Phi_{itime+1}
        if (L->part && L->part->rng_on_device) {
            int rc = rng_to_host(L);
            if (rc) return rc;
        }
        hit_walk(L);
        if (L->part && L->part->rng_on_device) {
            int rc = rng_to_device(L);
            if (rc) return rc;
        }
        int rc = hit_upload(L, (itime + 1) & 1);
        if (rc) return rc;
    }
    if (L->part) {
        int rc = particles_before_step(L, itime);
        if (rc) return rc;
    }
    KParams P = params_for(L, L->first, itime);
    // the particles of the previous step move on the fields this kernel gathers
    const bool fused = L->part && L->part->pending;
    if (fused) {
        int rc = ensure_float_fields(L);
        if (rc) return rc;
        P.uT_out = L->d_uT32;
    }
    int in = L->cur, out = 1 - L->cur;
    if (L->d_f0) {
        // NOZEROMODE: mean force of the previous step and of this one first
        if (L->dev32)
            launch_force_mean32(P, true, (const float *)L->g[in], (const float *)L->tlag, L->d_f0 + 6, L->d_f0,
                                L->stream);
        else
            launch_force_mean64(P, true, (const double *)L->g[in], (const double *)L->tlag, L->d_f0 + 6, L->d_f0,
                                L->stream);
        P.f0p = L->d_f0;
        P.f0n = L->d_f0 + 3;
    }
    if (P.les_fd) {
        // u of this step at every node first: the closure differentiates it
        if (L->dev32)
            launch_velocity_fp32(P, (const float *)L->f[in], (const float *)L->tlag, L->stream);
        else
            launch_velocity_fp64(P, (const double *)L->f[in], (const double *)L->tlag, L->stream);
    }
    if (L->dev32)
        launch_step_fp32(P, (const float *)L->f[in], (const float *)L->g[in], (float *)L->f[out], (float *)L->g[out],
                         (float *)L->tlag, (float *)L->taules, L->stream);
    else
        launch_step_fp64(P, (const double *)L->f[in], (const double *)L->g[in], (double *)L->f[out],
                         (double *)L->g[out], (double *)L->tlag, (double *)L->taules, L->stream);
    HIPCHK(hipGetLastError());
    L->cur = out;
    L->first = 0;
    L->itime = itime;
    if (L->part) {
        if (fused) {
            L->part->pending = 0;
            int rc = move_particles_on_fields(L, itime - 1);
            if (rc) return rc;
        }
        if (fields_fused(L)) {
            L->part->pending = 1;  // after the next kernel, or at the end of ch4g_step
            return 0;
        }
        return particles_after_step(L);
    }
    return 0;
}

}  // namespace

namespace ch4g {

// Two steady-state steps in one launch (2-D FP32-storage lattices, see
// two_step_ok): the lag moves to the other array.
bool two_step_eligible(const ch4g_lattice *L)
{
    return L->use_fusion && L->dev32 && !L->first && !L->part && two_step_ok(L->base);
}

int ensure_tlag_alt(ch4g_lattice *L)
{
    if (L->base.thermal && !L->tlag_alt) return dalloc(L, &L->tlag_alt, (size_t)L->N * L->rsz);
    return 0;
}

}  // namespace ch4g

namespace {

int launch_two_steps(ch4g_lattice *L)
{
    int rc = ensure_tlag_alt(L); // allocated before any capture (replay_graph)
    if (rc) return rc;
    KParams P = params_for(L, 0, (int)L->itime + 1);
    const int in = L->cur, out = 1 - L->cur;
    launch_two_steps_fp32(P, (const float *)L->f[in], (const float *)L->g[in], (float *)L->f[out],
                          (float *)L->g[out], (const float *)L->tlag, (float *)L->tlag_alt, L->stream);
    HIPCHK(hipGetLastError());
    L->cur = out;
    if (L->base.thermal) {
        std::swap(L->tlag, L->tlag_alt);
        L->tpar ^= 1;
    }
    L->itime += 2;
    return 0;
}

// Capture (once per starting buffer pair) and replay GRAPH_STEPS
// steady-state steps.  Kernel arguments of these steps depend only on the
// buffer parities.
int replay_graph(ch4g_lattice *L)
{
    if (two_step_eligible(L)) {
        int rc = ensure_tlag_alt(L);
        if (rc) return rc;
    }
    const int start = L->cur | (L->tpar << 1);
    if (!L->graph[start]) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
        const int64_t it0 = L->itime;
        const int cur0 = L->cur, tpar0 = L->tpar;
        void *tl0 = L->tlag, *ta0 = L->tlag_alt;
        int rc = 0;
        if (two_step_eligible(L)) {
            for (int k = 0; k < GRAPH_STEPS / 2 && !rc; k++) rc = launch_two_steps(L);
        } else {
            for (int k = 0; k < GRAPH_STEPS && !rc; k++) rc = launch_one_step(L);
        }
        hipError_t e = hipStreamEndCapture(L->stream, &g);
        // capture records work without running it: restore the counters
        L->itime = it0;
        L->cur = cur0;
        L->tpar = tpar0;
        L->tlag = tl0;
        if (ta0) L->tlag_alt = ta0;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(-2, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&L->graph[start], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            L->graph[start] = nullptr;
            return fail(-2, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
    }
    HIPCHK(hipGraphLaunch(L->graph[start], L->stream));
    L->itime += GRAPH_STEPS;
    return 0;
}

// Per-row base state of the FP32-deviation storage.  Temperature: the
// initial linear conduction profile when the run starts from it, else T0.
// Density: with buoyancy, the hydrostatic profile of that temperature
// (d ln rho / dy = 3 F_y, trapezoid rule), normalised to mean rho0; else rho0.
void base_profiles(const ch4g_config &c, const KParams &P, std::vector<double> &rb, std::vector<double> &tb)
{
    const int NY = c.NY;
    rb.assign(NY, c.rho0);
    tb.assign(NY, c.T0);
    if (has(c, CH4G_LB_TEMPERATURE) && has(c, CH4G_TEMPERATURE_INITIAL_LINEAR))
        for (int j = 0; j < NY; j++) tb[j] = (c.T_bot - c.T_ref) - ((c.T_bot - c.T_top) / NY) * (j + 0.5);
    if (P.buoy && c.gravity_y != 0.0) {
        auto Fy = [&](int j) {
            double th = tb[j] - P.T0b;
            return (c.beta_t * th + c.beta2_t * th * th) * c.gravity_y;
        };
        std::vector<double> lr(NY, 0.0);
        for (int j = 1; j < NY; j++) lr[j] = lr[j - 1] + 3.0 * 0.5 * (Fy(j - 1) + Fy(j));
        double mean = 0.0;
        for (int j = 0; j < NY; j++) mean += exp(lr[j]);
        mean /= NY;
        for (int j = 0; j < NY; j++) rb[j] = c.rho0 * exp(lr[j]) / mean;
    }
}

int validate(const ch4g_config *c)
{
    if (!c) return fail(-1, "null config");
    if (c->NX < 1 || c->NY < 1 || c->NZ < 1) return fail(-1, "lattice sizes must be >= 1");
    if ((int64_t)c->NX * c->NY * c->NZ > (int64_t)1 << 31) return fail(-1, "more than 2^31 nodes");
    if (!has(*c, CH4G_LB_FLUID) && !has(*c, CH4G_LB_TEMPERATURE)) return fail(-1, "neither LB_FLUID nor LB_TEMPERATURE");
    if (c->precision != CH4G_FP64 && c->precision != CH4G_FP32_DEV && c->precision != CH4G_FP32_STORE)
        return fail(-1, "unknown precision %d", c->precision);
    if (has(*c, CH4G_TEMPERATURE_BUOYANCY) && !has(*c, CH4G_LB_TEMPERATURE))
        return fail(-1, "LB_TEMPERATURE_BUOYANCY requires LB_TEMPERATURE");
    if (forced(*c) && !has(*c, CH4G_LB_FLUID)) return fail(-1, "fluid forcing requires LB_FLUID");
    if (c->tau_u <= 0.5 && has(*c, CH4G_LB_FLUID)) return fail(-1, "tau_u must exceed 1/2");
    if (c->tau_t <= 0.5 && has(*c, CH4G_LB_TEMPERATURE)) return fail(-1, "tau_t must exceed 1/2");
    if ((has(*c, CH4G_FLUID_BC_Y) || has(*c, CH4G_TEMPERATURE_BC_Y)) && c->NY < 3)
        return fail(-1, "y walls need NY >= 3");
    return 0;
}

}  // namespace

extern "C" {

const char *ch4g_last_error(void) { return g_err; }

int ch4g_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int64_t ch4g_footprint_bytes(const ch4g_config *c)
{
    if (!c) return -1;
    const int64_t N = (int64_t)c->NX * c->NY * c->NZ;
    const int64_t r = c->precision == CH4G_FP64 ? 8 : 4;
    const int64_t S = plane_stride(N, r);
    int64_t b = 0;
    if (has(*c, CH4G_LB_FLUID)) b += 2 * Q * S * r;
    if (has(*c, CH4G_LB_TEMPERATURE)) b += 2 * Q * S * r + N * r;
    if (has(*c, CH4G_FLUID_LES)) b += N * r;
    if (les_fd(*c)) b += 3 * N * r;
    return b;
}

int ch4g_create(const ch4g_config *cfg, ch4g_lattice **out)
{
    g_err[0] = 0;
    if (!out) return fail(-1, "null output handle");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-4, "no HIP device available");
    ch4g_lattice *L = new (std::nothrow) ch4g_lattice();
    if (!L) return fail(-3, "out of host memory");
    L->cfg = *cfg;
    HIPCHK(hipGetDevice(&L->device));
    HIPCHK(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
    L->N = (int64_t)cfg->NX * cfg->NY * cfg->NZ;
    L->dev32 = cfg->precision == CH4G_FP32_DEV || cfg->precision == CH4G_FP32_STORE;
    L->rsz = L->dev32 ? 4 : 8;
    const size_t N = (size_t)L->N, pop = (size_t)Q * (size_t)plane_stride(L->N, L->rsz) * L->rsz;
    const int fl = has(*cfg, CH4G_LB_FLUID), th = has(*cfg, CH4G_LB_TEMPERATURE);
    if (fl)
        for (int b = 0; b < 2; b++)
            if ((rc = dalloc(L, &L->f[b], pop))) goto err;
    if (th) {
        for (int b = 0; b < 2; b++)
            if ((rc = dalloc(L, &L->g[b], pop))) goto err;
        if ((rc = dalloc(L, &L->tlag, N * L->rsz))) goto err;
    }
    if (has(*cfg, CH4G_FLUID_LES))
        if ((rc = dalloc(L, &L->taules, N * L->rsz))) goto err;
    if (les_fd(*cfg))
        if ((rc = dalloc(L, &L->vel, 3 * N * L->rsz))) goto err;

    {
        // design_lb constants (lb.c:368-377) and parameters.c:179,285
        KParams &P = L->base;
        memset(&P, 0, sizeof P);
        P.NX = cfg->NX;
        P.NY = cfg->NY;
        P.NZ = cfg->NZ;
        P.N = L->N;
        P.S = plane_stride(L->N, L->rsz);
        P.fluid = fl;
        P.thermal = th;
        P.wall_f = has(*cfg, CH4G_FLUID_BC_Y);
        P.wall_t = has(*cfg, CH4G_TEMPERATURE_BC_Y);
        P.forced = forced(*cfg);
        P.les = has(*cfg, CH4G_FLUID_LES);
        P.les_fd = les_fd(*cfg);
        P.les_t = has(*cfg, CH4G_FLUID_LES) && has(*cfg, CH4G_TEMPERATURE_LES) && th && fl;
        P.vel = L->vel;
        double cs = 1.0 / sqrt(3.0);
        double cs2 = pow(cs, 2.0), cs4 = pow(cs, 4.0);
        P.cs2 = cs2;
        P.invcs2 = 1.0 / cs2;
        P.invcs4 = 1.0 / cs4;
        P.invtwocs2 = 1.0 / (2.0 * cs2);
        P.invtwocs4 = 1.0 / (2.0 * cs4);
        L->nu = (cfg->tau_u - 0.5 * 1.0) / 3.0;
        L->kappa = (cfg->tau_t - 0.5 * 1.0) / 3.0;
        L->deltaT = cfg->T_bot - cfg->T_top;
        P.nu = L->nu;
        P.tau_u = cfg->tau_u;
        P.invtau_u = 1.0 / cfg->tau_u;
        P.invtau_t = 1.0 / cfg->tau_t;
        P.guo_fac = (1.0 - 0.5 * 1.0 * (1.0 / cfg->tau_u));
        P.Cs2 = pow(cfg->les_cs * 1.0, 2.0);
        P.kappa = L->kappa;
        const double LX = cfg->NX, LY = cfg->NY;
        double Fc[3] = {0.0, 0.0, 0.0};
        if (has(*cfg, CH4G_FLUID_FORCING_POISEUILLE)) {
            Fc[0] += 2.0 * cfg->Amp_x * ((4. * L->nu) * pow(LY, -2.0));
            Fc[1] += 2.0 * cfg->Amp_y * ((4. * L->nu) * pow(LX, -2.0));
            Fc[2] += 2.0 * cfg->Amp_z * ((4. * L->nu) * pow(LY, -2.0));
        }
        if (has(*cfg, CH4G_FLUID_FORCING_CHANNEL)) {
            Fc[0] += 2.0 * pow(cfg->Amp_x, 2.0) / LY;
            Fc[1] += 2.0 * pow(cfg->Amp_y, 2.0) / LX;
            Fc[2] += 2.0 * pow(cfg->Amp_z, 2.0) / LY;
        }
        for (int a = 0; a < 3; a++) P.Fc[a] = Fc[a];
        P.hit = has(*cfg, CH4G_FLUID_FORCING_HIT);
        static const int VK[CH4G_HIT_NK][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
        for (int ii = 0; ii < CH4G_HIT_NK; ii++) {
            for (int a = 0; a < 3; a++) L->vk[ii][a] = VK[ii][a];
            double vk2 = L->vk[ii][0] * L->vk[ii][0] + L->vk[ii][1] * L->vk[ii][1] + L->vk[ii][2] * L->vk[ii][2];
            double fac = pow(vk2, -2. / 6.);
            P.hit_fa[0][ii] = fac * cfg->Amp_x;
            P.hit_fa[1][ii] = fac * cfg->Amp_y;
            P.hit_fa[2][ii] = fac * cfg->Amp_z;
        }
        P.buoy = has(*cfg, CH4G_TEMPERATURE_BUOYANCY);
        P.T0b = has(*cfg, CH4G_TEMPERATURE_BUOYANCY_T0_REF) ? cfg->T_ref : 0.5 * (cfg->T_bot + cfg->T_top);
        P.beta_t = cfg->beta_t;
        P.beta2_t = cfg->beta2_t;
        P.grav[0] = cfg->gravity_x;
        P.grav[1] = cfg->gravity_y;
        P.grav[2] = cfg->gravity_z;
        P.Tw_top = cfg->T_top - cfg->T_ref;
        P.Tw_bot = cfg->T_bot - cfg->T_ref;
        P.rho0 = cfg->rho0;
        P.T0 = cfg->T0;
        P.arith64 = cfg->precision == CH4G_FP32_STORE;
        P.nozero = has(*cfg, CH4G_FLUID_FORCING_NOZEROMODE) ? 1 : 0;
    }
    if (L->dev32) {
        std::vector<double> rb, tb;
        base_profiles(*cfg, L->base, rb, tb);
        const int NY = cfg->NY;
        std::vector<float> r32((size_t)NY * 8, 0.f);
        std::vector<double> r64((size_t)NY * 2);
        for (int j = 0; j < NY; j++) {
            int jm = (j + NY - 1) % NY, jp = (j + 1) % NY;
            float *r = &r32[(size_t)j * 8];
            r[0] = (float)rb[j];
            r[1] = (float)tb[j];
            r[2] = (float)(rb[jm] - rb[j]);
            r[3] = (float)(rb[jp] - rb[j]);
            r[4] = (float)(tb[jm] - tb[j]);
            r[5] = (float)(tb[jp] - tb[j]);
            r[6] = (float)(tb[j] - L->base.T0b);
            r64[2 * j] = rb[j];
            r64[2 * j + 1] = tb[j];
        }
        L->base.cw_top = L->base.Tw_top - tb[NY - 1];
        L->base.cw_bot = L->base.Tw_bot - tb[0];
        if ((rc = dalloc(L, (void **)&L->d_row32, r32.size() * sizeof(float)))) goto err;
        if ((rc = dalloc(L, (void **)&L->d_row64, r64.size() * sizeof(double)))) goto err;
        if (hipMemcpy(L->d_row32, r32.data(), r32.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(L->d_row64, r64.data(), r64.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(-2, "upload of the base-state rows failed");
            goto err;
        }
        L->base.row32 = L->d_row32;
        L->base.row64 = L->d_row64;
    }
    fill_float_params(L->base);
    // srand48(seed): X = (seed << 16) | 0x330E
    L->xsubi[0] = 0x330E;
    L->xsubi[1] = (unsigned short)((uint32_t)cfg->seed & 0xFFFFu);
    L->xsubi[2] = (unsigned short)(((uint32_t)cfg->seed >> 16) & 0xFFFFu);
    memset(L->phi, 0, sizeof L->phi);
    if (has(*cfg, CH4G_FLUID_FORCING_NOZEROMODE))
        if ((rc = dalloc(L, (void **)&L->d_f0, (size_t)(6 + 512 * 6) * sizeof(double)))) goto err;
    if (L->base.hit) {
        L->hit_len = (size_t)CH4G_HIT_NK * (cfg->NX + cfg->NY + cfg->NZ);
        for (int b = 0; b < 2; b++)
            if ((rc = dalloc(L, (void **)&L->d_hit[b], L->hit_len * sizeof(double)))) goto err;
        for (int r = 0; r < HIT_RING; r++) {
            hipError_t e = hipHostMalloc((void **)&L->h_ring[r], L->hit_len * sizeof(double), hipHostMallocDefault);
            if (e != hipSuccess) {
                rc = fail(-3, "hipHostMalloc failed: %s", hipGetErrorString(e));
                goto err;
            }
            e = hipEventCreateWithFlags(&L->ev_ring[r], hipEventDisableTiming);
            if (e != hipSuccess) {
                rc = fail(-2, "hipEventCreate failed: %s", hipGetErrorString(e));
                goto err;
            }
            (void)hipEventRecord(L->ev_ring[r], L->stream);
        }
    }
    *out = L;
    return 0;
err:
    ch4g_destroy(L);
    return rc;
}

void ch4g_destroy(ch4g_lattice *L)
{
    if (!L) return;
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    for (int b = 0; b < 2; b++) {
        if (L->f[b]) (void)hipFree(L->f[b]);
        if (L->g[b]) (void)hipFree(L->g[b]);
        if (L->d_hit[b]) (void)hipFree(L->d_hit[b]);
    }
    if (L->d_f0) (void)hipFree(L->d_f0);
    for (int r = 0; r < HIT_RING; r++) {
        if (L->h_ring[r]) (void)hipHostFree(L->h_ring[r]);
        if (L->ev_ring[r]) (void)hipEventDestroy(L->ev_ring[r]);
    }
    destroy_graphs(L);
    particles_free(L);
    void *scr[] = {L->tlag, L->tlag_alt, L->taules, L->vel, L->d_rul, L->d_uT32, L->d_dens, L->d_u, L->d_t, L->d_part, L->d_res, L->d_fs, L->d_gs,
                   L->d_row32, L->d_row64};
    for (void *p : scr)
        if (p) (void)hipFree(p);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    delete L;
}

int ch4g_initial_conditions(ch4g_lattice *L)
{
    if (!L) return fail(-1, "null lattice");
    particles_free(L);
    const ch4g_config &c = L->cfg;
    const int NX = c.NX, NY = c.NY, NZ = c.NZ;
    const double LY = NY;
    // fluid rows (initial_conditions.c:45, :72-84, :256-279)
    std::vector<double> prow((size_t)NY * Q);
    for (int j = 0; j < NY; j++)
        for (int q = 0; q < Q; q++) {
            double v = wq(q);
            if (has(c, CH4G_TEMPERATURE_BUOYANCY) && has(c, CH4G_INITIAL_BAROMETRIC)) {
                double L_ = LY, y = (double)j + 0.5;
                v = wq(q) * (exp(c.beta_t * c.gravity_y * y * (0.5 * L->deltaT - 0.5 * (L->deltaT / L_) * y) /
                                 L->base.cs2));
            }
            prow[(size_t)j * Q + q] = v;
        }
    if (has(c, CH4G_LB_FLUID) && has(c, CH4G_FLUID_INITIAL_UNIT_DENSITY)) {
        double norm = 0.0;
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++)
                for (int i = 0; i < NX; i++)
                    for (int q = 0; q < Q; q++) norm += prow[(size_t)j * Q + q];
        norm /= (double)NX * (double)NY * (double)NZ;
        for (auto &v : prow) v /= norm;
    }
    // temperature field (initial_conditions.c:294-424), drand48 in k,j,i order
    std::vector<double> t0;
    if (has(c, CH4G_LB_TEMPERATURE)) {
        t0.assign((size_t)L->N, 0.0);
        size_t m = 0;
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++)
                for (int i = 0; i < NX; i++, m++) {
                    double t = 0.0;
                    if (has(c, CH4G_TEMPERATURE_INITIAL_LINEAR)) {
                        double y = (double)j + 0.5;
                        t = ((c.T_bot - c.T_ref) - (L->deltaT / LY) * y);
                    }
                    if (has(c, CH4G_TEMPERATURE_INITIAL_ADD_NOISE)) t += c.noise_t * (2.0 * erand48(L->xsubi) - 1.0);
                    t0[m] = t;
                }
    }
    // HIT phases (initialization_forcing, initial_conditions.c:866-871)
    if (L->base.hit)
        for (int ii = 0; ii < CH4G_HIT_NK; ii++)
            for (int a = 0; a < 3; a++) L->phi[ii][a] = erand48(L->xsubi);

    double *d_prow = nullptr, *d_t0 = nullptr;
    HIPCHK(hipMalloc((void **)&d_prow, prow.size() * sizeof(double)));
    HIPCHK(hipMemcpyAsync(d_prow, prow.data(), prow.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
    if (!t0.empty()) {
        HIPCHK(hipMalloc((void **)&d_t0, t0.size() * sizeof(double)));
        HIPCHK(hipMemcpyAsync(d_t0, t0.data(), t0.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
    }
    L->cur = 0;
    L->first = 1;
    L->itime = 0;
    launch_init(L->base, d_prow, d_t0, L->f[0], L->g[0], L->tlag, L->dev32, L->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(L->stream));
    (void)hipFree(d_prow);
    if (d_t0) (void)hipFree(d_t0);
    return first_build_forcing(L);
}

int ch4g_set_populations(ch4g_lattice *L, const double *p, const double *g)
{
    if (!L) return fail(-1, "null lattice");
    particles_free(L);
    int rc = ensure_stage(L);
    if (rc) return rc;
    const size_t N = (size_t)L->N;
    std::vector<double> soa((size_t)Q * N);
    const double *src[2] = {p, g};
    double *dst[2] = {L->d_fs, L->d_gs};
    for (int s = 0; s < 2; s++) {
        if (!src[s]) continue;
        for (size_t n = 0; n < N; n++)
            for (int q = 0; q < Q; q++) soa[(size_t)q * N + n] = src[s][n * Q + q];
        HIPCHK(hipMemcpy(dst[s], soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    L->cur = 0;
    L->first = 1;
    L->itime = 0;
    launch_load(L->base, p ? L->d_fs : nullptr, g ? L->d_gs : nullptr, L->f[0], L->g[0], L->tlag, L->dev32,
                L->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(L->stream));
    return first_build_forcing(L);
}

int ch4g_get_populations(ch4g_lattice *L, double *p, double *g)
{
    if (!L) return fail(-1, "null lattice");
    int rc = ensure_stage(L);
    if (rc) return rc;
    const size_t N = (size_t)L->N;
    gather(L, p ? L->d_fs : nullptr, g ? L->d_gs : nullptr, nullptr, nullptr, nullptr);
    HIPCHK(hipGetLastError());
    std::vector<double> soa((size_t)Q * N);
    const double *src[2] = {L->d_fs, L->d_gs};
    double *dst[2] = {p, g};
    for (int s = 0; s < 2; s++) {
        if (!dst[s]) continue;
        if (s == 0 && !L->base.fluid) continue;
        if (s == 1 && !L->base.thermal) continue;
        HIPCHK(hipMemcpyAsync(soa.data(), src[s], soa.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
        for (size_t n = 0; n < N; n++)
            for (int q = 0; q < Q; q++) dst[s][n * Q + q] = soa[(size_t)q * N + n];
    }
    return 0;
}

int ch4g_step(ch4g_lattice *L, int nsteps)
{
    if (!L) return fail(-1, "null lattice");
    if (nsteps < 0) return fail(-1, "negative step count");
    int s = 0;
    while (s < nsteps) {
        // Steady-state steps (not the first, no per-step host forcing tables,
        // no particles) depend only on the buffer parities: replay a captured
        // GRAPH_STEPS-step graph instead of launching.
        if (L->use_graphs && !L->first && L->itime >= 1 && !L->base.hit && !L->part && nsteps - s >= GRAPH_STEPS) {
            int rc = replay_graph(L);
            if (rc) return rc;
            s += GRAPH_STEPS;
            continue;
        }
        if (nsteps - s >= 2 && two_step_eligible(L)) {
            int rc = launch_two_steps(L);
            if (rc) return rc;
            s += 2;
            continue;
        }
        int rc = launch_one_step(L);
        if (rc) return rc;
        s++;
    }
    return particles_flush(L);
}

int ch4g_set_graphs(ch4g_lattice *L, int enable)
{
    if (!L) return fail(-1, "null lattice");
    L->use_graphs = enable != 0;
    return 0;
}

int ch4g_steps_per_launch(const ch4g_lattice *L, int *steps)
{
    if (!L || !steps) return fail(-1, "null argument");
    ch4g::KParams P = L->base;
    P.first = 0;  // steady state
    *steps = (L->use_fusion && L->dev32 && !L->part && two_step_ok(P)) ? 2 : 1;
    return 0;
}

int ch4g_set_fusion(ch4g_lattice *L, int enable)
{
    if (!L) return fail(-1, "null lattice");
    if (L->use_fusion != (enable != 0)) {
        // captured step graphs hold the launches of the old setting
        HIPCHK(hipStreamSynchronize(L->stream));
        destroy_graphs(L);
    }
    L->use_fusion = enable != 0;
    return 0;
}

int ch4g_step_timed(ch4g_lattice *L, int nsteps, double *elapsed_ms)
{
    if (!L || !elapsed_ms) return fail(-1, "null argument");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, L->stream));
    int rc = ch4g_step(L, nsteps);
    if (rc) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }
    HIPCHK(hipEventRecord(e1, L->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *elapsed_ms = ms;
    return 0;
}

int ch4g_hydro_fields(ch4g_lattice *L, double *dens, double *u, double *t)
{
    if (!L) return fail(-1, "null lattice");
    int rc = ensure_scratch(L);
    if (rc) return rc;
    gather(L, nullptr, nullptr, L->d_dens, L->d_u, L->d_t);
    HIPCHK(hipGetLastError());
    const size_t N = (size_t)L->N;
    if (dens && L->base.fluid)
        HIPCHK(hipMemcpyAsync(dens, L->d_dens, N * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    if (u && L->base.fluid) HIPCHK(hipMemcpyAsync(u, L->d_u, 3 * N * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    if (t && L->base.thermal)
        HIPCHK(hipMemcpyAsync(t, L->d_t, N * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return 0;
}

namespace {

// normalisation of output.c:499-690: fluid members by fnorm, temperature
// members by tnorm, Nusselt numbers by tnorm / (kappa deltaT / SY), lb from
// the normalised eps and epst with buoyancy
void normalise(const ch4g_lattice *L, ch4g_output &o, double fnorm, double tnorm)
{
    const ch4g_config &c = L->cfg;
    if (L->base.fluid) {
        double *v[] = {&o.ene, &o.rho, &o.ux, &o.uy, &o.uz, &o.ux2, &o.uy2, &o.uz2, &o.eps, &o.rho2};
        for (double *x : v) *x *= fnorm;
    }
    if (L->base.thermal) {
        double *v[] = {&o.t, &o.t2, &o.epst, &o.dxt, &o.dyt, &o.dzt, &o.uxt, &o.uyt, &o.uzt};
        for (double *x : v) *x *= tnorm;
        double nn = tnorm / (L->kappa * L->deltaT / (double)c.NY);
        o.nux *= nn;
        o.nuy *= nn;
        o.nuz *= nn;
        o.lb = 0.0;
        if (L->base.buoy)
            o.lb = pow(o.eps, 5. / 4.) * pow(o.epst, -3. / 4.) * pow(c.gravity_y * c.beta_t, -3. / 2.);
    }
}

}  // namespace

int ch4g_dump_averages(ch4g_lattice *L, ch4g_output *out)
{
    if (!L || !out) return fail(-1, "null argument");
    int rc = ensure_scratch(L);
    if (rc) return rc;
    gather(L, nullptr, nullptr, L->d_dens, L->d_u, L->d_t);
    launch_averages(L->base, L->d_dens, L->d_u, L->d_t, L->nu, L->kappa, L->d_part, L->d_res, L->stream);
    HIPCHK(hipGetLastError());
    double a[NOUT];
    HIPCHK(hipMemcpyAsync(a, L->d_res, sizeof a, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const ch4g_config &c = L->cfg;
    const double norm = 1.0 / ((double)c.NX * (double)c.NY * (double)c.NZ);
    ch4g_output o;
    memcpy(&o, a, sizeof o);
    normalise(L, o, norm, norm);
    *out = o;
    return 0;
}

}  // extern "C"

namespace {

// Block partials of the averages reduction (reduce_blocks caps its grid at
// AVG_MAX_BLOCKS) and, after them, the NOUT sums, for the allocation-free
// entry point.  One copy per device, held until the process ends; host code
// uses the address hipGetSymbolAddress returns for the device, never the
// array's name.
constexpr int AVG_MAX_BLOCKS = 2048;
constexpr size_t AVG_SCRATCH_DOUBLES = (size_t)(AVG_MAX_BLOCKS + 1) * NOUT;
__device__ __align__(256) double scratch[AVG_SCRATCH_DOUBLES];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_dev[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on device dev, looked up once per device with
// that device current (the caller's current device is restored).
int scratch_on(int dev, double **out)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(-1, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        int prev = 0;
        HIPCHK(hipGetDevice(&prev));
        if (prev != dev) HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != dev) HIPCHK(hipSetDevice(prev));
        if (e != hipSuccess || !p) return fail(-2, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        scratch_dev[dev] = (double *)p;
    }
    *out = scratch_dev[dev];
    return 0;
}

}  // namespace

extern "C" {

int ch4g_dump_averages_noalloc(ch4g_lattice *L, double *dens, double *u, double *t, ch4g_output *out)
{
    if (!L || !out) return fail(-1, "null argument");
    if (L->base.fluid && (!dens || !u)) return fail(-1, "a fluid lattice needs device buffers dens[N] and u[3N]");
    if (L->base.thermal && !t) return fail(-1, "a thermal lattice needs a device buffer t[N]");
    const int nb = reduce_blocks(L->base);
    if (nb < 1 || nb > AVG_MAX_BLOCKS)
        return fail(-3, "averages reduction of %d blocks exceeds the %d-block scratch", nb, AVG_MAX_BLOCKS);
    double *s = nullptr;
    int rc = scratch_on(L->device, &s);
    if (rc) return rc;
    double *res = s + (size_t)AVG_MAX_BLOCKS * NOUT;
    // every partial the final sum reads is written first: nothing to clear
    gather(L, nullptr, nullptr, dens, u, t);
    launch_averages(L->base, dens, u, t, L->nu, L->kappa, s, res, L->stream);
    HIPCHK(hipGetLastError());
    double a[NOUT];
    HIPCHK(hipMemcpyAsync(a, res, sizeof a, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const ch4g_config &c = L->cfg;
    const double norm = 1.0 / ((double)c.NX * (double)c.NY * (double)c.NZ);
    ch4g_output o;
    memcpy(&o, a, sizeof o);
    normalise(L, o, norm, norm);
    *out = o;
    return 0;
}

int ch4g_dump_rulers(ch4g_lattice *L, ch4g_output *rx, ch4g_output *ry, ch4g_output *rz)
{
    if (!L) return fail(-1, "null lattice");
    int rc = ensure_scratch(L);
    if (rc) return rc;
    const ch4g_config &c = L->cfg;
    const size_t planes = (size_t)c.NX + c.NY + c.NZ;
    if (!L->d_rul && (rc = dalloc(L, (void **)&L->d_rul, planes * NOUT * sizeof(double)))) return rc;
    double *dx = L->d_rul, *dy = dx + (size_t)c.NX * NOUT, *dz = dy + (size_t)c.NY * NOUT;
    gather(L, nullptr, nullptr, L->d_dens, L->d_u, L->d_t);
    launch_rulers(L->base, L->d_dens, L->d_u, L->d_t, L->nu, L->kappa, rx ? dx : nullptr, ry ? dy : nullptr,
                  rz ? dz : nullptr, L->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> h(planes * NOUT);
    HIPCHK(hipMemcpyAsync(h.data(), L->d_rul, h.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    static_assert(sizeof(ch4g_output) == NOUT * sizeof(double), "ch4g_output layout");
    const double SX = c.NX, SY = c.NY, SZ = c.NZ;
    // output.c:520-650: fluid 1/(SY SZ), 1/(SX SZ), 1/(SX SY); temperature
    // 1/(SY SZ), 1/(SX SZ) and again 1/(SY SZ) for z; the ruler's own
    // coordinate by 1/(NY NZ), 1/(NX NZ), 1/(NX NY)
    struct Axis {
        ch4g_output *r;
        int n;
        const double *src;
        double fnorm, tnorm, cnorm;
        int coord;
    } ax[3] = {{rx, c.NX, h.data(), 1.0 / (SY * SZ), 1.0 / (SY * SZ), 1.0 / (double)(c.NY * c.NZ), 0},
               {ry, c.NY, h.data() + (size_t)c.NX * NOUT, 1.0 / (SX * SZ), 1.0 / (SX * SZ),
                1.0 / (double)(c.NX * c.NZ), 1},
               {rz, c.NZ, h.data() + (size_t)(c.NX + c.NY) * NOUT, 1.0 / (SX * SY), 1.0 / (SY * SZ),
                1.0 / (double)(c.NX * c.NY), 2}};
    for (const Axis &A : ax) {
        if (!A.r) continue;
        for (int i = 0; i < A.n; i++) {
            ch4g_output o;
            memcpy(&o, A.src + (size_t)i * NOUT, sizeof o);
            double *xyz = &o.x;
            const double cc = L->base.fluid ? xyz[A.coord] * A.cnorm : 0.0;
            normalise(L, o, A.fnorm, A.tnorm);
            o.x = o.y = o.z = 0.0;
            xyz[A.coord] = cc;
            A.r[i] = o;
        }
    }
    return 0;
}

int ch4g_get_hit_phases(ch4g_lattice *L, double *phi)
{
    if (!L || !phi) return fail(-1, "null argument");
    memcpy(phi, L->phi, sizeof L->phi);
    return 0;
}

int ch4g_set_hit_phases(ch4g_lattice *L, const double *phi)
{
    if (!L || !phi) return fail(-1, "null argument");
    if (!L->base.hit) return fail(-1, "lattice has no HIT forcing");
    memcpy(L->phi, phi, sizeof L->phi);
    return hit_upload(L, (int)((L->itime + 1) & 1));
}

int ch4g_synchronize(ch4g_lattice *L)
{
    if (!L) return fail(-1, "null lattice");
    HIPCHK(hipStreamSynchronize(L->stream));
    return 0;
}

int64_t ch4g_itime(const ch4g_lattice *L) { return L ? L->itime : -1; }

int64_t ch4g_device_bytes(const ch4g_lattice *L) { return L ? L->bytes : -1; }

}  // extern "C"
