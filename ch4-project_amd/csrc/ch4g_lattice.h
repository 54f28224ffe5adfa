// Host-side record of one device lattice, shared by the C-ABI translation
// units (ch4g_api.cpp: lattice and step; ch4g_particles.cpp: particles).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ch4g_internal.h"

namespace ch4g {

constexpr int HIT_RING = 8;
// steps between reorderings of the particle rows by cell
constexpr int REORDER_STEPS = 32;
// steps per captured graph (even: the graph starts and ends on the same buffer)
constexpr int GRAPH_STEPS = 32;

// Lagrangian particles of one lattice (device SoA, see PartField).
struct Particles {
    ch4g_particle_config pc{};
    int64_t np = 0;          // particles in the arrays (alive or dropped)
    int64_t cap = 0;         // capacity of each SoA row
    double *d_state = nullptr;   // [PF_COUNT][cap]
    int *d_alive = nullptr;      // 1 alive, 0 dropped by sendrecv_particles
    int *d_ndraw = nullptr;      // drand48 draws of the current sub-step
    int *d_bsum = nullptr;       // per-block draw sums, then block offsets
    uint64_t *d_rng = nullptr;   // 48-bit drand48 state while particles own the stream
    int rng_on_device = 0;
    // the particle update of the last step is still to run (fields_fused:
    // it runs after the next step kernel has written that step's fields)
    int pending = 0;
    double dt = 1.0;             // time_dt_lagr
    int substeps = 1;
    // two-way coupling: feedback force planes [3][N] per step parity (double
    // on FP64 lattices, float on float storage), and the sorted stencil cells
    // of the particles of each (whose heads the next call clears)
    void *d_fb[2] = {nullptr, nullptr};
    uint32_t *d_keys[2] = {nullptr, nullptr};
    int64_t nkeys[2] = {0, 0};
    uint32_t *d_kin = nullptr, *d_vin = nullptr, *d_vout = nullptr;
    double *d_rec = nullptr;     // [np][8] fac f_p, position and reference index of each row
    double *d_dep = nullptr;     // [8][np][3] corner deposits of the particles sorted by stencil cell
    uint32_t *d_ord = nullptr;   // [np] their reference indices
    int *d_head = nullptr;       // [(NX+1)(NY+1)(NZ+1)] first sorted position of each stencil cell, or -1
    // rows kept in cell order (reordered every REORDER_STEPS steps): the
    // reference-order index of each row, and the buffers of the reordering
    int *d_orig = nullptr, *d_orig2 = nullptr, *d_alive2 = nullptr;
    double *d_state2 = nullptr;
    uint32_t *d_ck = nullptr, *d_ck2 = nullptr, *d_cv = nullptr, *d_cv2 = nullptr;
    void *d_ro_tmp = nullptr;
    size_t ro_tmp_bytes = 0;
    int cell_bits = 32;
    int64_t last_reorder = 0;
    PermRows perm{};             // rows the reordering moves (set by each upload)
    std::vector<ch4g_particle_type> types;  // per-type values of the last initialisation
    void *d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int key_bits = 32;
};

}  // namespace ch4g

struct ch4g_lattice {
    ch4g_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t N = 0;
    int dev32 = 0;           // FP32-deviation storage
    size_t rsz = 8;          // bytes per stored real
    void *f[2] = {nullptr, nullptr};
    void *g[2] = {nullptr, nullptr};
    void *tlag = nullptr;
    void *taules = nullptr;
    void *vel = nullptr;     // [3][N] velocity of the step (finite-difference LES)
    int cur = 0;             // buffer holding the current state
    int first = 1;           // state holds pre-collision populations (no gather)
    int64_t itime = 0;
    // constants of the step
    ch4g::KParams base;
    double nu = 0, kappa = 0, deltaT = 0;
    // HIT forcing: phases of the current force (Phi_{itime+1}), device tables
    // for the two most recent steps, pinned staging ring
    double phi[CH4G_HIT_NK][3];
    double vk[CH4G_HIT_NK][3];
    double *d_hit[2] = {nullptr, nullptr};
    // LB_FLUID_FORCING_NOZEROMODE: means [6] + block partial sums [512][6]
    double *d_f0 = nullptr;
    double *h_ring[ch4g::HIT_RING] = {};
    hipEvent_t ev_ring[ch4g::HIT_RING] = {};
    int ring = 0;
    size_t hit_len = 0;      // doubles per table
    unsigned short xsubi[3]; // drand48 stream (srand48(seed) state)
    // lazily allocated FP64 scratch
    double *d_dens = nullptr, *d_u = nullptr, *d_t = nullptr;
    double *d_part = nullptr, *d_res = nullptr;
    double *d_rul = nullptr;  // ruler plane sums [NX + NY + NZ][NOUT]
    float4 *d_uT32 = nullptr;  // [N] (u_x, u_y, u_z, T) in float for particles (float storage)
    double *d_fs = nullptr, *d_gs = nullptr;
    // FP32-deviation per-row base state
    float *d_row32 = nullptr;
    double *d_row64 = nullptr;
    // captured GRAPH_STEPS-step sequences, one per starting buffer and
    // temperature-lag buffer (index cur | tpar << 1)
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    // two-step launches (2-D FP32 storage) write the lag into a second array;
    // tlag stays the current one, tpar counts the swaps mod 2
    void *tlag_alt = nullptr;
    int tpar = 0;
    int use_graphs = 1;
    int use_fusion = 1;      // ch4g_set_fusion
    int64_t bytes = 0;
    ch4g::Particles *part = nullptr;
};

namespace ch4g {

// error reporting shared by the C-ABI units (message via ch4g_last_error)
int fail(int code, const char *fmt, ...);
int dalloc(ch4g_lattice *L, void **p, size_t bytes);
int ensure_scratch(ch4g_lattice *L);
KParams params_for(const ch4g_lattice *L, int first, int itime);
void gather(ch4g_lattice *L, double *fpost, double *gpost, double *dens, double *u, double *t);
void destroy_graphs(ch4g_lattice *L);

// particle hooks of the step sequence (ch4g_particles.cpp)
int particles_before_step(ch4g_lattice *L, int itime);  // add_particle_feedbacks
bool two_step_eligible(const ch4g_lattice *L);           // steady steps go two per launch
int ensure_tlag_alt(ch4g_lattice *L);                    // the second lag array of two-step launches
int particles_after_step(ch4g_lattice *L);               // main.c:163-183
bool fields_fused(const ch4g_lattice *L);                // the step kernel writes the particles' fields
int ensure_float_fields(ch4g_lattice *L);
int move_particles_on_fields(ch4g_lattice *L, int64_t step);
int particles_flush(ch4g_lattice *L);                    // the update owed by a fused sequence
void particles_free(ch4g_lattice *L);
int rng_to_host(ch4g_lattice *L);                        // before host drand48 use
int rng_to_device(ch4g_lattice *L);                      // after it

}  // namespace ch4g

#define HIPCHK(call)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return ::ch4g::fail(-2, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,    \
                                __LINE__);                                                               \
    } while (0)
