// Internal declarations of the HIP lattice-Boltzmann library (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "../../include/ch4g.h"

namespace ch4g {

constexpr int Q = 19;

// D3Q19 velocities in the reference's population order (src/lb.c:163-181).
__host__ __device__ constexpr int cxq(int q)
{
    constexpr int v[Q] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    return v[q];
}
__host__ __device__ constexpr int cyq(int q)
{
    constexpr int v[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
    return v[q];
}
__host__ __device__ constexpr int czq(int q)
{
    constexpr int v[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1, -1};
    return v[q];
}
// opposite direction (src/lb.c:188-206)
__host__ __device__ constexpr int invq(int q)
{
    constexpr int v[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
    return v[q];
}
// weights 1/3, 1/18, 1/36 (src/lb.c:141-159), as the reference writes them
__host__ __device__ constexpr double wq(int q) { return q == 0 ? 1. / 3. : (q <= 6 ? 1. / 18. : 1. / 36.); }

// HIT forcing: per-step tables of sin(2 pi (k x / L + phi)) for the six modes,
// layout [6][NX] | [6][NY] | [6][NZ].
struct HitTable {
    const double *sx, *sy, *sz;
};

// Everything a step / gather kernel needs, passed by value.
struct KParams {
    int NX, NY, NZ;
    int64_t N;          // interior nodes
    int64_t S;          // population plane stride in elements (plane_stride: N, padded off 64 KiB multiples)
    int fluid, thermal; // population sets present
    int wall_f, wall_t; // y walls: fluid no-slip, temperature fixed
    int forced;         // any fluid forcing (hydro_fields half-force shift + Guo)
    int first;          // inputs are pre-collision p_0/g_0: no gather
    int les;            // Smagorinsky relaxation time per node (tau_u_les)
    int les_fd;         // |S| from finite differences of the stored velocity (else local strain)
    int les_t;          // temperature relaxation time from tau_u_les (tau_t_les)
    int itime;          // step being computed (1 for the first step)
    // sound-speed constants computed on the host exactly as the reference does
    double invcs2, invcs4, invtwocs2, invtwocs4, cs2;
    double invtau_u, invtau_t, guo_fac, nu, tau_u, Cs2, kappa;
    // velocity planes [3][N] of the step (double, or float for float storage),
    // written by the velocity kernel for the finite-difference LES closure
    const void *vel;
    // forcing
    double Fc[3];       // constant part (Poiseuille + channel), summed in reference order
    int hit;
    double hit_fa[3][CH4G_HIT_NK];  // pow(k^2,-1/3) * Amp_{x,y,z}
    HitTable hit_prev, hit_now;
    // LB_FLUID_FORCING_NOZEROMODE: the mean of the built force of the
    // previous (f0p) and of this step (f0n), removed at every node; null
    // without the switch (nozero set: the lattice has it)
    const double *f0p, *f0n;
    int nozero;
    int buoy;
    double T0b, beta_t, beta2_t, grav[3];
    double Tw_top, Tw_bot;  // (T_wall - T_ref)
    // FP32-deviation base state: per-row profiles rho_b(y), T_b(y).
    // row32[y*8 + k]: 0 rho_b, 1 T_b, 2 rho_b(y-1)-rho_b(y), 3 rho_b(y+1)-rho_b(y),
    //                 4 T_b(y-1)-T_b(y), 5 T_b(y+1)-T_b(y), 6 T_b(y)-T0b, 7 unused
    // (rows y-1, y+1 periodic); row64[y*2 + k]: 0 rho_b, 1 T_b in double.
    const float *row32;
    const double *row64;
    double cw_top, cw_bot;  // (T_wall - T_ref) - T_b(wall row)
    double rho0, T0;
    int arith64;            // float storage with double collision arithmetic
    // two-way particle feedback added to the built force (lagrange_feedback.c:245-254):
    // planes [3][N] of the previous step (u shift) and of this step (Guo term)
    int twoway;
    const double *fb_prev, *fb_now;       // FP64 lattices
    const float *fbf_prev, *fbf_now;      // float storage: the same sums rounded to float
    // float copies of the step constants used by the float collision (the
    // same round-to-nearest values the kernels would convert per use);
    // filled by fill_float_params once the double values are final
    float Fc_f[3], grav_f[3], beta_t_f, beta2_t_f, invtau_u_f, invtau_t_f, guo_fac_f, cw_top_f, cw_bot_f;
    // when set (float storage, particles), the step kernel also writes the
    // hydro fields of the previous step it has just gathered: [N][4] float
    // (u_x, u_y, u_z, T; 0 for a population set the lattice lacks), the
    // fields the particle kernel interpolates
    float4 *uT_out;
    int part;  // particles ride on the step kernel's fields (the instance that can write them)
};

// Stride of the population planes of a lattice of N nodes with rsz-byte
// reals.  Planes whose byte size is a multiple of 64 KiB would start at the
// same offset modulo large powers of two, and the 38 streams of a step then
// collide on the same memory channels (a copy of 38 planes at 512^3: 4.72
// TB/s with stride N, 5.34 TB/s padded).  The pad (4160 elements) keeps the
// planes 256-byte aligned.
inline int64_t plane_stride(int64_t N, int64_t rsz)
{
    return (N * rsz) % 65536 == 0 ? N + 4160 : N;
}

inline void fill_float_params(KParams &P)
{
    for (int a = 0; a < 3; a++) {
        P.Fc_f[a] = (float)P.Fc[a];
        P.grav_f[a] = (float)P.grav[a];
    }
    P.beta_t_f = (float)P.beta_t;
    P.beta2_t_f = (float)P.beta2_t;
    P.invtau_u_f = (float)P.invtau_u;
    P.invtau_t_f = (float)P.invtau_t;
    P.guo_fac_f = (float)P.guo_fac;
    P.cw_top_f = (float)P.cw_top;
    P.cw_bot_f = (float)P.cw_bot;
}

// Neighbour coordinates of a node for the pull gather: the source of
// population q is x - c_q, periodic in every direction; crossings of y walls
// are flagged (uniform across a wavefront, whose lanes share y and z).
struct Nbr {
    int x, y, z;
    int xm, xp, ym, yp, zm, zp;
    bool bot, top;  // y == 0, y == NY-1
};

__device__ __forceinline__ Nbr make_nbr(int x, int y, int z, int NX, int NY, int NZ)
{
    Nbr b;
    b.x = x;
    b.y = y;
    b.z = z;
    b.xm = x == 0 ? NX - 1 : x - 1;
    b.xp = x == NX - 1 ? 0 : x + 1;
    b.ym = y == 0 ? NY - 1 : y - 1;
    b.yp = y == NY - 1 ? 0 : y + 1;
    b.zm = z == 0 ? NZ - 1 : z - 1;
    b.zp = z == NZ - 1 ? 0 : z + 1;
    b.bot = y == 0;
    b.top = y == NY - 1;
    return b;
}

// linear index of the pull source x - c_q (periodic wrap)
template <int q>
__device__ __forceinline__ uint32_t src_index(const Nbr &b, int NX, int NY)
{
    const int sx = cxq(q) > 0 ? b.xm : (cxq(q) < 0 ? b.xp : b.x);
    const int sy = cyq(q) > 0 ? b.ym : (cyq(q) < 0 ? b.yp : b.y);
    const int sz = czq(q) > 0 ? b.zm : (czq(q) < 0 ? b.zp : b.z);
    return ((uint32_t)sz * (uint32_t)NY + (uint32_t)sy) * (uint32_t)NX + (uint32_t)sx;
}

// does the pull source of q lie beyond a y wall?
template <int q>
__device__ __forceinline__ bool crosses_wall(const Nbr &b)
{
    return (cyq(q) > 0 && b.bot) || (cyq(q) < 0 && b.top);
}

// Block order for XCD locality: the GPU hands blocks to its 8 XCDs in turn
// (block b to XCD b mod 8); renumbering them so that XCD x receives a
// contiguous range makes neighbouring work units share that XCD's L2.  A
// bijection of [0, nb) for any nb.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    const uint32_t x = b % 8u, q = nb / 8u, r = nb % 8u;
    return x * q + (x < r ? x : r) + b / 8u;
}

// compile-time loop over q
template <int q, int Q_, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (q < Q_) {
        f(std::integral_constant<int, q>{});
        static_for<q + 1, Q_>(f);
    }
}

// ---------------------------------------------------------------------------
// Lagrangian particles: FP64 structure-of-arrays rows of one device buffer.
enum PartField {
    PX, PY, PZ, PVX, PVY, PVZ, PVXO, PVYO, PVZO, PAX, PAY, PAZ, PAXO, PAYO, PAZO,
    PUX, PUY, PUZ, PUXO, PUYO, PUZO,
    PG,                 // 9 rows: dx_ux, dy_ux, dz_ux, dx_uy, dy_uy, dz_uy, dx_uz, dy_uz, dz_uz
    PDTU = PG + 9,      // 3 rows: Dt_ux, Dt_uy, Dt_uz
    PT = PDTU + 3, PTOLD, PDTT,
    PGT,                // 3 rows: dx_t, dy_t, dz_t
    PTAU = PGT + 3, PBETA,
    PVOL,               // (4/3) pi (3 nu beta tau)^(3/2), host pow (lagrange_feedback.c:87-90)
    PDRAT,              // (3 - beta) / (2 beta) (lagrange_feedback.c:84)
    PNAME,
    PPO,                // 3 rows: px, py, pz (orientation)
    PDTP = PPO + 3,     // 3 rows: dt_px, dt_py, dt_pz
    PASP = PDTP + 3,    // aspect_ratio
    PGYV,               // gyrotaxis_velocity
    PSWIM,              // swim_velocity
    PNO,                // 3 rows: nx, ny, nz (second orientation)
    PDTN = PNO + 3,     // 3 rows: dt_nx, dt_ny, dt_nz
    PTP = PDTN + 3,     // t_p, the particle's temperature
    PCP,                // cp_p
    PTFAC,              // time_dt_lagr / tau_heat of lagrange.c:4011-4013, host arithmetic
    PTEXP,              // exp(-PTFAC), glibc (the same bits as the reference's call)
    PGC,                // gravity_coeff (LAGRANGE_GRAVITY_VARIABLE)
    PCPE,               // c_perp of the spheroid drag, host libm (lagrange.c:3277-3288)
    PCPA,               // c_par
    PSED,               // sediment (bottom-wall depositions)
    PCT,                // 6 rows: cxx, cyy, czz, cxy, cyz, cxz (polymer conformation)
    PDCT = PCT + 6,     // 6 rows: dt_cxx .. dt_cxz
    PCSR = PDCT + 6,    // critical_shear_rate (jumps)
    PSR,                // shear_rate
    PJT,                // jump_time
    PTFJ,               // time_from_jump
    PPJ,                // 3 rows: px_jump, py_jump, pz_jump
    PF_COUNT = PPJ + 3
};

struct PartParams {
    int NX, NY, NZ;
    int64_t N;
    int64_t np, ld;          // particles, row stride
    double *s;               // state rows
    int *alive;
    const int *orig;         // reference-order index of each particle (rows may be kept in cell order)
    int *ndraw;
    int *bsum;
    uint64_t *rng;
    const double *u, *t;     // fields of the step: u interleaved [3N], t [N]
    const float4 *uT;        // the same in float, [N] (u_x, u_y, u_z, T) (float-storage lattices), else null
    uint32_t flags;          // ch4g_lagrange_flag
    uint32_t flags2;         // ch4g_lagrange_flag2
    double nu;               // kinematic viscosity (wake drag)
    int fluid, thermal, wall_f, wall_t;
    double Tw_bot, Tw_top;   // T_wall - T_ref
    double T_bot, T_top;     // the walls' temperatures a re-injected particle takes (LAGRANGE_TEMPERATURE)
    double inv_tau_polymer;  // 1 / tau_polymer (LAGRANGE_POLYMER)
    double log_001;          // log(0.01), host libm (the jump duration -jump_time log(0.01))
    int first;               // itime == 1 (Euler start)
    double dt;               // time_dt_lagr
    double grav[3];
    double beta_t;
};

// one sub-step: interpolation + move + walls (+ re-injection draws) + wrap
int launch_particles(const PartParams &P, hipStream_t s);
// Fields of the initial state at the particles from the FP64 planes P.u /
// P.t: with `velocity` u (and grad u) into the rows, v = u when `set_v`
// (LAGRANGE_INITIAL_VELOCITY_FLUID); otherwise t (and grad t) and t_p = t
// (LAGRANGE_INITIAL_TEMPERATURE_FLUID)
int launch_particle_finit(const PartParams &P, bool velocity, bool set_v, hipStream_t s);
// Two-way feedback (lagrange_feedback.c:11-256): every particle's deposits
// summed per node in reference particle order, written to dense planes
// fb[3][N] (double on FP64 lattices, float on float storage).  The particles
// are sorted by stencil cell (stable over reference order), a head array
// marks where each cell's list starts, and one pass over the nodes merges
// the lists of the eight cells that have the node as a corner.
struct FeedbackBufs {
    uint32_t *kin, *vin;     // [np] by reference index: stencil cell (or the sentinel), row
    uint32_t *kout, *vout;   // the same sorted by cell
    const uint32_t *old_keys;  // sorted cells of the previous call, whose heads are cleared
    int64_t old_n;
    int *head;               // [(NX+1)(NY+1)(NZ+1)] first sorted position of each cell, -1 if empty
    double *rec;             // [np][8] by row: fac f_p (3 components), wrapped position (3), reference index
    double *dep;             // [8][np][3] deposit of corner c of the t-th sorted particle
    uint32_t *ord;           // [np] reference index of the t-th sorted particle
    void *tmp;
    size_t tmp_bytes;
    int key_bits;
    void *fb;                // [3][N]
    int fb_float;
    double gv[3];            // gravity vector of the feedback (g_x, -g_y, g_z) or 0
};
int launch_feedback(const PartParams &P, const FeedbackBufs &B, hipStream_t s);
size_t feedback_sort_bytes(int64_t nitems, int key_bits);
// Rows, alive flags and reference-order indices of the particles reordered
// by cell (z, y, x of the position's node), for locality of the corner reads.
// The rows a reordering moves: every row a kernel of the configuration
// writes, and every row whose values differ between particles; a row that
// is uniform and never written holds the same value in both buffers.
struct PermRows {
    uint8_t r[PF_COUNT];
    int n;
};
struct ReorderBufs {
    PermRows rows;
    uint32_t *kin, *kout, *vin, *vout;   // cell keys / current indices, sorted
    void *tmp;
    size_t tmp_bytes;
    int key_bits;
    double *s_out;                       // state rows in the new order
    int *alive_out, *orig_out, *orig_in;
};
int launch_reorder(const PartParams &P, const ReorderBufs &B, hipStream_t s);

// The batched step works in wave tiles of BATCH_WAVE lanes: up to 64 x of
// one (y, z) row, so the row, its base state and its wall flags are uniform
// across the wave as in the single-lattice kernel; or, for "dense" lattices
// (dense_rows), 64 consecutive nodes in linear order, each lane decoding its
// own row.  A block runs BATCH_WAVES_PER_BLOCK tiles.
constexpr int BATCH_WAVE = 64;
constexpr int BATCH_WAVES_PER_BLOCK = 4;

// Lattices whose rows fill less than 90 % of the 64-lane waves that cover
// them (NX = 78: 61 %) are stepped with one node per lane in linear order,
// each lane decoding its own row, instead of row-aligned waves.
inline bool dense_rows(int NX)
{
    const int64_t lanes = (int64_t)(NX + BATCH_WAVE - 1) / BATCH_WAVE * BATCH_WAVE;
    return (int64_t)NX * 10 < lanes * 9;
}

// wave tiles of a lattice in the batched step
inline int64_t batch_tiles(int NX, int NY, int NZ)
{
    const int64_t N = (int64_t)NX * NY * NZ;
    if (dense_rows(NX)) return (N + BATCH_WAVE - 1) / BATCH_WAVE;
    return (int64_t)NY * NZ * ((NX + BATCH_WAVE - 1) / BATCH_WAVE);
}

// One lattice of a batched step: its constants and its two state buffers
// (and, for two-step launches, its two lag arrays).
struct BatchEntry {
    KParams P;
    void *f[2], *g[2];
    void *tlag, *taules;
    void *tlag2;        // second lag array (two-step launches), else null
    int cur;            // buffer holding the state at the start of the chunk
};
// One lattice taking part in a batched launch: entry index (low 29 bits;
// bit 31: parity of the lattice's population buffer at this launch; for
// two-step launches bit 30: a single step, bit 29: parity of its lag array)
// and its first tile in the launch (wave tiles, or 64 x 4 block tiles for
// two-step launches).  A launch's slots are sorted by tile0; the last one is
// a sentinel whose tile0 is the launch's tile count.
struct BatchSlot {
    uint32_t entry;
    uint32_t tile0;
};
constexpr uint32_t BATCH_ENTRY_MASK = 0x1fffffffu;
constexpr uint32_t BATCH_SINGLE = 1u << 30;
constexpr uint32_t BATCH_LAG_PARITY = 1u << 29;
// 64 x 4 block tiles of a 2-D lattice in the batched two-step launch
inline int64_t batch2_tiles(int NX, int NY) { return (int64_t)((NX + 63) / 64) * ((NY + 3) / 4); }
// one launch covers the lattices of a storage kind (FP32 float/double
// arithmetic, FP64) listed in its slots
// population sets of a batched launch (one kernel instance per kind)
enum { BATCH_FLUID = 1, BATCH_THERMAL = 2, BATCH_FLUID_THERMAL = 3 };
// gen: some lattice of the launch uses the HIT / LES / two-way paths
void launch_batch_fp32(const BatchEntry *tab, const BatchSlot *slots, int nslots, int64_t nblocks, int kind,
                       bool dense, bool gen, int arith64, bool ntl, hipStream_t s);
void launch_batch_fp64(const BatchEntry *tab, const BatchSlot *slots, int nslots, int64_t nblocks, int kind,
                       bool dense, hipStream_t s);
// two steps per launch for 2-D float-storage lattices (two_step_ok); nblocks
// = 64 x 4 tiles of the launch
void launch_batch2_fp32(const BatchEntry *tab, const BatchSlot *slots, int nslots, int64_t nblocks, int kind,
                        int arith64, bool ntl, hipStream_t s);

// Kernel launchers (stream-ordered).
void launch_step_fp64(const KParams &P, const double *fin, const double *gin, double *fout, double *gout,
                      double *tlag, double *taules, hipStream_t s);
void launch_gather_fp64(const KParams &P, const double *fin, const double *gin, const double *tlag,
                        double *fpost, double *gpost, double *dens, double *u, double *t, hipStream_t s);
void launch_step_fp32(const KParams &P, const float *fin, const float *gin, float *fout, float *gout, float *tlag,
                      float *taules, hipStream_t s);
void launch_gather_fp32(const KParams &P, const float *fin, const float *gin, const float *tlag, double *fpost,
                        double *gpost, double *dens, double *u, double *t, hipStream_t s);
// u [N][3] and t of the current step in float (particle interpolation on
// float-storage lattices), either may be null
void launch_fields_f32(const KParams &P, const float *fin, const float *gin, const float *tlag, float4 *uT,
                       hipStream_t s);

// Two steps in one launch for 2-D FP32-storage lattices (NZ = 1, NX a
// multiple of 64, steady steps without HIT tables, LES or particles): reads
// the state (f, g in; temperature lag tl_in), writes the state two steps on
// (f, g out; lag tl_out, a different array).
bool two_step_ok(const KParams &P);
void launch_two_steps_fp32(const KParams &P, const float *fin, const float *gin, float *fout, float *gout,
                           const float *tl_in, float *tl_out, hipStream_t s);

// Velocity of the step at every node into P.vel (the finite-difference LES
// closure reads it at the six neighbours); same gather and moments as the
// step kernel.
void launch_velocity_fp64(const KParams &P, const double *fin, const double *tlag, hipStream_t s);
// LB_FLUID_FORCING_NOZEROMODE means into f0[6] (scratch part: 512 * 6 doubles)
// the same for float-storage lattices, with the float force of the FP32 kernels
void launch_force_mean32(const KParams &P, bool step, const float *gin, const float *tlag, double *part, double *f0,
                         hipStream_t s);
void launch_force_mean64(const KParams &P, bool step, const double *gin, const double *tlag, double *part, double *f0,
                         hipStream_t s);
void launch_velocity_fp32(const KParams &P, const float *fin, const float *tlag, hipStream_t s);

// Initial state from a per-row fluid profile prow[NY][19] and a temperature
// field t0[N]; restart load from FP64 SoA staging.  dev = FP32-deviation.
void launch_init(const KParams &P, const double *prow, const double *t0, void *f, void *g, void *tlag, int dev,
                 hipStream_t s);
void launch_load(const KParams &P, const double *fs, const double *gs, void *f, void *g, void *tlag, int dev,
                 hipStream_t s);

// Reductions for dump_averages: FP64 per-block partials then a fixed-order
// final sum (deterministic for a given lattice size).
constexpr int NOUT = 26;
int reduce_blocks(const KParams &P);
void launch_averages(const KParams &P, const double *dens, const double *u, const double *t, double nu, double kappa,
                     double *partials, double *result, hipStream_t s);
// plane sums of the same terms along x, y, z: [NX|NY|NZ][NOUT], null to skip
void launch_rulers(const KParams &P, const double *dens, const double *u, const double *t, double nu, double kappa,
                   double *rx, double *ry, double *rz, hipStream_t s);

}  // namespace ch4g
