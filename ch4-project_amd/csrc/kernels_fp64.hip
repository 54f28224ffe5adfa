// FP64 fused stream+collide kernel and gather (hydro_fields) kernel.
//
// Arithmetic is written in the reference's evaluation order and this file is
// compiled with -ffp-contract=off (the reference builds with -std=c11, which
// disables contraction), so every field is bit-identical to the reference's
// FP64 step on the same inputs.
//
// State between steps (pull scheme): the post-collision populations of the
// last step (the reference's rhs_p/rhs_g after add_forcing), plus one lag
// scalar per node, tlag = the temperature of the previous step.  The kernel
// for step s first performs the streaming of step s-1 (lb.c:544-583) with the
// wall rules of boundary_conditions.c applied as a gather, reconstructs
// rho, u, T of step s-1 (lb.c:420-540; the u shift uses the force built from
// the lagged temperature, as main.c's ordering implies), then collides
// (fluid.c:1033-1225) and adds the Guo term (forcing.c:1424-1435) of step s.
#include "ch4g_internal.h"

namespace ch4g {

// typedef.h:286-291, strict left-to-right sums
__device__ __forceinline__ double m_sum(const double *p)
{
    return p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8] + p[9] + p[10] + p[11] + p[12] + p[13] +
           p[14] + p[15] + p[16] + p[17] + p[18];
}
__device__ __forceinline__ double m_x(const double *p)
{
    return p[1] - p[2] + p[7] + p[8] - p[9] - p[10] + p[11] - p[12] + p[13] - p[14];
}
__device__ __forceinline__ double m_y(const double *p)
{
    return p[3] - p[4] + p[7] - p[8] + p[9] - p[10] + p[15] + p[16] - p[17] - p[18];
}
__device__ __forceinline__ double m_z(const double *p)
{
    return p[5] - p[6] + p[11] + p[12] - p[13] - p[14] + p[15] - p[16] + p[17] - p[18];
}

// build_forcing for one node (forcing.c:361-667), terms added in source order;
// with NOZEROMODE the mean of those terms is removed (forcing.c:1123-1152)
// before the particle feedback is added (main.c:118-131).
__device__ __forceinline__ void force_at(const KParams &P, const HitTable &tab, const double *fb, const double *f0,
                                         uint32_t n, int x, int y, int z, double T, double F[3])
{
    F[0] = P.Fc[0];
    F[1] = P.Fc[1];
    F[2] = P.Fc[2];
    if (P.hit) {
#pragma unroll
        for (int ii = 0; ii < CH4G_HIT_NK; ii++) {
            double ax = tab.sx[ii * P.NX + x];
            double ay = tab.sy[ii * P.NY + y];
            double az = tab.sz[ii * P.NZ + z];
            F[0] += P.hit_fa[0][ii] * (ay + az);
            F[1] += P.hit_fa[1][ii] * (ax + az);
            F[2] += P.hit_fa[2][ii] * (ay + ax);
        }
    }
    if (P.buoy) {
        double temp = T - P.T0b;
        double fac = P.beta_t * temp;
        if (P.beta2_t != 0.0) fac += P.beta2_t * temp * temp;
        F[0] += fac * P.grav[0];
        F[1] += fac * P.grav[1];
        F[2] += fac * P.grav[2];
    }
    if (f0) {
        F[0] -= f0[0];
        F[1] -= f0[1];
        F[2] -= f0[2];
    }
    if (P.twoway) {
        const size_t N = (size_t)P.N;
        F[0] += fb[n];
        F[1] += fb[N + n];
        F[2] += fb[2 * N + n];
    }
}

// Streaming of the previous step as a pull, with halfway bounce-back at y
// walls (boundary_conditions.c:714-873: the ghost receives the node's own
// post-collision population of the opposite direction).  Branch-free so all
// loads are in flight together.
template <bool FIRST>
__device__ __forceinline__ void gather_fluid(const KParams &P, const Nbr &b, const double *__restrict__ in,
                                             uint32_t n, double f[Q])
{
    const size_t N = (size_t)P.S;  // plane stride
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (FIRST) {
            f[q] = in[(size_t)q * N + n];
        } else {
            const bool cross = P.wall_f && crosses_wall<q>(b);
            const double *base = in + (size_t)(cross ? invq(q) : q) * N;
            const uint32_t off = cross ? n : src_index<q>(b, P.NX, P.NY);
            f[q] = base[off];
        }
    });
}

// Fixed-temperature y walls (boundary_conditions.c:966-1078): the ghost
// receives the node's own post-collision population of the SAME direction
// plus w_q * 2((T_wall - T_ref) - t), t being the node temperature of the
// previous hydro_fields call (the lag scalar).
template <bool FIRST>
__device__ __forceinline__ void gather_thermal(const KParams &P, const Nbr &b, const double *__restrict__ in,
                                               uint32_t n, double Tlag, double g[Q])
{
    const size_t N = (size_t)P.S;  // plane stride
    const double fac_top = 2.0 * (P.Tw_top - Tlag);
    const double fac_bot = 2.0 * (P.Tw_bot - Tlag);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (FIRST) {
            g[q] = in[(size_t)q * N + n];
        } else {
            const bool cross = P.wall_t && crosses_wall<q>(b);
            const uint32_t off = cross ? n : src_index<q>(b, P.NX, P.NY);
            const double v = in[(size_t)q * N + off];
            if constexpr (cyq(q) > 0)
                g[q] = cross ? v + wq(q) * fac_bot : v;
            else if constexpr (cyq(q) < 0)
                g[q] = cross ? v + wq(q) * fac_top : v;
            else
                g[q] = v;
        }
    });
}

// equilibrium (lb.c:382-415)
__device__ __forceinline__ double feq_q(const KParams &P, int q, double rho, double ux, double uy, double uz,
                                        double u2)
{
    double cu = ((double)cxq(q) * ux + (double)cyq(q) * uy + (double)czq(q) * uz);
    return rho * wq(q) * (1.0 + P.invcs2 * cu + P.invtwocs4 * cu * cu - P.invtwocs2 * u2);
}

// tau_u_les, local-strain branch (les.c:20-30,:144-180,:253) with the
// stress_tensor assignments of gradients.c:42-69.
__device__ __forceinline__ double les_tau(const KParams &P, const double f[Q], double rho, double ux, double uy,
                                          double uz, double u2, double tau_prev)
{
    double xx = 0.0, yy = 0.0, zz = 0.0, xy = 0.0, xz = 0.0, yz = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        double d = f[q] - feq_q(P, q, rho, ux, uy, uz, u2);
        const double cx = cxq(q), cy = cyq(q), cz = czq(q);
        xx += cx * cx * d;
        yy += cy * cy * d;
        zz += cz * cz * d;
        xy += cx * cy * d;
        xz += cx * cz * d;
        yz += cy * cz * d;
    }
    (void)zz;
    const double yx = xy, zx = 0.0, zy = yz;
    zz = xz;
    double fac = -1.0 / (2 * P.cs2 * rho * tau_prev * 1.0);
    double s_norm = fabs(fac) * sqrt(2.0 * (xx * xx + xy * xy + xz * xz + yx * yx + yy * yy + yz * yz + zx * zx +
                                            zy * zy + zz * zz));
    double nu_les = s_norm * P.Cs2;
    return 3.0 * (nu_les + P.nu) + 0.5;
}

// tau_u_les, default finite-difference branch (les.c:31-43, :160-180, :253):
// gradient_vector of the stored velocity (gradients.c:78-499: centred
// differences with periodic images, one-sided second order at y walls,
// components along a size-one axis zeroed), |S| = sqrt(2 S_ij S_ij) summed
// in the reference's term order.  tens.ab = d u_a / d x_b.
__device__ __forceinline__ double les_tau_fd(const KParams &P, const double *__restrict__ vel, const Nbr &b,
                                             uint32_t n)
{
    const size_t N = (size_t)P.N;
    const int NX = P.NX, NY = P.NY;
    auto id = [&](int i, int j, int k) -> size_t { return ((size_t)k * NY + j) * NX + i; };
    double G[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double *u = vel + c * N;
        G[c][0] = (u[id(b.xp, b.y, b.z)] - u[id(b.xm, b.y, b.z)]) / 2.0;
        G[c][1] = (u[id(b.x, b.yp, b.z)] - u[id(b.x, b.ym, b.z)]) / 2.0;
        G[c][2] = (u[id(b.x, b.y, b.zp)] - u[id(b.x, b.y, b.zm)]) / 2.0;
        if (P.wall_f) {
            if (b.bot) G[c][1] = -1.5 * u[n] + 2.0 * u[id(b.x, 1, b.z)] + -0.5 * u[id(b.x, 2, b.z)];
            if (b.top) G[c][1] = 1.5 * u[n] + -2.0 * u[id(b.x, NY - 2, b.z)] + 0.5 * u[id(b.x, NY - 3, b.z)];
        }
        if (P.NX == 1) G[c][0] = 0.0;
        if (NY == 1) G[c][1] = 0.0;
        if (P.NZ == 1) G[c][2] = 0.0;
    }
    const double xx = G[0][0], xy = G[0][1], xz = G[0][2];
    const double yx = G[1][0], yy = G[1][1], yz = G[1][2];
    const double zx = G[2][0], zy = G[2][1], zz = G[2][2];
    double s_norm = sqrt(0.5 * ((xx + xx) * (xx + xx) + (xy + yx) * (xy + yx) + (xz + zx) * (xz + zx) +
                                (yx + xy) * (yx + xy) + (yy + yy) * (yy + yy) + (yz + zy) * (yz + zy) +
                                (zx + xz) * (zx + xz) + (zy + yz) * (zy + yz) + (zz + zz) * (zz + zz)));
    double nu_les = s_norm * P.Cs2;
    return 3.0 * (nu_les + P.nu) + 0.5;
}

// tau_t_les (les.c:273-297, METHOD_STREAMING, prandtl = 1.0)
__device__ __forceinline__ double les_tau_t(const KParams &P, double tau_les)
{
    const double prandtl = 1.0;
    return 3.0 * (P.kappa + (tau_les - 0.5 - 3 * P.nu) / (3 * prandtl)) + 0.5;
}

// hydro_fields of the previous step for one node (lb.c:420-540): density and
// velocity with the half force built from the lagged temperature.
__device__ __forceinline__ void fluid_velocity64(const KParams &P, const double f[Q], uint32_t n, int x, int y, int z,
                                                 double Tlag, double &rho, double &ux, double &uy, double &uz)
{
    rho = m_sum(f);
    if (P.forced) {
        double Fp[3];
        force_at(P, P.hit_prev, P.fb_prev, P.f0p, n, x, y, z, Tlag, Fp);
        ux = m_x(f) / rho + 0.5 * 1.0 * Fp[0];
        uy = m_y(f) / rho + 0.5 * 1.0 * Fp[1];
        uz = m_z(f) / rho + 0.5 * 1.0 * Fp[2];
    } else {
        ux = m_x(f) / rho;
        uy = m_y(f) / rho;
        uz = m_z(f) / rho;
    }
}

// One node of the fused step (shared by the single-lattice and the batched
// kernels, so a lattice advances bit-identically in either).
template <bool FIRST, bool FLUID, bool THERMAL>
__device__ __forceinline__ void step_node64(const KParams &P, const double *__restrict__ fin,
                                            const double *__restrict__ gin, double *__restrict__ fout,
                                            double *__restrict__ gout, double *__restrict__ tlag,
                                            double *__restrict__ taules, int x, int y, int z, uint32_t n)
{
    const size_t N = (size_t)P.S;  // plane stride
    const Nbr b = make_nbr(x, y, z, P.NX, P.NY, P.NZ);

    double Tlag = 0.0, T = 0.0;
    double g[Q], f[Q];
    if constexpr (THERMAL) {
        Tlag = tlag[n];
        gather_thermal<FIRST>(P, b, gin, n, Tlag, g);
    }
    if constexpr (FLUID) gather_fluid<FIRST>(P, b, fin, n, f);
    if constexpr (THERMAL) T = m_sum(g);

    double ux = 0.0, uy = 0.0, uz = 0.0;
    double tau_les = 0.0;
    if constexpr (FLUID) {
        double rho;
        fluid_velocity64(P, f, n, x, y, z, Tlag, rho, ux, uy, uz);
        const double u2 = (ux * ux + uy * uy + uz * uz);
        double invtau = P.invtau_u;
        if (P.les) {
            if (P.les_fd) {
                tau_les = les_tau_fd(P, (const double *)P.vel, b, n);
            } else {
                double tau_prev = (P.itime == 1) ? P.tau_u : taules[n];
                tau_les = les_tau(P, f, rho, ux, uy, uz, u2, tau_prev);
            }
            taules[n] = tau_les;
            invtau = 1.0 / tau_les;
        }
        double Fn[3] = {0.0, 0.0, 0.0};
        if (P.forced) force_at(P, P.hit_now, P.fb_now, P.f0n, n, x, y, z, T, Fn);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            double r = f[q];
            double fcoll = -invtau * (f[q] - feq_q(P, q, rho, ux, uy, uz, u2));
            r += fcoll;
            if (P.forced) {
                const double cx = cxq(q), cy = cyq(q), cz = czq(q);
                double cu = (cx * ux + cy * uy + cz * uz);
                double dx = (cx - ux) * P.invcs2 + cx * cu * P.invcs4;
                double dy = (cy - uy) * P.invcs2 + cy * cu * P.invcs4;
                double dz = (cz - uz) * P.invcs2 + cz * cu * P.invcs4;
                r += P.guo_fac * wq(q) * rho * Fn[0] * dx;
                r += P.guo_fac * wq(q) * rho * Fn[1] * dy;
                r += P.guo_fac * wq(q) * rho * Fn[2] * dz;
            }
            fout[(size_t)q * N + n] = r;
        }
    }

    if constexpr (THERMAL) {
        const double u2 = (ux * ux + uy * uy + uz * uz);
        double invtau_t = P.invtau_t;
        if constexpr (FLUID)
            if (P.les_t) invtau_t = 1.0 / les_tau_t(P, tau_les);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            double r = g[q];
            r += -invtau_t * (g[q] - feq_q(P, q, T, ux, uy, uz, u2));
            gout[(size_t)q * N + n] = r;
        }
        tlag[n] = T;
    }
}

// Row grid or, DENSE, a linear grid with one node per lane (see dense_rows).
template <bool FIRST, bool FLUID, bool THERMAL, bool DENSE>
__global__ void __launch_bounds__(256) k_step_fp64(KParams P, const double *__restrict__ fin,
                                                   const double *__restrict__ gin, double *__restrict__ fout,
                                                   double *__restrict__ gout, double *__restrict__ tlag,
                                                   double *__restrict__ taules)
{
    int x, y, z;
    uint32_t n;
    if constexpr (DENSE) {
        n = blockIdx.x * blockDim.x + threadIdx.x;
        if (n >= (uint32_t)P.N) return;
        const uint32_t row = n / (uint32_t)P.NX;
        x = (int)(n - row * (uint32_t)P.NX);
        y = (int)(row % (uint32_t)P.NY);
        z = (int)(row / (uint32_t)P.NY);
    } else {
        x = blockIdx.x * blockDim.x + threadIdx.x;
        if (x >= P.NX) return;
        y = blockIdx.y;
        z = blockIdx.z;
        n = ((uint32_t)z * P.NY + y) * P.NX + x;
    }
    step_node64<FIRST, FLUID, THERMAL>(P, fin, gin, fout, gout, tlag, taules, x, y, z, n);
}

// Velocity of the step at every node (the u[] that add_collision reads,
// fluid.c:1056-1080), for the finite-difference LES closure.
template <bool FIRST>
__global__ void __launch_bounds__(256) k_velocity64(KParams P, const double *__restrict__ fin,
                                                    const double *__restrict__ tlag)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= P.NX) return;
    const uint32_t n = ((uint32_t)z * P.NY + y) * P.NX + x;
    const size_t N = (size_t)P.N;
    const Nbr b = make_nbr(x, y, z, P.NX, P.NY, P.NZ);
    double f[Q];
    gather_fluid<FIRST>(P, b, fin, n, f);
    const double Tlag = P.thermal ? tlag[n] : 0.0;
    double rho, ux, uy, uz;
    fluid_velocity64(P, f, n, x, y, z, Tlag, rho, ux, uy, uz);
    double *vel = (double *)P.vel;
    vel[n] = ux;
    vel[N + n] = uy;
    vel[2 * N + n] = uz;
}

// Many lattices in one launch (see k_step_batch32).
template <bool FLUID, bool THERMAL, bool DENSE>
__global__ void __launch_bounds__(256) k_step_batch64(const BatchEntry *__restrict__ tab,
                                                      const BatchSlot *__restrict__ slots, int nslots)
{
    const uint32_t wt = __builtin_amdgcn_readfirstlane(blockIdx.x * BATCH_WAVES_PER_BLOCK + threadIdx.x / BATCH_WAVE);
    if (wt >= slots[nslots].tile0) return;
    int lo = 0, hi = nslots - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slots[mid].tile0 <= wt)
            lo = mid;
        else
            hi = mid - 1;
    }
    const BatchSlot sl = slots[lo];
    const BatchEntry &E = tab[sl.entry & BATCH_ENTRY_MASK];
    const KParams &P = E.P;
    const int in = E.cur ^ (int)(sl.entry >> 31);
    const double *fin = (const double *)E.f[in], *gin = (const double *)E.g[in];
    double *fout = (double *)E.f[in ^ 1], *gout = (double *)E.g[in ^ 1];
    double *tl = (double *)E.tlag, *tau = (double *)E.taules;
    // DENSE: wave tile t = nodes 64 t .. 64 t + 63 of the lattice, each lane
    // decoding its row; else tile t = up to 64 x of one row (row
    // coordinates wave-uniform), ceil(NX / 64) tiles per row
    const uint32_t t = wt - sl.tile0, lane = threadIdx.x % BATCH_WAVE;
    uint32_t n, x, y, z;
    if constexpr (DENSE) {
        n = t * BATCH_WAVE + lane;
        if (n >= (uint32_t)P.N) return;
        const uint32_t row = n / (uint32_t)P.NX;
        x = n - row * (uint32_t)P.NX;
        y = row % (uint32_t)P.NY;
        z = row / (uint32_t)P.NY;
    } else {
        const uint32_t xt = ((uint32_t)P.NX + BATCH_WAVE - 1) / BATCH_WAVE, row = t / xt;
        x = (t - row * xt) * BATCH_WAVE + lane;
        if (x >= (uint32_t)P.NX) return;
        y = row % (uint32_t)P.NY;
        z = row / (uint32_t)P.NY;
        n = row * (uint32_t)P.NX + x;
    }
    step_node64<false, FLUID, THERMAL>(P, fin, gin, fout, gout, tl, tau, (int)x, (int)y, (int)z, n);
}

// Streaming of the last step + hydro_fields, without a new collision:
// produces the reference's p, g (AoS order not applied here: SoA planes),
// dens, u and t of the current step.  Any output may be null.
__global__ void __launch_bounds__(256) k_gather_fp64(KParams P, const double *__restrict__ fin,
                                                     const double *__restrict__ gin,
                                                     const double *__restrict__ tlag, double *__restrict__ fpost,
                                                     double *__restrict__ gpost, double *__restrict__ dens,
                                                     double *__restrict__ u, double *__restrict__ t)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P.NX) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const uint32_t n = ((uint32_t)z * P.NY + y) * P.NX + x;
    const size_t N = (size_t)P.N;
    const Nbr b = make_nbr(x, y, z, P.NX, P.NY, P.NZ);
    double Tlag = 0.0;
    if (P.thermal) {
        double g[Q];
        Tlag = tlag[n];
        if (P.first)
            gather_thermal<true>(P, b, gin, n, Tlag, g);
        else
            gather_thermal<false>(P, b, gin, n, Tlag, g);
        if (gpost)
            for (int q = 0; q < Q; q++) gpost[q * N + n] = g[q];
        if (t) t[n] = m_sum(g);
    }
    if (P.fluid) {
        double f[Q];
        if (P.first)
            gather_fluid<true>(P, b, fin, n, f);
        else
            gather_fluid<false>(P, b, fin, n, f);
        if (fpost)
            for (int q = 0; q < Q; q++) fpost[q * N + n] = f[q];
        const double rho = m_sum(f);
        double ux, uy, uz;
        if (P.forced) {
            double F[3];
            force_at(P, P.hit_now, P.fb_now, P.f0n, n, x, y, z, Tlag, F);
            ux = m_x(f) / rho + 0.5 * 1.0 * F[0];
            uy = m_y(f) / rho + 0.5 * 1.0 * F[1];
            uz = m_z(f) / rho + 0.5 * 1.0 * F[2];
        } else {
            ux = m_x(f) / rho;
            uy = m_y(f) / rho;
            uz = m_z(f) / rho;
        }
        if (dens) dens[n] = rho;
        if (u) {
            u[3 * n] = ux;
            u[3 * n + 1] = uy;
            u[3 * n + 2] = uz;
        }
    }
}

static dim3 grid_of(const KParams &P, int bx) { return dim3((P.NX + bx - 1) / bx, P.NY, P.NZ); }
static int block_x(const KParams &P) { return P.NX >= 256 ? 256 : (P.NX >= 128 ? 128 : 64); }

void launch_step_fp64(const KParams &P, const double *fin, const double *gin, double *fout, double *gout,
                      double *tlag, double *taules, hipStream_t s)
{
    int bx = block_x(P);
    const bool dense = !P.first && dense_rows(P.NX);
    dim3 grid = grid_of(P, bx), block(bx);
    if (dense) {
        grid = dim3((unsigned)((P.N + 255) / 256));
        block = dim3(256);
    }
#define CH4G_LAUNCH64(F_, FL_, TH_)                                                                                   \
    do {                                                                                                              \
        if (!(F_) && dense)                                                                                           \
            hipLaunchKernelGGL((k_step_fp64<false, FL_, TH_, true>), grid, block, 0, s, P, fin, gin, fout, gout,     \
                               tlag, taules);                                                                         \
        else                                                                                                          \
            hipLaunchKernelGGL((k_step_fp64<F_, FL_, TH_, false>), grid, block, 0, s, P, fin, gin, fout, gout, tlag, \
                               taules);                                                                               \
    } while (0)
    if (P.fluid && P.thermal) {
        if (P.first) CH4G_LAUNCH64(true, true, true); else CH4G_LAUNCH64(false, true, true);
    } else if (P.fluid) {
        if (P.first) CH4G_LAUNCH64(true, true, false); else CH4G_LAUNCH64(false, true, false);
    } else {
        if (P.first) CH4G_LAUNCH64(true, false, true); else CH4G_LAUNCH64(false, false, true);
    }
#undef CH4G_LAUNCH64
}

void launch_velocity_fp64(const KParams &P, const double *fin, const double *tlag, hipStream_t s)
{
    int bx = block_x(P);
    dim3 grid = grid_of(P, bx);
    if (P.first)
        hipLaunchKernelGGL(k_velocity64<true>, grid, dim3(bx), 0, s, P, fin, tlag);
    else
        hipLaunchKernelGGL(k_velocity64<false>, grid, dim3(bx), 0, s, P, fin, tlag);
}

void launch_batch_fp64(const BatchEntry *tab, const BatchSlot *slots, int nslots, int64_t nblocks, int kind,
                       bool dense, hipStream_t s)
{
    dim3 grid((unsigned)nblocks), block(256);
#define CH4G_BATCH64(FL_, TH_)                                                                                        \
    do {                                                                                                              \
        if (dense)                                                                                                    \
            hipLaunchKernelGGL((k_step_batch64<FL_, TH_, true>), grid, block, 0, s, tab, slots, nslots);            \
        else                                                                                                          \
            hipLaunchKernelGGL((k_step_batch64<FL_, TH_, false>), grid, block, 0, s, tab, slots, nslots);           \
    } while (0)
    if (kind == BATCH_FLUID_THERMAL)
        CH4G_BATCH64(true, true);
    else if (kind == BATCH_FLUID)
        CH4G_BATCH64(true, false);
    else
        CH4G_BATCH64(false, true);
#undef CH4G_BATCH64
}

// LB_FLUID_FORCING_NOZEROMODE (forcing.c:1123-1152): sums over all nodes of
// the built force (no particle feedback) of the previous step, F(Tlag) with
// the previous HIT table, and of this step, F(T) with T gathered as the step
// kernel gathers it (STEP), or of the last step alone, F(Tlag) with the
// current table (the hydro-fields gather).  Fixed grid and a tree per block:
// the partial sums are the same on every run.
constexpr int FM_THREADS = 256, FM_BLOCKS = 512;

template <bool STEP, bool FIRST>
__global__ void __launch_bounds__(FM_THREADS) k_force_sum64(KParams P, const double *__restrict__ gin,
                                                            const double *__restrict__ tlag, double *__restrict__ part)
{
    P.twoway = 0;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int64_t N = P.N, stride = (int64_t)gridDim.x * FM_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * FM_THREADS + threadIdx.x; i < N; i += stride) {
        const uint32_t n = (uint32_t)i;
        const int x = (int)(i % P.NX), y = (int)((i / P.NX) % P.NY), z = (int)(i / ((int64_t)P.NX * P.NY));
        const double Tlag = P.thermal ? tlag[n] : 0.0;
        double F[3];
        if constexpr (STEP) {
            force_at(P, P.hit_prev, nullptr, nullptr, n, x, y, z, Tlag, F);
            for (int c = 0; c < 3; c++) acc[c] += F[c];
            double T = 0.0;
            if (P.thermal && P.buoy) {
                double g[Q];
                gather_thermal<FIRST>(P, make_nbr(x, y, z, P.NX, P.NY, P.NZ), gin, n, Tlag, g);
                T = m_sum(g);
            }
            force_at(P, P.hit_now, nullptr, nullptr, n, x, y, z, T, F);
        } else {
            force_at(P, P.hit_now, nullptr, nullptr, n, x, y, z, Tlag, F);
        }
        for (int c = 0; c < 3; c++) acc[3 + c] += F[c];
    }
    __shared__ double red[6][FM_THREADS];
    for (int c = 0; c < 6; c++) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = FM_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 6; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Block sums in block order, times 1/(NX NY NZ) as forcing.c:1140 forms it.
__global__ void k_force_mean64(int nblocks, double NX, double NY, double NZ, const double *__restrict__ part,
                               double *__restrict__ f0)
{
    const int c = threadIdx.x;
    if (c >= 6) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += part[(size_t)b * 6 + c];
    const double norm = 1.0 / (NX * NY * NZ);
    f0[c] = s * norm;
}

// f0[0..2]: mean force of the previous step, f0[3..5]: of this step (STEP);
// without STEP only f0[3..5] (the last step's).  part: FM_BLOCKS * 6 doubles.
void launch_force_mean64(const KParams &P, bool step, const double *gin, const double *tlag, double *part, double *f0,
                         hipStream_t s)
{
    const int64_t need = (P.N + FM_THREADS - 1) / FM_THREADS;
    const int nb = (int)(need < FM_BLOCKS ? need : FM_BLOCKS);
    if (!step)
        hipLaunchKernelGGL((k_force_sum64<false, false>), dim3(nb), dim3(FM_THREADS), 0, s, P, gin, tlag, part);
    else if (P.first)
        hipLaunchKernelGGL((k_force_sum64<true, true>), dim3(nb), dim3(FM_THREADS), 0, s, P, gin, tlag, part);
    else
        hipLaunchKernelGGL((k_force_sum64<true, false>), dim3(nb), dim3(FM_THREADS), 0, s, P, gin, tlag, part);
    hipLaunchKernelGGL(k_force_mean64, dim3(1), dim3(64), 0, s, nb, (double)P.NX, (double)P.NY, (double)P.NZ, part, f0);
}

void launch_gather_fp64(const KParams &P, const double *fin, const double *gin, const double *tlag, double *fpost,
                        double *gpost, double *dens, double *u, double *t, hipStream_t s)
{
    int bx = block_x(P);
    hipLaunchKernelGGL(k_gather_fp64, grid_of(P, bx), dim3(bx), 0, s, P, fin, gin, tlag, fpost, gpost, dens, u, t);
}

}  // namespace ch4g
