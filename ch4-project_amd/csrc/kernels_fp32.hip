// FP32-deviation fused stream+collide kernel and its gather kernel.
//
// Storage: df_i = f_i - w_i rho_b(y) and dg_i = g_i - w_i T_b(y) in float,
// SoA planes [q][z][y][x] (planes P.S >= N elements apart), where
// rho_b(y), T_b(y) is a per-row base state
// chosen on the host (uniform rho0/T0, or the conduction profile with its
// hydrostatic density for buoyancy-driven runs).  Subtracting the base keeps
// the stored numbers of the size of the dynamics, so float rounding acts on
// the signal and not on the O(1) background.  Streaming a population from row
// y' to row y adds w_i (rho_b(y') - rho_b(y)) (exact algebra; the row
// differences are tabulated).  Collision arithmetic is float:
//   drho = sum df,  j = sum c df,  rho = rho_b + drho,  u = j/rho + F_prev/2
//   df_eq = w (drho + rho (3 c.u + 9/2 (c.u)^2 - 3/2 u^2))
//   df*   = df - (df - df_eq)/tau + (1 - 1/(2 tau)) w rho (3 (c-u).F + 9 (c.u)(c.F))
// and the same for the temperature (T = T_b + dT, no forcing term).  Step
// semantics (lag scalar, wall rules, force lag) are those of kernels_fp64.hip.
#include "ch4g_internal.h"

namespace ch4g {

namespace {

// Population planes are streamed: every value is read once and written once
// per step.  Stores are always non-temporal; loads are non-temporal when the
// two lattices fit the last-level cache (measured faster there), plain
// otherwise (x-shifted reads of neighbouring wavefronts share L2 lines).
template <bool NTL>
__device__ __forceinline__ float ldp(const float *p)
{
    if constexpr (NTL)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Stores are non-temporal except where the single-step kernels update a
// lattice resident in the last-level cache (NT = false there: the next
// step's reads then find the state in the cache; channel 78^3 21.6 -> 18.8
// us per step, while the two-step kernel and streamed lattices are faster
// with non-temporal stores: 512x256 7.24 vs 7.45 us, 256^3 957 vs 999 us).
template <bool NT = true>
__device__ __forceinline__ void stp(float *p, float v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

__device__ __forceinline__ float wf(int q) { return q == 0 ? 1.f / 3.f : (q <= 6 ? 1.f / 18.f : 1.f / 36.f); }

// per-row constants of the base state
struct Row {
    float rb, tb, dr_up, dr_dn, dt_up, dt_dn, offT;
};

__device__ __forceinline__ Row row_of(const KParams &P, int y)
{
    const float *r = P.row32 + 8 * y;
    Row w;
    w.rb = r[0];
    w.tb = r[1];
    w.dr_up = r[2];
    w.dr_dn = r[3];
    w.dt_up = r[4];
    w.dt_dn = r[5];
    w.offT = r[6];
    return w;
}

// force of one node from its temperature deviation dT (buoyancy) and the
// HIT tables; same term order as build_forcing.  GEN: the HIT, LES and
// two-way coupling paths are compiled in (lattices without them run an
// instance without those paths and their register cost).
template <bool GEN>
__device__ __forceinline__ void force_at32(const KParams &P, const HitTable &tab, const float *fb, const double *f0,
                                           uint32_t n, int x, int y, int z, float theta, float F[3])
{
    F[0] = P.Fc_f[0];
    F[1] = P.Fc_f[1];
    F[2] = P.Fc_f[2];
    if (GEN && P.hit) {
#pragma unroll
        for (int ii = 0; ii < CH4G_HIT_NK; ii++) {
            float ax = (float)tab.sx[ii * P.NX + x];
            float ay = (float)tab.sy[ii * P.NY + y];
            float az = (float)tab.sz[ii * P.NZ + z];
            F[0] += (float)P.hit_fa[0][ii] * (ay + az);
            F[1] += (float)P.hit_fa[1][ii] * (ax + az);
            F[2] += (float)P.hit_fa[2][ii] * (ay + ax);
        }
    }
    if (P.buoy) {
        float fac = P.beta_t_f * theta + P.beta2_t_f * theta * theta;
        F[0] += fac * P.grav_f[0];
        F[1] += fac * P.grav_f[1];
        F[2] += fac * P.grav_f[2];
    }
    if (GEN && f0) {
        // NOZEROMODE: the mean of the built force removed (forcing.c:1123-1152)
        F[0] -= (float)f0[0];
        F[1] -= (float)f0[1];
        F[2] -= (float)f0[2];
    }
    if (GEN && P.twoway) {
        // per-node sums of the deposits in particle order, rounded to float
        const size_t N = (size_t)P.N;
        F[0] += fb[n];
        F[1] += fb[N + n];
        F[2] += fb[2 * N + n];
    }
}

// Streaming of the previous step (pull) with halfway bounce-back at y walls.
// Branch-free: every load address is computed first (wall crossings are a
// wave-uniform select), so all 19 loads are in flight together.
template <bool FIRST, bool NTL>
__device__ __forceinline__ void gather_f32(const KParams &P, const Row &R, const Nbr &b, const float *__restrict__ in,
                                           uint32_t n, float f[Q])
{
    const size_t N = (size_t)P.S;  // plane stride
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (FIRST) {
            f[q] = ldp<NTL>(in + (size_t)q * N + n);
        } else {
            const bool cross = P.wall_f && crosses_wall<q>(b);
            const float *base = in + (size_t)(cross ? invq(q) : q) * N;
            const uint32_t off = cross ? n : src_index<q>(b, P.NX, P.NY);
            float v = ldp<NTL>(base + off);
            if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? 0.f : R.dr_up, v);
            if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? 0.f : R.dr_dn, v);
            f[q] = v;
        }
    });
}

// Fixed-T y walls: same-direction population + w 2((T_w - T_ref) - T_lag).
template <bool FIRST, bool NTL>
__device__ __forceinline__ void gather_g32(const KParams &P, const Row &R, const Nbr &b, const float *__restrict__ in,
                                           uint32_t n, float dTlag, float g[Q])
{
    const size_t N = (size_t)P.S;  // plane stride
    const float fac_top = 2.f * (P.cw_top_f - dTlag);
    const float fac_bot = 2.f * (P.cw_bot_f - dTlag);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (FIRST) {
            g[q] = ldp<NTL>(in + (size_t)q * N + n);
        } else {
            const bool cross = P.wall_t && crosses_wall<q>(b);
            const uint32_t off = cross ? n : src_index<q>(b, P.NX, P.NY);
            float v = ldp<NTL>(in + (size_t)q * N + off);
            if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? fac_bot : R.dt_up, v);
            if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? fac_top : R.dt_dn, v);
            g[q] = v;
        }
    });
}

// The same gathers with 32-bit byte offsets from the array base (arrays of
// less than 4 GiB): a load is the array's scalar base plus one 32-bit vector
// offset instead of a 64-bit address per lane and population.
template <bool NTL>
__device__ __forceinline__ float ldo(const float *base, uint32_t byteoff)
{
    return ldp<NTL>((const float *)((const char *)base + byteoff));
}

template <bool NTL>
__device__ __forceinline__ void gather_f32_o32(const KParams &P, const Row &R, const Nbr &b,
                                               const float *__restrict__ in, uint32_t n, float f[Q])
{
    const uint32_t S4 = (uint32_t)P.S * 4u;
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const bool cross = P.wall_f && crosses_wall<q>(b);
        const uint32_t off = cross ? (uint32_t)invq(q) * S4 + n * 4u : (uint32_t)q * S4 + src_index<q>(b, P.NX, P.NY) * 4u;
        float v = ldo<NTL>(in, off);
        if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? 0.f : R.dr_up, v);
        if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? 0.f : R.dr_dn, v);
        f[q] = v;
    });
}

template <bool NTL>
__device__ __forceinline__ void gather_g32_o32(const KParams &P, const Row &R, const Nbr &b,
                                               const float *__restrict__ in, uint32_t n, float dTlag, float g[Q])
{
    const uint32_t S4 = (uint32_t)P.S * 4u;
    const float fac_top = 2.f * (P.cw_top_f - dTlag);
    const float fac_bot = 2.f * (P.cw_bot_f - dTlag);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const bool cross = P.wall_t && crosses_wall<q>(b);
        const uint32_t off = (uint32_t)q * S4 + (cross ? n : src_index<q>(b, P.NX, P.NY)) * 4u;
        float v = ldo<NTL>(in, off);
        if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? fac_bot : R.dt_up, v);
        if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? fac_top : R.dt_dn, v);
        g[q] = v;
    });
}

__device__ __forceinline__ float sum19(const float *p)
{
    // pairwise: keeps the float rounding of the moment small
    float a = (p[1] + p[2]) + (p[3] + p[4]) + (p[5] + p[6]);
    float b = ((p[7] + p[8]) + (p[9] + p[10])) + ((p[11] + p[12]) + (p[13] + p[14]));
    float c = (p[15] + p[16]) + (p[17] + p[18]);
    return p[0] + (a + (b + c));
}

}  // namespace

// representative q of the k-th (q, inv q) pair
__host__ __device__ constexpr int pair_q(int k)
{
    constexpr int v[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
    return v[k];
}

// c_q . (x, y, z) for c_q with components in {0, +-1}, written as the one
// or two signed terms (no multiplications by zero)
template <int q, typename A>
__device__ __forceinline__ A cdot(A x, A y, A z)
{
    constexpr int a = cxq(q), b = cyq(q), c = czq(q);
    const A tx = a > 0 ? x : -x, ty = b > 0 ? y : -y, tz = c > 0 ? z : -z;
    if constexpr (a != 0 && b != 0)
        return tx + ty;
    else if constexpr (a != 0 && c != 0)
        return tx + tz;
    else if constexpr (b != 0 && c != 0)
        return ty + tz;
    else if constexpr (a != 0)
        return tx;
    else if constexpr (b != 0)
        return ty;
    else
        return tz;
}

// a step constant in the collision's arithmetic type: the host's float copy
// for float arithmetic, the double value otherwise
template <typename A>
__device__ __forceinline__ A kc(double d, float f)
{
    if constexpr (sizeof(A) == 4)
        return f;
    else
        return (A)d;
}

template <typename A>
__device__ __forceinline__ A wa(int q)
{
    return q == 0 ? A(1) / A(3) : (q <= 6 ? A(1) / A(18) : A(1) / A(36));
}

// Where a node's post-collision state goes.  NodeOut: the global planes of
// the single-step kernels (non-temporal stores when NT), with the previous
// LES relaxation time read from tau_src[tau_at]; NodeOut32 (below): the
// second step of k_step2_fp32.  TileOut: the LDS tile of k_step2_fp32's
// first step: fluid planes from 0, temperature planes from gofs, then the
// temperature lag, `stride` nodes apart.
template <bool NT>
struct NodeOut {
    float *__restrict__ f;
    float *__restrict__ g;
    float *__restrict__ tlag;
    float *__restrict__ taules;
    const float *tau_src;
    uint32_t n, tau_at;
    size_t S;  // plane stride
    __device__ void put_f(int q, float v) const { stp<NT>(f + (size_t)q * S + n, v); }
    __device__ void put_g(int q, float v) const { stp<NT>(g + (size_t)q * S + n, v); }
    __device__ void put_tlag(float v) const { tlag[n] = v; }
    __device__ float get_tau() const { return tau_src[tau_at]; }
    __device__ void put_tau(float v) const { taules[n] = v; }
};

// NodeOut with 32-bit byte offsets (arrays of less than 4 GiB)
struct NodeOut32 {
    float *__restrict__ f;
    float *__restrict__ g;
    float *__restrict__ tlag;
    uint32_t n4, S4;  // byte offset of the node, plane stride in bytes
    __device__ void put_f(int q, float v) const { stp((float *)((char *)f + ((uint32_t)q * S4 + n4)), v); }
    __device__ void put_g(int q, float v) const { stp((float *)((char *)g + ((uint32_t)q * S4 + n4)), v); }
    __device__ void put_tlag(float v) const { *(float *)((char *)tlag + n4) = v; }
    __device__ float get_tau() const { return 0.f; }  // no LES on the two-step path
    __device__ void put_tau(float) const {}
};

struct TileOut {
    float *t;
    int at, stride, gofs;  // tile node, plane stride, first temperature plane
    __device__ void put_f(int q, float v) const { t[q * stride + at] = v; }
    __device__ void put_g(int q, float v) const { t[(gofs + q) * stride + at] = v; }
    __device__ void put_tlag(float v) const { t[(gofs + Q) * stride + at] = v; }
    __device__ float get_tau() const { return 0.f; }  // no LES on the two-step path
    __device__ void put_tau(float) const {}
};

// Moments of the gathered fluid populations (deviation form): density
// deviation and density, momentum, and the velocity with the lagged half
// force (hydro_fields, lb.c:420-540).
template <typename A>
struct Moments {
    A fa[Q];
    A drho, rho, jx, jy, jz, irho, ux, uy, uz;
};

template <typename A, bool GEN>
__device__ __forceinline__ void fluid_moments(const KParams &P, const Row &R, int x, int y, int z, uint32_t n,
                                              const float f[Q], float dTlag, Moments<A> &m)
{
    A *fa = m.fa;
    for (int q = 0; q < Q; q++) fa[q] = (A)f[q];
    if constexpr (sizeof(A) == 4) {
        m.drho = sum19(f);
    } else {
        m.drho = 0;
        for (int q = 0; q < Q; q++) m.drho += fa[q];
    }
    m.rho = (sizeof(A) == 4 ? (A)R.rb : (A)P.row64[2 * y]) + m.drho;
    // momentum from (q, inv q) differences: exact for nearby values
    m.jx = (fa[1] - fa[2]) + (fa[7] - fa[10]) + (fa[8] - fa[9]) + (fa[11] - fa[14]) + (fa[13] - fa[12]);
    m.jy = (fa[3] - fa[4]) + (fa[7] - fa[10]) + (fa[9] - fa[8]) + (fa[15] - fa[18]) + (fa[16] - fa[17]);
    m.jz = (fa[5] - fa[6]) + (fa[11] - fa[14]) + (fa[12] - fa[13]) + (fa[15] - fa[18]) + (fa[17] - fa[16]);
    m.irho = A(1) / m.rho;
    m.ux = m.jx * m.irho;
    m.uy = m.jy * m.irho;
    m.uz = m.jz * m.irho;
    if (P.forced) {
        float Fp[3];
        force_at32<GEN>(P, P.hit_prev, P.fbf_prev, P.f0p, n, x, y, z, R.offT + dTlag, Fp);
        m.ux += A(0.5) * Fp[0];
        m.uy += A(0.5) * Fp[1];
        m.uz += A(0.5) * Fp[2];
    }
}

// tau_u_les, finite-difference branch (les.c:31-43): gradient_vector of the
// stored velocity (centred, periodic images, one-sided second order at y
// walls, size-one axes zeroed), |S| = sqrt(2 S_ij S_ij).
template <typename A>
__device__ __forceinline__ A les_tau_fd32(const KParams &P, const float *__restrict__ vel, int x, int y, int z,
                                          uint32_t n)
{
    const Nbr b = make_nbr(x, y, z, P.NX, P.NY, P.NZ);
    const size_t N = (size_t)P.N;
    const int NX = P.NX, NY = P.NY;
    auto id = [&](int i, int j, int k) -> size_t { return ((size_t)k * NY + j) * NX + i; };
    A G[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float *u = vel + c * N;
        G[c][0] = A(0.5) * ((A)u[id(b.xp, y, z)] - (A)u[id(b.xm, y, z)]);
        G[c][1] = A(0.5) * ((A)u[id(x, b.yp, z)] - (A)u[id(x, b.ym, z)]);
        G[c][2] = A(0.5) * ((A)u[id(x, y, b.zp)] - (A)u[id(x, y, b.zm)]);
        if (P.wall_f) {
            if (b.bot) G[c][1] = A(-1.5) * (A)u[n] + A(2) * (A)u[id(x, 1, z)] - A(0.5) * (A)u[id(x, 2, z)];
            if (b.top) G[c][1] = A(1.5) * (A)u[n] - A(2) * (A)u[id(x, NY - 2, z)] + A(0.5) * (A)u[id(x, NY - 3, z)];
        }
        if (P.NX == 1) G[c][0] = A(0);
        if (NY == 1) G[c][1] = A(0);
        if (P.NZ == 1) G[c][2] = A(0);
    }
    A ss = A(0);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const A e = G[c][d] + G[d][c];
            ss += e * e;
        }
    return A(3) * (sqrt(A(0.5) * ss) * (A)P.Cs2 + (A)P.nu) + A(0.5);
}

// Fluid part of the step for one node: moments, velocity with the lagged
// half force, optional LES, Guo-forced BGK in FMA form, restoration of the
// collision invariants, non-temporal stores; the velocity used is returned
// in ux, uy, uz and the LES relaxation time (if any) in tau_les.
template <typename A, bool GEN, class Out>
__device__ __forceinline__ void collide_fluid(const KParams &P, const Row &R, int x, int y, int z, uint32_t n,
                                              const float f[Q], float dTlag, A dT, const Out &out, A &ux, A &uy,
                                              A &uz, A &tau_les)
{
    {
        Moments<A> m;
        fluid_moments<A, GEN>(P, R, x, y, z, n, f, dTlag, m);
        const A *fa = m.fa;
        const A drho = m.drho, rho = m.rho, jx = m.jx, jy = m.jy, jz = m.jz, irho = m.irho;
        ux = m.ux;
        uy = m.uy;
        uz = m.uz;
        const A u2 = ux * ux + uy * uy + uz * uz;
        // weight-class products (w rho, w drho) shared by the populations
        const A wr[3] = {wa<A>(0) * rho, wa<A>(1) * rho, wa<A>(7) * rho};
        const A wd[3] = {wa<A>(0) * drho, wa<A>(1) * drho, wa<A>(7) * drho};
        A omega = kc<A>(P.invtau_u, P.invtau_u_f);
        if (GEN && P.les && P.les_fd) {
            const A tau = les_tau_fd32<A>(P, (const float *)P.vel, x, y, z, n);
            out.put_tau((float)tau);
            tau_les = tau;
            omega = A(1) / tau;
        } else if (GEN && P.les) {
            A xx = 0, yy = 0, xy = 0, xz = 0, yz = 0;
#pragma unroll
            for (int q = 1; q < Q; q++) {
                const int wc = q <= 6 ? 1 : 2;
                const A cx = cxq(q), cy = cyq(q), cz = czq(q);
                const A cu = cx * ux + cy * uy + cz * uz;
                const A eq = fma(wr[wc], A(3) * cu + A(4.5) * cu * cu - A(1.5) * u2, wd[wc]);
                const A d = fa[q] - eq;
                xx += cx * cx * d;
                yy += cy * cy * d;
                xy += cx * cy * d;
                xz += cx * cz * d;
                yz += cy * cz * d;
            }
            // stress_tensor of gradients.c:65-67: S.zz takes S.xz, S.zx stays 0
            const A ss = xx * xx + A(2) * xy * xy + xz * xz + yy * yy + A(2) * yz * yz + xz * xz;
            A tau_prev = (P.itime == 1) ? (A)P.tau_u : (A)out.get_tau();
            A fac = A(1) / (A(2) * (A)P.cs2 * rho * tau_prev);
            A tau = A(3) * (fac * sqrt(A(2) * ss) * (A)P.Cs2 + (A)P.nu) + A(0.5);
            out.put_tau((float)tau);
            tau_les = tau;
            omega = A(1) / tau;
        }
        float Fn[3] = {0.f, 0.f, 0.f};
        if (P.forced) force_at32<GEN>(P, P.hit_now, P.fbf_now, P.f0n, n, x, y, z, R.offT + (float)dT, Fn);
        const A uF = ux * Fn[0] + uy * Fn[1] + uz * Fn[2];
        // Guo term folded into the equilibrium: f* = f - omega (f - eqG) with
        // eqG = f_eq + G/omega, so the O(population) value is rounded twice.
        // Pair form: for the opposite pair (q, inv q), eqG = E +- D with
        //   E = w rho (9/2 (c.u)^2 - 3/2 u^2 + kg (9 (c.u)(c.F) - 3 u.F)) + w drho
        //   D = w rho (3 c.u + 3 kg c.F),   kg = (1 - omega/2) / omega
        // (E even, D odd in c); the rest population has c = 0.
        const A kg = kc<A>(P.guo_fac, P.guo_fac_f) / omega;
        const A k3 = A(3) * kg, k9 = A(9) * kg;
        const A c0 = -A(1.5) * u2 - k3 * uF;
        A r[Q];
        {
            const A e = fma(wr[0], c0, wd[0]);
            r[0] = fma(-omega, fa[0] - e, fa[0]);
        }
        static_for<0, 9>([&](auto kk) {
            constexpr int q = pair_q(decltype(kk)::value), qb = invq(q);
            constexpr int wc = q <= 6 ? 1 : 2;
            const A cu = cdot<q>(ux, uy, uz);
            A sym = fma(A(4.5) * cu, cu, c0);
            A anti = A(3) * cu;
            if (P.forced) {
                const A cF = cdot<q>((A)Fn[0], (A)Fn[1], (A)Fn[2]);
                sym = fma(k9 * cu, cF, sym);
                anti = fma(k3, cF, anti);
            }
            const A E = fma(wr[wc], sym, wd[wc]), D = wr[wc] * anti;
            r[q] = fma(-omega, fa[q] - (E + D), fa[q]);
            r[qb] = fma(-omega, fa[qb] - (E - D), fa[qb]);
        });
        if constexpr (sizeof(A) == 4) {
            // Float weights miss the lattice identities (18 fl(1/18) != 1, ...)
            // by ~1e-8, which would scale mass and momentum systematically at
            // every step.  Restore the exact collision invariants: mass is
            // conserved; momentum becomes j + omega rho F_prev/2 + guo rho F,
            // both targets evaluated at velocity magnitude.
            A Fp2[3] = {A(0), A(0), A(0)};
            if (P.forced) {
                Fp2[0] = (ux - jx * irho) * rho;  // = rho F_prev / 2
                Fp2[1] = (uy - jy * irho) * rho;
                Fp2[2] = (uz - jz * irho) * rho;
            }
            const A gf = kc<A>(P.guo_fac, P.guo_fac_f);
            const A tx = jx + omega * Fp2[0] + gf * rho * (A)Fn[0];
            const A ty = jy + omega * Fp2[1] + gf * rho * (A)Fn[1];
            const A tz = jz + omega * Fp2[2] + gf * rho * (A)Fn[2];
            const A mx = (r[1] - r[2]) + (r[7] - r[10]) + (r[8] - r[9]) + (r[11] - r[14]) + (r[13] - r[12]);
            const A my = (r[3] - r[4]) + (r[7] - r[10]) + (r[9] - r[8]) + (r[15] - r[18]) + (r[16] - r[17]);
            const A mz = (r[5] - r[6]) + (r[11] - r[14]) + (r[12] - r[13]) + (r[15] - r[18]) + (r[17] - r[16]);
            const A hx = A(0.5) * (tx - mx), hy = A(0.5) * (ty - my), hz = A(0.5) * (tz - mz);
            r[1] += hx;
            r[2] -= hx;
            r[3] += hy;
            r[4] -= hy;
            r[5] += hz;
            r[6] -= hz;
            r[0] += drho - sum19(r);
        }
#pragma unroll
        for (int q = 0; q < Q; q++) out.put_f(q, (float)r[q]);
    }

}

// Temperature part of the step for one node: BGK with the fluid velocity,
// exact conservation of the temperature, non-temporal stores.
template <typename A, bool GEN, class Out>
__device__ __forceinline__ void collide_thermal(const KParams &P, const Row &R, int y, const float g[Q], A dT, A ux,
                                                A uy, A uz, A tau_les, const Out &out)
{
    {
        const A u2 = ux * ux + uy * uy + uz * uz;
        const A T = (sizeof(A) == 4 ? (A)R.tb : (A)P.row64[2 * y + 1]) + dT;
        // tau_t_les (les.c:273-297) with prandtl 1: tau_t = 3(kappa + nu_les) + 1/2
        const A om = (GEN && P.les_t && P.fluid) ? A(1) / (A(3) * ((A)P.kappa + (tau_les - A(0.5) - A(3) * (A)P.nu) / A(3)) +
                                                     A(0.5))
                                          : kc<A>(P.invtau_t, P.invtau_t_f);
        const A wT[3] = {wa<A>(0) * T, wa<A>(1) * T, wa<A>(7) * T};
        const A wdT[3] = {wa<A>(0) * dT, wa<A>(1) * dT, wa<A>(7) * dT};
        // pair form as for the fluid (no forcing term): eq = E +- D with
        // E = w T (9/2 (c.u)^2 - 3/2 u^2) + w dT, D = w T 3 c.u
        const A c0 = -A(1.5) * u2;
        A r[Q];
        {
            const A e = fma(wT[0], c0, wdT[0]);
            const A g0 = (A)g[0];
            r[0] = fma(-om, g0 - e, g0);
        }
        static_for<0, 9>([&](auto kk) {
            constexpr int q = pair_q(decltype(kk)::value), qb = invq(q);
            constexpr int wc = q <= 6 ? 1 : 2;
            const A cu = cdot<q>(ux, uy, uz);
            const A E = fma(wT[wc], fma(A(4.5) * cu, cu, c0), wdT[wc]), D = wT[wc] * (A(3) * cu);
            const A ga = (A)g[q], gb = (A)g[qb];
            r[q] = fma(-om, ga - (E + D), ga);
            r[qb] = fma(-om, gb - (E - D), gb);
        });
        // the collision conserves the temperature exactly (float weights do not)
        if constexpr (sizeof(A) == 4) r[0] += dT - sum19(r);
#pragma unroll
        for (int q = 0; q < Q; q++) out.put_g(q, (float)r[q]);
        out.put_tlag((float)dT);
    }
}

// One node of the fused step; shared by the single-lattice kernel and the
// batched kernel, so a lattice advances bit-identically in either.
// A = arithmetic type of the collision (float for CH4G_FP32_DEV, double for
// CH4G_FP32_STORE); storage is float in both.
template <bool FIRST, bool FLUID, bool THERMAL, typename A, bool NTL, bool GEN>
__device__ __forceinline__ void step_node32(const KParams &P, const float *__restrict__ fin,
                                            const float *__restrict__ gin, float *__restrict__ fout,
                                            float *__restrict__ gout, float *__restrict__ tlag,
                                            float *__restrict__ taules, int x, int y, int z, uint32_t n)
{
    const Row R = row_of(P, y);
    const Nbr b = make_nbr(x, y, z, P.NX, P.NY, P.NZ);

    float dTlag = 0.f;
    float g[Q];
    if constexpr (THERMAL) {
        dTlag = tlag[n];
        gather_g32<FIRST, NTL>(P, R, b, gin, n, dTlag, g);
    }
    float f[Q];
    if constexpr (FLUID) gather_f32<FIRST, NTL>(P, R, b, fin, n, f);
    A dT = A(0);
    if constexpr (THERMAL) {
        if constexpr (sizeof(A) == 4) {
            dT = sum19(g);
        } else {
            for (int q = 0; q < Q; q++) dT += (A)g[q];
        }
    }
    A ux = 0, uy = 0, uz = 0, tau_les = A(0);
    const NodeOut<!NTL> out{fout, gout, tlag, taules, taules, n, n, (size_t)P.S};
    if constexpr (FLUID) collide_fluid<A, GEN>(P, R, x, y, z, n, f, dTlag, dT, out, ux, uy, uz, tau_les);
    if constexpr (THERMAL) collide_thermal<A, GEN>(P, R, y, g, dT, ux, uy, uz, tau_les, out);
    if constexpr (GEN && !FIRST) {
        // hydro fields of the previous step (hydro_fields, lb.c:420-540) for
        // the particles: the gathered moments and the velocity with its half
        // force, as the collision used them
        if (P.uT_out) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if constexpr (FLUID) {
                v.x = (float)ux;
                v.y = (float)uy;
                v.z = (float)uz;
            }
            if constexpr (THERMAL) v.w = (float)((sizeof(A) == 4 ? (A)R.tb : (A)P.row64[2 * y + 1]) + dT);
            P.uT_out[n] = v;
        }
    }
}

// Row grid (blocks along x of one (y, z) row: row quantities block-uniform)
// or, DENSE, a linear grid with one node per lane (see dense_rows).
template <bool FIRST, bool FLUID, bool THERMAL, typename A, bool NTL, bool DENSE, bool GEN>
__global__ void __launch_bounds__(256) k_step_fp32(KParams P, const float *__restrict__ fin,
                                                   const float *__restrict__ gin, float *__restrict__ fout,
                                                   float *__restrict__ gout, float *__restrict__ tlag,
                                                   float *__restrict__ taules)
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
    step_node32<FIRST, FLUID, THERMAL, A, NTL, GEN>(P, fin, gin, fout, gout, tlag, taules, x, y, z, n);
}

// ---------------------------------------------------------------------------
// Two steps per launch for 2-D lattices (NZ = 1).  A block owns a 64 x TY
// tile (TY = 4 or 5; the last tile of a row is narrower when NX is not a
// multiple of 64) and has one thread per node of the tile with its one-node
// halo.  First step: the tile and the halo (periodic
// images; rows beyond a wall are computed from their periodic images and
// never read) are advanced from the global state into LDS with the
// single-step per-node code.  Second step: the tile's nodes gather from LDS
// (same wall rules and row terms) and collide into the global output.  Per
// two steps the state is read 1.55 (TY = 4) or 1.44 (TY = 5) times and
// written once instead of twice each, and one launch boundary (ramp-up,
// drain) is paid instead of two.  The temperature lag is read from tl_in and
// written to tl_out (halo nodes of one block are tile nodes of another).
constexpr int T2_TX = 64, T2_W = T2_TX + 2;
template <int TY>
struct Tile2 {
    static constexpr int H = TY + 2, M = T2_W * H;
    static constexpr int THREADS = (M + 63) / 64 * 64;  // 448 (TY = 4), 512 (TY = 5)
    static_assert(THREADS >= T2_TX * TY && THREADS <= 512, "two-step block");
};

template <int M>
__device__ __forceinline__ void gather_f_tile(const KParams &P, const Row &R, const Nbr &b, const float *t, int at,
                                              float f[Q])
{
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const bool cross = P.wall_f && crosses_wall<q>(b);
        float v = cross ? t[invq(q) * M + at] : t[q * M + at - cxq(q) - cyq(q) * T2_W];
        if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? 0.f : R.dr_up, v);
        if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? 0.f : R.dr_dn, v);
        f[q] = v;
    });
}

template <int M>
__device__ __forceinline__ void gather_g_tile(const KParams &P, const Row &R, const Nbr &b, const float *t, int gofs,
                                              int at, float dTlag, float g[Q])
{
    const float fac_top = 2.f * (P.cw_top_f - dTlag);
    const float fac_bot = 2.f * (P.cw_bot_f - dTlag);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const bool cross = P.wall_t && crosses_wall<q>(b);
        float v = t[(gofs + q) * M + (cross ? at : at - cxq(q) - cyq(q) * T2_W)];
        if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? fac_bot : R.dt_up, v);
        if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? fac_top : R.dt_dn, v);
        g[q] = v;
    });
}

// LDS planes of a two-step tile: fluid populations, then temperature
// populations and the temperature lag
template <bool FLUID, bool THERMAL>
constexpr int tile2_planes()
{
    return (FLUID ? Q : 0) + (THERMAL ? Q + 1 : 0);
}

// The two steps of one 64 x TY tile whose first node is (x0, y0), through
// the block's LDS array t (shared by the single-lattice and batched kernels).
template <bool FLUID, bool THERMAL, typename A, bool NTL, int TY>
__device__ __forceinline__ void step2_tile(const KParams &P, const float *__restrict__ fin,
                                           const float *__restrict__ gin, float *__restrict__ fout,
                                           float *__restrict__ gout, const float *__restrict__ tl_in,
                                           float *__restrict__ tl_out, int x0, int y0, float *t)
{
    constexpr int M = Tile2<TY>::M;
    constexpr int GOFS = FLUID ? Q : 0;
    const int NX = P.NX, NY = P.NY;
    // columns of this tile: 64, or fewer in the last tile of a row when NX
    // is not a multiple of 64 (its right halo column is then x = 0)
    const int w = NX - x0 < T2_TX ? NX - x0 : T2_TX;
    // first step: tile + halo into LDS, one node per thread (LDS columns
    // beyond the halo of a narrower last tile are never read: skipped)
    if (const int m = threadIdx.x; m < M && m % T2_W <= w + 1) {
        const int lx = m % T2_W, ly = m / T2_W;
        int x = x0 - 1 + lx;  // in [-1, NX]
        x = x < 0 ? x + NX : (x >= NX ? x - NX : x);
        int y = (y0 - 1 + ly) % NY;
        y = y < 0 ? y + NY : y;
        const uint32_t n = (uint32_t)y * NX + x;
        const Row R = row_of(P, y);
        const Nbr b = make_nbr(x, y, 0, NX, NY, 1);
        float dTlag = 0.f;
        float g[Q], f[Q];
        if constexpr (THERMAL) {
            dTlag = tl_in[n];
            gather_g32_o32<NTL>(P, R, b, gin, n, dTlag, g);
        }
        if constexpr (FLUID) gather_f32_o32<NTL>(P, R, b, fin, n, f);
        A dT = A(0);
        if constexpr (THERMAL) {
            if constexpr (sizeof(A) == 4) {
                dT = sum19(g);
            } else {
                for (int q = 0; q < Q; q++) dT += (A)g[q];
            }
        }
        A ux = 0, uy = 0, uz = 0, tau_les = A(0);
        const TileOut out{t, m, M, GOFS};
        if constexpr (FLUID) collide_fluid<A, false>(P, R, x, y, 0, n, f, dTlag, dT, out, ux, uy, uz, tau_les);
        if constexpr (THERMAL) collide_thermal<A, false>(P, R, y, g, dT, ux, uy, uz, tau_les, out);
    }
    __syncthreads();
    // second step: the tile, from LDS into the global state
    const int tx = threadIdx.x % T2_TX, ty = threadIdx.x / T2_TX;
    const int x = x0 + tx, y = y0 + ty;
    if (ty >= TY || y >= NY || tx >= w) return;
    const int at = (ty + 1) * T2_W + tx + 1;
    const uint32_t n = (uint32_t)y * NX + x;
    const Row R = row_of(P, y);
    const Nbr b = make_nbr(x, y, 0, NX, NY, 1);
    float dTlag = 0.f;
    float g[Q], f[Q];
    if constexpr (THERMAL) {
        dTlag = t[(GOFS + Q) * M + at];
        gather_g_tile<M>(P, R, b, t, GOFS, at, dTlag, g);
    }
    if constexpr (FLUID) gather_f_tile<M>(P, R, b, t, at, f);
    A dT = A(0);
    if constexpr (THERMAL) {
        if constexpr (sizeof(A) == 4) {
            dT = sum19(g);
        } else {
            for (int q = 0; q < Q; q++) dT += (A)g[q];
        }
    }
    A ux = 0, uy = 0, uz = 0, tau_les = A(0);
    const NodeOut32 out{fout, gout, tl_out, n * 4u, (uint32_t)P.S * 4u};
    if constexpr (FLUID) collide_fluid<A, false>(P, R, x, y, 0, n, f, dTlag, dT, out, ux, uy, uz, tau_les);
    if constexpr (THERMAL) collide_thermal<A, false>(P, R, y, g, dT, ux, uy, uz, tau_les, out);
}

template <bool FLUID, bool THERMAL, typename A, bool NTL, int TY>
__global__ void __launch_bounds__(Tile2<TY>::THREADS) k_step2_fp32(KParams P, const float *__restrict__ fin,
                                                                   const float *__restrict__ gin,
                                                                   float *__restrict__ fout, float *__restrict__ gout,
                                                                   const float *__restrict__ tl_in,
                                                                   float *__restrict__ tl_out)
{
    __shared__ float t[tile2_planes<FLUID, THERMAL>() * Tile2<TY>::M];
    step2_tile<FLUID, THERMAL, A, NTL, TY>(P, fin, gin, fout, gout, tl_in, tl_out, blockIdx.x * T2_TX,
                                           blockIdx.y * TY, t);
}

// ---------------------------------------------------------------------------
// Two steps per launch for 3-D lattices (NZ > 1), marching along z.  A block
// owns a 64 x TY column of x-y tiles over a range of z planes and has one
// thread per node of the tile with its one-node x-y halo.  Iteration j:
//   A  step 1 at plane j for the tile + halo nodes (gather from the global
//      state, collide into registers); the populations with c_z = -1 go to
//      the LDS ring M;
//   B  step 2 at plane j - 1 for the tile nodes, gathering from LDS: c_z = +1
//      from plane j - 2 (ring P), c_z = 0 from plane j - 1 (ring Z), c_z = -1
//      from plane j (ring M); collide into the global output;
//   C  the registers of step A go to ring Z (c_z = 0 and the step-1
//      temperature lag) and ring P (c_z = +1, two slots by plane parity);
//      the tile nodes also keep their populations 16 and 18 (c_z = -1) in
//      ring W, which the wall rules of step B read at the node itself one
//      plane later (fluid bounce-back at y = 0 takes f_16, f_18 there, the
//      fixed-T rule at y = NY - 1 takes g_16, g_18).
// Planes just below and above the block's range are computed in step A
// (z-halo) but not written.  Per two steps the state is read (1 + 2/ZC)
// times over the tile + halo area (x-y halo: 396 / 256 nodes for TY = 4) and
// written once; the temperature lag is read from tl_in and written to tl_out
// (halo nodes of one block are tile nodes of another).
namespace z2 {
// index of population q within its c_z class: c_z = 0: {0,1,2,3,4,7,8,9,10},
// c_z = +1: {5,11,12,15,17}, c_z = -1: {6,13,14,16,18}
__host__ __device__ constexpr int slot(int q)
{
    constexpr int v[Q] = {0, 1, 2, 3, 4, 0, 0, 5, 6, 7, 8, 1, 2, 1, 2, 3, 3, 4, 4};
    return v[q];
}
// ring W index of population 16 / 18
__host__ __device__ constexpr int wslot(int q) { return q == 16 ? 0 : 1; }
}  // namespace z2

// LDS rings of a z-marching tile of TX x TY nodes (one-node x-y halo).
// HOLD: step A keeps its c_z >= 0 results in registers until step C (one
// slot of ring Z, two of ring P, one of ring W); otherwise step A writes
// every result at once into rings with one more slot each (two of Z, three
// of P, two of W).
template <bool FLUID, bool THERMAL, int TX_, int TY_, bool HOLD_>
struct TileZ {
    static constexpr int TX = TX_, TY = TY_, W = TX + 2;
    static constexpr bool HOLD = HOLD_;
    static constexpr int M = W * (TY + 2);            // tile + halo nodes
    static constexpr int T = TX * TY;                 // tile nodes
    static constexpr int THREADS = (M + 63) / 64 * 64;
    static constexpr int F = FLUID ? 1 : 0, H = THERMAL ? 1 : 0;
    static constexpr int KM = 5 * F + 5 * H;          // ring M planes
    static constexpr int KZ = 9 * F + 9 * H + H;      // planes of one ring Z slot (+ lag)
    static constexpr int KP = 5 * F + 5 * H;          // planes of one ring P slot
    static constexpr int KW = 2 * F + 2 * H;          // planes of one ring W slot (tile nodes)
    static constexpr int NZS = HOLD ? 1 : 2, NPS = HOLD ? 2 : 3, NWS = HOLD ? 1 : 2;
    static constexpr int OZ = KM, OP = KM + NZS * KZ, OW = (KM + NZS * KZ + NPS * KP) * M;
    static constexpr int FLOATS = OW + NWS * KW * T;
    static_assert(THREADS >= T && THREADS <= 1024, "z-marching block");
    // plane of population q (fluid: gq = 0, temperature: gq = 1) in its ring
    static __device__ constexpr int pm(int q, int gq) { return gq * 5 * F + z2::slot(q); }
    static __device__ constexpr int pz(int q, int gq, int s) { return OZ + s * KZ + gq * 9 * F + z2::slot(q); }
    static __device__ constexpr int lag(int s) { return OZ + s * KZ + 9 * F + 9 * H; }
    static __device__ constexpr int pp(int q, int gq, int s) { return OP + s * KP + gq * 5 * F + z2::slot(q); }
    static __device__ constexpr int pw(int q, int gq, int s) { return s * KW + gq * 2 * F + z2::wslot(q); }
};

// step-1 output of one node, kept in registers until it goes to the rings
struct RegOut {
    float *f, *g, *tl;
    __device__ void put_f(int q, float v) const { f[q] = v; }
    __device__ void put_g(int q, float v) const { g[q] = v; }
    __device__ void put_tlag(float v) const { *tl = v; }
    __device__ float get_tau() const { return 0.f; }  // no LES on the two-step path
    __device__ void put_tau(float) const {}
};

// Ring slots read by step B at iteration it (plane p = it - 1 of the
// iteration count): ring Z and W of plane p, ring P of plane p and p - 1.
template <class TZ>
struct RingSlots {
    int z, w, p_own, p_below;
    __device__ RingSlots(int it)
    {
        if constexpr (TZ::HOLD) {
            z = 0;
            w = 0;
            p_own = (it - 1) & 1;
            p_below = it & 1;
        } else {
            z = (it - 1) & 1;
            w = (it - 1) & 1;
            p_own = (it - 1) % 3;
            p_below = (it + 1) % 3;  // (it - 2) mod 3
        }
    }
};

// Step-2 gather of population q (gq = 0 fluid, 1 temperature) at tile node
// `at` (tile + halo index) / `tn` (tile index) from the rings.
template <class TZ, int q, int gq>
__device__ __forceinline__ float ring_src(const float *t, int at, int tn, const RingSlots<TZ> &rs, bool cross)
{
    constexpr int M = TZ::M;
    float v;
    // pull source in the streaming direction
    const int src = at - cxq(q) - cyq(q) * TZ::W;
    if constexpr (czq(q) > 0)
        v = t[TZ::pp(q, gq, rs.p_below) * M + src];
    else if constexpr (czq(q) < 0)
        v = t[TZ::pm(q, gq) * M + src];
    else
        v = t[TZ::pz(q, gq, rs.z) * M + src];
    if constexpr (cyq(q) != 0) {
        // wall rules: fluid bounce-back takes the opposite population at the
        // node itself, the fixed-T rule the same population there
        constexpr int qs = gq == 0 ? invq(q) : q;
        if (cross) {
            if constexpr (czq(qs) > 0)
                v = t[TZ::pp(qs, gq, rs.p_own) * M + at];
            else if constexpr (czq(qs) < 0)
                v = t[TZ::OW + TZ::pw(qs, gq, rs.w) * TZ::T + tn];
            else
                v = t[TZ::pz(qs, gq, rs.z) * M + at];
        }
    }
    return v;
}

// Step A's loads of one node (load_raw_buf) and the row and wall terms
// applied to them afterwards (finish_raw): together the same values as
// gather_f32 / gather_g32.
struct RawNode {
    float f[Q], g[Q], tl;
};

// The Q planes of a population array as three buffer resources of up to
// seven planes each: a plane is addressed by the resource of its group, a
// scalar offset of its place in the group and the node's 32-bit byte offset
// (planes of less than 4 GiB / 7).  Keeps the scalar registers free of 38
// 64-bit plane addresses and the vector registers of 64-bit lane addresses.
struct Planes {
    __amdgpu_buffer_rsrc_t r[3];
    uint32_t pb;  // plane stride in bytes
};

__device__ __forceinline__ Planes make_planes(const float *base, int64_t S)
{
    Planes p;
    p.pb = (uint32_t)(S * 4);
    static_for<0, 3>([&](auto gc) {
        constexpr int gi = decltype(gc)::value;
        constexpr int np = gi < 2 ? 7 : Q - 14;
        p.r[gi] = __builtin_amdgcn_make_buffer_rsrc((void *)(base + (size_t)gi * 7 * (size_t)S), (short)0,
                                                    (int)(uint32_t)((uint64_t)np * p.pb), 0x00020000);
    });
    return p;
}

// cache-policy bits of buffer instructions: non-temporal
constexpr int CPOL_NT = 2;

template <int q, bool NT>
__device__ __forceinline__ float bload(const Planes &p, uint32_t soff, uint32_t voff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(p.r[q / 7], (int)voff, (int)soff,
                                                                          NT ? CPOL_NT : 0));
}

template <bool FLUID, bool THERMAL, bool NTL>
__device__ __forceinline__ void load_raw_buf(const KParams &P, const Nbr &b, const Planes &pf, const Planes &pg,
                                             const float *__restrict__ tl_in, uint32_t n, RawNode &r)
{
    const uint32_t n4 = n * 4u;
    if constexpr (THERMAL) {
        r.tl = tl_in[n];
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const bool cross = P.wall_t && crosses_wall<q>(b);
            r.g[q] = bload<q, NTL>(pg, (uint32_t)(q % 7) * pg.pb, cross ? n4 : src_index<q>(b, P.NX, P.NY) * 4u);
        });
    }
    if constexpr (FLUID) {
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (cyq(q) == 0) {
                r.f[q] = bload<q, NTL>(pf, (uint32_t)(q % 7) * pf.pb, src_index<q>(b, P.NX, P.NY) * 4u);
            } else {
                // bounce-back reads the opposite plane (same group of seven)
                constexpr int iq = q % 7, ii = invq(q) % 7, mn = iq < ii ? iq : ii;
                static_assert(q / 7 == invq(q) / 7, "opposite planes share a resource");
                const bool cross = P.wall_f && crosses_wall<q>(b);
                const uint32_t voff = cross ? n4 + (uint32_t)(ii - mn) * pf.pb
                                            : src_index<q>(b, P.NX, P.NY) * 4u + (uint32_t)(iq - mn) * pf.pb;
                r.f[q] = bload<q, NTL>(pf, (uint32_t)mn * pf.pb, voff);
            }
        });
    }
}

// NodeOut through buffer resources (non-temporal stores)
struct BufOut {
    Planes f, g;
    float *__restrict__ tlag;
    uint32_t n, n4;
    template <int q>
    __device__ void st(const Planes &p, float v) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), p.r[q / 7], (int)n4,
                                              (int)((uint32_t)(q % 7) * p.pb), CPOL_NT);
    }
    __device__ void put_f(int q, float v) const
    {
        static_for<0, Q>([&](auto qc) {
            constexpr int qq = decltype(qc)::value;
            if (qq == q) st<qq>(f, v);
        });
    }
    __device__ void put_g(int q, float v) const
    {
        static_for<0, Q>([&](auto qc) {
            constexpr int qq = decltype(qc)::value;
            if (qq == q) st<qq>(g, v);
        });
    }
    __device__ void put_tlag(float v) const { tlag[n] = v; }
    __device__ float get_tau() const { return 0.f; }  // no LES on the two-step path
    __device__ void put_tau(float) const {}
};

template <bool FLUID, bool THERMAL>
__device__ __forceinline__ void finish_raw(const KParams &P, const Row &R, const Nbr &b, RawNode &r)
{
    if constexpr (THERMAL) {
        const float fac_top = 2.f * (P.cw_top_f - r.tl);
        const float fac_bot = 2.f * (P.cw_bot_f - r.tl);
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const bool cross = P.wall_t && crosses_wall<q>(b);
            if constexpr (cyq(q) > 0) r.g[q] = fmaf(wf(q), cross ? fac_bot : R.dt_up, r.g[q]);
            if constexpr (cyq(q) < 0) r.g[q] = fmaf(wf(q), cross ? fac_top : R.dt_dn, r.g[q]);
        });
    }
    if constexpr (FLUID) {
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const bool cross = P.wall_f && crosses_wall<q>(b);
            if constexpr (cyq(q) > 0) r.f[q] = fmaf(wf(q), cross ? 0.f : R.dr_up, r.f[q]);
            if constexpr (cyq(q) < 0) r.f[q] = fmaf(wf(q), cross ? 0.f : R.dr_dn, r.f[q]);
        });
    }
}

// Step A / C: the results of plane j (iteration it) of node m into the
// rings: c_z = -1 into ring M (always), the others when `rest`.
template <class TZ, bool FLUID, bool THERMAL>
__device__ __forceinline__ void ring_put(float *t, int m, int it, const float rf[Q], const float rg[Q], float rt,
                                         bool tile, int tn, bool minus, bool rest)
{
    constexpr int M = TZ::M;
    const int sz = TZ::HOLD ? 0 : (it & 1), sw = sz;
    const int sp = TZ::HOLD ? (it & 1) : (it % 3);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (czq(q) < 0) {
            if (minus) {
                if constexpr (FLUID) t[TZ::pm(q, 0) * M + m] = rf[q];
                if constexpr (THERMAL) t[TZ::pm(q, 1) * M + m] = rg[q];
            }
        } else if constexpr (czq(q) == 0) {
            if (rest) {
                if constexpr (FLUID) t[TZ::pz(q, 0, sz) * M + m] = rf[q];
                if constexpr (THERMAL) t[TZ::pz(q, 1, sz) * M + m] = rg[q];
            }
        } else {
            if (rest) {
                if constexpr (FLUID) t[TZ::pp(q, 0, sp) * M + m] = rf[q];
                if constexpr (THERMAL) t[TZ::pp(q, 1, sp) * M + m] = rg[q];
            }
        }
    });
    if (rest) {
        if constexpr (THERMAL) t[TZ::lag(sz) * M + m] = rt;
        if (tile) {
            if constexpr (FLUID) {
                t[TZ::OW + TZ::pw(16, 0, sw) * TZ::T + tn] = rf[16];
                t[TZ::OW + TZ::pw(18, 0, sw) * TZ::T + tn] = rf[18];
            }
            if constexpr (THERMAL) {
                t[TZ::OW + TZ::pw(16, 1, sw) * TZ::T + tn] = rg[16];
                t[TZ::OW + TZ::pw(18, 1, sw) * TZ::T + tn] = rg[18];
            }
        }
    }
}

// Two steps per launch for 3-D lattices (NZ > 1), marching along z.  A block
// owns a TX x TY column of x-y tiles over a range of zc planes and has one
// thread per node of the tile with its one-node x-y halo.  Iteration j:
//   A  step 1 at plane j for the tile + halo nodes (gather from the global
//      state, collision);
//   B  step 2 at plane j - 1 for the tile nodes, gathering from LDS: c_z = +1
//      from plane j - 2 (ring P), c_z = 0 from plane j - 1 (ring Z), c_z = -1
//      from plane j (ring M); collision into the global output.
// Ring Z also holds the step-1 temperature (the lag of step 2); ring W keeps
// the tile nodes' populations 16 and 18 (c_z = -1), which the wall rules of
// step B read at the node itself one plane later (fluid bounce-back at y = 0
// takes f_16, f_18 there... see ring_src).  With HOLD the c_z >= 0 results of
// step A wait in registers and go to the rings after step B (C); without it
// step A writes them at once into the slots step B does not read.  The planes
// just below and above the block's range are computed in step A (z halo) but
// not written.  Per two steps the state is read (1 + 2/zc) x M/T times and
// written once; the temperature lag is read from tl_in and written to tl_out
// (halo nodes of one block are tile nodes of another).
template <bool FLUID, bool THERMAL, typename A, bool NTL, class TZ>
__global__ void __launch_bounds__(TZ::THREADS) k_step2z_fp32(KParams P, const float *__restrict__ fin,
                                                             const float *__restrict__ gin, float *__restrict__ fout,
                                                             float *__restrict__ gout,
                                                             const float *__restrict__ tl_in,
                                                             float *__restrict__ tl_out, int zc, int remap)
{
    constexpr int M = TZ::M, W = TZ::W, TX = TZ::TX, TY = TZ::TY;
    __shared__ float t[TZ::FLOATS];
    const int NX = P.NX, NY = P.NY, NZ = P.NZ;
    // block -> (x tile, y tile, z range); with remap, the blocks that one XCD
    // receives (every 8th in dispatch order) cover consecutive tiles, so
    // neighbouring tiles share their halo rows in that XCD's L2
    const uint32_t nbx = gridDim.x, nby = gridDim.y, total = nbx * nby * gridDim.z;
    uint32_t bid = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    if (remap) bid = xcd_block(bid, total);
    const int x0 = (int)(bid % nbx) * TX, y0 = (int)((bid / nbx) % nby) * TY;
    const int zb = (int)(bid / (nbx * nby)) * zc, ze = min(zb + zc, NZ);
    const int w = NX - x0 < TX ? NX - x0 : TX;
    // step A node of this thread (tile + halo)
    const int m = threadIdx.x, lx = m % W, ly = m / W;
    const bool act_a = m < M && lx <= w + 1;
    int xa = x0 - 1 + lx;
    xa = xa < 0 ? xa + NX : (xa >= NX ? xa - NX : xa);
    int ya = (y0 - 1 + ly) % NY;
    ya = ya < 0 ? ya + NY : ya;
    const bool tile_a = act_a && lx >= 1 && lx <= w && ly >= 1 && ly <= TY;
    const int tn_a = (ly - 1) * TX + (lx - 1);
    // step B node of this thread (tile)
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int xb = x0 + tx, yb = y0 + ty;
    const bool act_b = ty < TY && tx < w && yb < NY;
    const int at = (ty + 1) * W + tx + 1, tn = ty * TX + tx;

    // plane of step A at iteration it (periodic in z) and its node
    auto node_a = [&](int it_, uint32_t &n, Nbr &b) {
        const int j = zb - 1 + it_;
        const int za = j < 0 ? j + NZ : (j >= NZ ? j - NZ : j);
        n = ((uint32_t)za * NY + ya) * NX + xa;
        b = make_nbr(xa, ya, za, NX, NY, NZ);
    };
    const int nit = ze - zb + 2;
    Planes pin_f, pin_g, pout_f, pout_g;
    if constexpr (FLUID) {
        pin_f = make_planes(fin, P.S);
        pout_f = make_planes(fout, P.S);
    }
    if constexpr (THERMAL) {
        pin_g = make_planes(gin, P.S);
        pout_g = make_planes(gout, P.S);
    }
    for (int it = 0; it < nit; it++) {
        const int j = zb - 1 + it;
        float rf[Q], rg[Q], rt = 0.f;
        // ---- A: step 1 at plane j
        if (act_a) {
            uint32_t n;
            Nbr b;
            node_a(it, n, b);
            const Row R = row_of(P, ya);
            RawNode cur;
            load_raw_buf<FLUID, THERMAL, NTL>(P, b, pin_f, pin_g, tl_in, n, cur);
            finish_raw<FLUID, THERMAL>(P, R, b, cur);
            const float dTlag = THERMAL ? cur.tl : 0.f;
            A dT = A(0);
            if constexpr (THERMAL) {
                if constexpr (sizeof(A) == 4) {
                    dT = sum19(cur.g);
                } else {
                    for (int q = 0; q < Q; q++) dT += (A)cur.g[q];
                }
            }
            A ux = 0, uy = 0, uz = 0, tau_les = A(0);
            const RegOut out{rf, rg, &rt};
            const int za = b.z;
            if constexpr (FLUID)
                collide_fluid<A, false>(P, R, xa, ya, za, n, cur.f, dTlag, dT, out, ux, uy, uz, tau_les);
            if constexpr (THERMAL) collide_thermal<A, false>(P, R, ya, cur.g, dT, ux, uy, uz, tau_les, out);
            ring_put<TZ, FLUID, THERMAL>(t, m, it, rf, rg, rt, tile_a, tn_a, true, !TZ::HOLD);
        }
        __syncthreads();
        // ---- B: step 2 at plane j - 1 (planes of the block's range only)
        const int p = j - 1;
        if (act_b && p >= zb) {
            const RingSlots<TZ> rs(it);
            const uint32_t n = ((uint32_t)p * NY + yb) * NX + xb;
            const Row R = row_of(P, yb);
            const Nbr b = make_nbr(xb, yb, p, NX, NY, NZ);
            float dTlag = 0.f;
            if constexpr (THERMAL) dTlag = t[TZ::lag(rs.z) * M + at];
            // temperature populations from the rings (read twice: for the
            // temperature the fluid's force needs, then for their collision,
            // so they do not occupy registers during the fluid's)
            auto gather_g = [&](float g[Q]) {
                const float fac_top = 2.f * (P.cw_top_f - dTlag);
                const float fac_bot = 2.f * (P.cw_bot_f - dTlag);
                static_for<0, Q>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const bool cross = P.wall_t && crosses_wall<q>(b);
                    float v = ring_src<TZ, q, 1>(t, at, tn, rs, cross);
                    if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? fac_bot : R.dt_up, v);
                    if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? fac_top : R.dt_dn, v);
                    g[q] = v;
                });
            };
            A dT = A(0);
            if constexpr (THERMAL) {
                float g[Q];
                gather_g(g);
                if constexpr (sizeof(A) == 4) {
                    dT = sum19(g);
                } else {
                    for (int q = 0; q < Q; q++) dT += (A)g[q];
                }
            }
            A ux = 0, uy = 0, uz = 0, tau_les = A(0);
            auto collide2 = [&](const auto &out) {
            if constexpr (FLUID) {
                float f[Q];
                static_for<0, Q>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const bool cross = P.wall_f && crosses_wall<q>(b);
                    float v = ring_src<TZ, q, 0>(t, at, tn, rs, cross);
                    if constexpr (cyq(q) > 0) v = fmaf(wf(q), cross ? 0.f : R.dr_up, v);
                    if constexpr (cyq(q) < 0) v = fmaf(wf(q), cross ? 0.f : R.dr_dn, v);
                    f[q] = v;
                });
                collide_fluid<A, false>(P, R, xb, yb, p, n, f, dTlag, dT, out, ux, uy, uz, tau_les);
            }
            if constexpr (THERMAL) {
                float g[Q];
                gather_g(g);
                collide_thermal<A, false>(P, R, yb, g, dT, ux, uy, uz, tau_les, out);
            }
            };
            collide2(BufOut{pout_f, pout_g, tl_out, n, n * 4u});
        }
        __syncthreads();
        // ---- C (HOLD): the c_z >= 0 results of plane j into rings Z, P, W
        if constexpr (TZ::HOLD) {
            if (act_a) ring_put<TZ, FLUID, THERMAL>(t, m, it, rf, rg, rt, tile_a, tn_a, false, true);
        }
    }
}

// Many 2-D lattices in one launch, two steps each: block b advances one
// 64 x 4 tile of the lattice whose block range in this launch holds it
// (block-uniform binary search over the launch's few slots).  A slot
// flagged single advances its tiles by one step instead (the odd remainder
// of the lattice's share of a chunk; the lag stays in place).  Population
// buffers alternate with every launch of a lattice, its lag arrays with
// every two-step launch (parities in the slot).
template <bool FLUID, bool THERMAL, typename A, bool NTL>
__global__ void __launch_bounds__(Tile2<4>::THREADS) k_step2_batch32(const BatchEntry *__restrict__ tab,
                                                                     const BatchSlot *__restrict__ slots, int nslots)
{
    constexpr int TY = 4;
    __shared__ float t[tile2_planes<FLUID, THERMAL>() * Tile2<TY>::M];
    const uint32_t bt = blockIdx.x;
    if (bt >= slots[nslots].tile0) return;
    int lo = 0, hi = nslots - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slots[mid].tile0 <= bt)
            lo = mid;
        else
            hi = mid - 1;
    }
    const BatchSlot sl = slots[lo];
    const BatchEntry &E = tab[sl.entry & BATCH_ENTRY_MASK];
    const KParams &P = E.P;
    const int in = E.cur ^ (int)(sl.entry >> 31);
    const bool single = (sl.entry & BATCH_SINGLE) != 0;
    const bool tp = (sl.entry & BATCH_LAG_PARITY) != 0;
    const float *fin = (const float *)E.f[in], *gin = (const float *)E.g[in];
    float *fout = (float *)E.f[in ^ 1], *gout = (float *)E.g[in ^ 1];
    float *tl_a = (float *)(tp ? E.tlag2 : E.tlag), *tl_b = (float *)(tp ? E.tlag : E.tlag2);
    const uint32_t tile = bt - sl.tile0, txs = ((uint32_t)P.NX + T2_TX - 1) / T2_TX;
    const int x0 = (int)(tile % txs) * T2_TX, y0 = (int)(tile / txs) * TY;
    if (single) {
        const int x = x0 + (int)threadIdx.x % T2_TX, y = y0 + (int)threadIdx.x / T2_TX;
        if ((int)threadIdx.x >= T2_TX * TY || x >= P.NX || y >= P.NY) return;
        step_node32<false, FLUID, THERMAL, A, NTL, false>(P, fin, gin, fout, gout, tl_a, (float *)E.taules, x, y, 0,
                                                           (uint32_t)y * (uint32_t)P.NX + (uint32_t)x);
        return;
    }
    step2_tile<FLUID, THERMAL, A, NTL, TY>(P, fin, gin, fout, gout, tl_a, tl_b, x0, y0, t);
}

// Velocity of the step at every node (the u[] that add_collision reads,
// fluid.c:1056-1080), for the finite-difference LES closure.
template <bool FIRST, typename A>
__global__ void __launch_bounds__(256) k_velocity32(KParams P, const float *__restrict__ fin,
                                                    const float *__restrict__ tlag)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= P.NX) return;
    const uint32_t n = ((uint32_t)z * P.NY + y) * P.NX + x;
    const size_t N = (size_t)P.N;
    const Row R = row_of(P, y);
    const Nbr b = make_nbr(x, y, z, P.NX, P.NY, P.NZ);
    float f[Q];
    gather_f32<FIRST, false>(P, R, b, fin, n, f);
    const float dTlag = P.thermal ? tlag[n] : 0.f;
    Moments<A> m;
    fluid_moments<A, true>(P, R, x, y, z, n, f, dTlag, m);
    float *vel = (float *)P.vel;
    vel[n] = (float)m.ux;
    vel[N + n] = (float)m.uy;
    vel[2 * N + n] = (float)m.uz;
}

// Many lattices in one launch: each wave advances one wave tile (up to 64
// consecutive x of one row) of the lattice whose tile range in this launch
// holds it (wave-uniform binary search over the launch's few slots), with
// that lattice's step parity; the state buffer of each lattice alternates
// with the step parity.  One instance per population-set kind.
template <bool FLUID, bool THERMAL, typename A, bool NTL, bool DENSE, bool GEN>
__global__ void __launch_bounds__(256) k_step_batch32(const BatchEntry *__restrict__ tab,
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
    const float *fin = (const float *)E.f[in], *gin = (const float *)E.g[in];
    float *fout = (float *)E.f[in ^ 1], *gout = (float *)E.g[in ^ 1];
    float *tl = (float *)E.tlag, *tau = (float *)E.taules;
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
    step_node32<false, FLUID, THERMAL, A, NTL, GEN>(P, fin, gin, fout, gout, tl, tau, (int)x, (int)y, (int)z, n);
}

__global__ void __launch_bounds__(256) k_gather_fp32(KParams P, const float *__restrict__ fin,
                                                     const float *__restrict__ gin, const float *__restrict__ tlag,
                                                     double *__restrict__ fpost, double *__restrict__ gpost,
                                                     double *__restrict__ dens, double *__restrict__ u,
                                                     double *__restrict__ t, float4 *__restrict__ uT)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P.NX) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const uint32_t n = ((uint32_t)z * P.NY + y) * P.NX + x;
    const size_t N = (size_t)P.N;
    const Row R = row_of(P, y);
    const Nbr b = make_nbr(x, y, z, P.NX, P.NY, P.NZ);
    const double rb = P.row64[2 * y], tb = P.row64[2 * y + 1];
    float dTlag = 0.f;
    float3 fu = make_float3(0.f, 0.f, 0.f);  // the float fields, for uT
    float ft = 0.f;
    if (P.thermal) {
        float g[Q];
        dTlag = tlag[n];
        if (P.first)
            gather_g32<true, false>(P, R, b, gin, n, dTlag, g);
        else
            gather_g32<false, false>(P, R, b, gin, n, dTlag, g);
        double s = 0.0;
        for (int q = 0; q < Q; q++) {
            s += (double)g[q];
            if (gpost) gpost[q * N + n] = wq(q) * tb + (double)g[q];
        }
        if (t) t[n] = tb + s;
        ft = (float)(tb + s);
    }
    if (P.fluid) {
        float f[Q];
        if (P.first)
            gather_f32<true, false>(P, R, b, fin, n, f);
        else
            gather_f32<false, false>(P, R, b, fin, n, f);
        double drho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
        for (int q = 0; q < Q; q++) {
            double d = (double)f[q];
            drho += d;
            jx += cxq(q) * d;
            jy += cyq(q) * d;
            jz += czq(q) * d;
            if (fpost) fpost[q * N + n] = wq(q) * rb + d;
        }
        double rho = rb + drho;
        double ux = jx / rho, uy = jy / rho, uz = jz / rho;
        if (P.forced) {
            float F[3];
            force_at32<true>(P, P.hit_now, P.fbf_now, P.f0n, n, x, y, z, R.offT + dTlag, F);
            ux += 0.5 * (double)F[0];
            uy += 0.5 * (double)F[1];
            uz += 0.5 * (double)F[2];
        }
        if (dens) dens[n] = rho;
        if (u) {
            u[3 * n] = ux;
            u[3 * n + 1] = uy;
            u[3 * n + 2] = uz;
        }
        fu = make_float3((float)ux, (float)uy, (float)uz);
    }
    if (uT) uT[n] = make_float4(fu.x, fu.y, fu.z, ft);
}

// x extent of a row-grid block: 256 threads, halved for short rows
static int block_x32(const KParams &P)
{
    int bx = 256;
    while (bx > 64 && bx / 2 >= P.NX) bx /= 2;
    return bx;
}

void launch_step_fp32(const KParams &P, const float *fin, const float *gin, float *fout, float *gout, float *tlag,
                      float *taules, hipStream_t s)
{
    int bx = block_x32(P);
    const bool dense = !P.first && dense_rows(P.NX);
    dim3 grid((P.NX + bx - 1) / bx, P.NY, P.NZ), block(bx);
    if (dense) {
        grid = dim3((unsigned)((P.N + 255) / 256));
        block = dim3(256);
    }
    // both lattices resident in the 256 MB last-level cache?
    const int64_t lattice_bytes = 2 * (int64_t)Q * P.N * 4 * ((P.fluid ? 1 : 0) + (P.thermal ? 1 : 0));
    const bool ntl = lattice_bytes <= ((int64_t)192 << 20);
    const bool gen = P.hit || P.les || P.twoway || P.part || P.f0p || P.f0n;
#define CH4G_LAUNCH32G_(F_, FL_, TH_, A_, N_, D_)                                                                     \
    do {                                                                                                              \
        if (gen)                                                                                                      \
            hipLaunchKernelGGL((k_step_fp32<F_, FL_, TH_, A_, N_, D_, true>), grid, block, 0, s, P, fin, gin, fout,  \
                               gout, tlag, taules);                                                                   \
        else                                                                                                          \
            hipLaunchKernelGGL((k_step_fp32<F_, FL_, TH_, A_, N_, D_, false>), grid, block, 0, s, P, fin, gin, fout, \
                               gout, tlag, taules);                                                                   \
    } while (0)
#define CH4G_LAUNCH32D_(F_, FL_, TH_, A_, D_)                                                                          \
    do {                                                                                                              \
        if (ntl)                                                                                                      \
            CH4G_LAUNCH32G_(F_, FL_, TH_, A_, true, D_);                                                              \
        else                                                                                                          \
            CH4G_LAUNCH32G_(F_, FL_, TH_, A_, false, D_);                                                             \
    } while (0)
#define CH4G_LAUNCH32_(F_, FL_, TH_, A_)                                                                              \
    do {                                                                                                              \
        if (!(F_) && dense)                                                                                           \
            CH4G_LAUNCH32D_(false, FL_, TH_, A_, true);                                                               \
        else                                                                                                          \
            CH4G_LAUNCH32D_(F_, FL_, TH_, A_, false);                                                                 \
    } while (0)
#define CH4G_LAUNCH32(F_, FL_, TH_)                                                                                  \
    do {                                                                                                              \
        if (P.arith64)                                                                                                \
            CH4G_LAUNCH32_(F_, FL_, TH_, double);                                                                     \
        else                                                                                                          \
            CH4G_LAUNCH32_(F_, FL_, TH_, float);                                                                      \
    } while (0)
    if (P.fluid && P.thermal) {
        if (P.first) CH4G_LAUNCH32(true, true, true); else CH4G_LAUNCH32(false, true, true);
    } else if (P.fluid) {
        if (P.first) CH4G_LAUNCH32(true, true, false); else CH4G_LAUNCH32(false, true, false);
    } else {
        if (P.first) CH4G_LAUNCH32(true, false, true); else CH4G_LAUNCH32(false, false, true);
    }
#undef CH4G_LAUNCH32
#undef CH4G_LAUNCH32_
#undef CH4G_LAUNCH32D_
#undef CH4G_LAUNCH32G_
}

// Two steps per launch (k_step2_fp32) for 2-D lattices whose planes span
// less than 4 GiB, without the HIT / LES / two-way paths.  Measured on 2-D
// fluid+T lattices against one launch per step (us per step, bit-identical):
// 512x256 9.19 -> 7.97, 2048x1024 103.0 -> 88.7, 4096x2048 361.4 -> 325.6;
// with narrower last tiles 78x40 4.69 -> 3.85, 1000x500 38.6 -> 28.9,
// 1500x700 76.3 -> 48.2 (against the dense single-step grid).
//
// 3-D lattices (k_step2z_fp32) need enough tile columns to fill the GPU and
// rows that fill their 64-wide tiles; measured against single steps with
// graph replay (us per step): channel 64^3 8.4 vs 20.3, 96^3 28.6 vs 46.0,
// 128^3 62.2 vs 49.5, 160^3 (rows of 2.5 tiles) 116 vs 122, 192^3 196 vs 185;
// Rayleigh-Benard 96^3 57.1 vs 66.9, 128^3 126.9 vs 92.6, 192^3 406 vs 344.
bool two_step_ok(const KParams &P)
{
    // 2-D: 32-bit byte offsets into the Q population planes; 3-D: 32-bit
    // byte offsets within a group of seven planes
    const bool off32 = (int64_t)Q * P.S * 4 < ((int64_t)1 << 32);
    const int64_t tiles_x = (P.NX + 63) / 64;
    const bool big3d = P.N >= ((int64_t)1 << 21) && 10 * (int64_t)P.NX >= 9 * 64 * tiles_x;
    const bool fits = P.NZ == 1 ? off32 : (big3d && 7 * P.S * 4 < ((int64_t)1 << 32));
    return fits && !P.first && !P.hit && !P.les && !P.twoway && !P.nozero;
}

namespace {

// Two steps per launch for 3-D lattices (k_step2z_fp32), z ranges of zc
// planes.
template <class TZ>
void launch_z2(const KParams &P, const float *fin, const float *gin, float *fout, float *gout, const float *tl_in,
               float *tl_out, int zc, int remap, bool ntl, hipStream_t s)
{
    const unsigned bx = (unsigned)((P.NX + TZ::TX - 1) / TZ::TX), by = (unsigned)((P.NY + TZ::TY - 1) / TZ::TY);
    const dim3 grid(bx, by, (unsigned)((P.NZ + zc - 1) / zc));
#define CH4G_Z2N_(FL_, TH_, A_, N_)                                                                                   \
    hipLaunchKernelGGL((k_step2z_fp32<FL_, TH_, A_, N_, TZ>), grid, dim3(TZ::THREADS), 0, s, P, fin, gin, fout,     \
                       gout, tl_in, tl_out, zc, remap)
#define CH4G_Z2_(FL_, TH_, A_)                                                                                        \
    do {                                                                                                              \
        if (ntl)                                                                                                      \
            CH4G_Z2N_(FL_, TH_, A_, true);                                                                            \
        else                                                                                                          \
            CH4G_Z2N_(FL_, TH_, A_, false);                                                                           \
    } while (0)
    constexpr bool FL = TZ::F == 1, TH = TZ::H == 1;
    if (P.arith64)
        CH4G_Z2_(FL, TH, double);
    else
        CH4G_Z2_(FL, TH, float);
#undef CH4G_Z2_
#undef CH4G_Z2N_
}

template <bool FL, bool TH, int TX, int TY, bool HOLD>
using TZ_ = TileZ<FL, TH, TX, TY, HOLD>;

// Tiles (measured at 512^3 fluid + T, FP32-deviation, ms per step, single
// steps 7.9): 64 x 4 held in registers 6.9, with the next plane's loads in
// flight during step B 6.5, 32 x 5 all in LDS (two blocks per CU) 7.3,
// 64 x 8 held 6.0, the same through buffer resources 5.6-5.9; 3-D channel
// 256^3: single 0.46, 64 x 8 held 0.35.  Double arithmetic needs more
// registers than an 11-wave block allows: 64 x 4.  z ranges: at least 1024
// blocks (four per CU) and at most 128 planes (channel 192^3 with ranges of
// 128 planes: 144 blocks, 0.30 ms; of 32 planes: 0.18 ms).
void launch_two_steps_3d(const KParams &P, const float *fin, const float *gin, float *fout, float *gout,
                         const float *tl_in, float *tl_out, hipStream_t s)
{
    const int64_t lattice_bytes = 2 * (int64_t)Q * P.N * 4 * ((P.fluid ? 1 : 0) + (P.thermal ? 1 : 0));
    const bool ntl = lattice_bytes <= ((int64_t)192 << 20);
    const bool tall = P.fluid && !(P.thermal && P.arith64);
    const int TY = tall ? 8 : 4;
    const int64_t columns = (int64_t)((P.NX + 63) / 64) * ((P.NY + TY - 1) / TY);
    int64_t zc = (int64_t)P.NZ * columns / 1024;
    zc = zc < 8 ? 8 : (zc > 128 ? 128 : zc);
    zc = zc > P.NZ ? P.NZ : zc;
    const int z = (int)zc, remap = 1;
#define Z2(...) launch_z2<TZ_<__VA_ARGS__>>(P, fin, gin, fout, gout, tl_in, tl_out, z, remap, ntl, s)
    if (P.fluid && P.thermal) {
        if (P.arith64)
            Z2(true, true, 64, 4, true);
        else
            Z2(true, true, 64, 8, true);
    } else if (P.fluid) {
        Z2(true, false, 64, 8, true);
    } else {
        Z2(false, true, 64, 4, false);
    }
#undef Z2
}

}  // namespace

void launch_two_steps_fp32(const KParams &P, const float *fin, const float *gin, float *fout, float *gout,
                           const float *tl_in, float *tl_out, hipStream_t s)
{
    if (P.NZ > 1) {
        launch_two_steps_3d(P, fin, gin, fout, gout, tl_in, tl_out, s);
        return;
    }
    // non-temporal loads while both lattices fit the last-level cache (as the
    // single-step kernel); 64x5 tiles (fewer halo nodes) once there are at
    // least two rounds of blocks (two blocks per CU by LDS), else 64x4 tiles
    // (more, smaller blocks fill the CUs evenly; 64x3 tiles, three blocks per
    // CU, were slower at 512x256: 9.6 against 7.3 us per step)
    const int64_t lattice_bytes = 2 * (int64_t)Q * P.N * 4 * ((P.fluid ? 1 : 0) + (P.thermal ? 1 : 0));
    const bool ntl = lattice_bytes <= ((int64_t)192 << 20);
    const unsigned bx = (unsigned)((P.NX + T2_TX - 1) / T2_TX);
    const bool ty5 = (int64_t)bx * ((P.NY + 4) / 5) >= 1024;
    const dim3 g4(bx, (unsigned)((P.NY + 3) / 4)), g5(bx, (unsigned)((P.NY + 4) / 5));
#define CH4G_T2N_(FL_, TH_, A_, N_)                                                                                   \
    do {                                                                                                              \
        if (ty5)                                                                                                      \
            hipLaunchKernelGGL((k_step2_fp32<FL_, TH_, A_, N_, 5>), g5, dim3(Tile2<5>::THREADS), 0, s, P, fin, gin,  \
                               fout, gout, tl_in, tl_out);                                                            \
        else                                                                                                          \
            hipLaunchKernelGGL((k_step2_fp32<FL_, TH_, A_, N_, 4>), g4, dim3(Tile2<4>::THREADS), 0, s, P, fin, gin,  \
                               fout, gout, tl_in, tl_out);                                                            \
    } while (0)
#define CH4G_T2_(FL_, TH_, A_)                                                                                        \
    do {                                                                                                              \
        if (ntl)                                                                                                      \
            CH4G_T2N_(FL_, TH_, A_, true);                                                                            \
        else                                                                                                          \
            CH4G_T2N_(FL_, TH_, A_, false);                                                                           \
    } while (0)
#define CH4G_T2(FL_, TH_)                                                                                             \
    do {                                                                                                              \
        if (P.arith64)                                                                                                \
            CH4G_T2_(FL_, TH_, double);                                                                               \
        else                                                                                                          \
            CH4G_T2_(FL_, TH_, float);                                                                                \
    } while (0)
    if (P.fluid && P.thermal)
        CH4G_T2(true, true);
    else if (P.fluid)
        CH4G_T2(true, false);
    else
        CH4G_T2(false, true);
#undef CH4G_T2
#undef CH4G_T2_
#undef CH4G_T2N_
}

void launch_velocity_fp32(const KParams &P, const float *fin, const float *tlag, hipStream_t s)
{
    int bx = block_x32(P);
    dim3 grid((P.NX + bx - 1) / bx, P.NY, P.NZ);
    if (P.arith64) {
        if (P.first)
            hipLaunchKernelGGL((k_velocity32<true, double>), grid, dim3(bx), 0, s, P, fin, tlag);
        else
            hipLaunchKernelGGL((k_velocity32<false, double>), grid, dim3(bx), 0, s, P, fin, tlag);
    } else {
        if (P.first)
            hipLaunchKernelGGL((k_velocity32<true, float>), grid, dim3(bx), 0, s, P, fin, tlag);
        else
            hipLaunchKernelGGL((k_velocity32<false, float>), grid, dim3(bx), 0, s, P, fin, tlag);
    }
}

void launch_batch_fp32(const BatchEntry *tab, const BatchSlot *slots, int nslots, int64_t nblocks, int kind,
                       bool dense, bool gen, int arith64, bool ntl, hipStream_t s)
{
    dim3 grid((unsigned)nblocks), block(256);
#define CH4G_BATCH32G_(FL_, TH_, A_, N_, D_)                                                                          \
    do {                                                                                                              \
        if (gen)                                                                                                      \
            hipLaunchKernelGGL((k_step_batch32<FL_, TH_, A_, N_, D_, true>), grid, block, 0, s, tab, slots, nslots); \
        else                                                                                                          \
            hipLaunchKernelGGL((k_step_batch32<FL_, TH_, A_, N_, D_, false>), grid, block, 0, s, tab, slots,         \
                               nslots);                                                                               \
    } while (0)
#define CH4G_BATCH32D_(FL_, TH_, A_, D_)                                                                              \
    do {                                                                                                              \
        if (ntl)                                                                                                      \
            CH4G_BATCH32G_(FL_, TH_, A_, true, D_);                                                                   \
        else                                                                                                          \
            CH4G_BATCH32G_(FL_, TH_, A_, false, D_);                                                                  \
    } while (0)
#define CH4G_BATCH32_(FL_, TH_, A_)                                                                                   \
    do {                                                                                                              \
        if (dense)                                                                                                    \
            CH4G_BATCH32D_(FL_, TH_, A_, true);                                                                       \
        else                                                                                                          \
            CH4G_BATCH32D_(FL_, TH_, A_, false);                                                                      \
    } while (0)
#define CH4G_BATCH32(FL_, TH_)                                                                                        \
    do {                                                                                                              \
        if (arith64)                                                                                                  \
            CH4G_BATCH32_(FL_, TH_, double);                                                                          \
        else                                                                                                          \
            CH4G_BATCH32_(FL_, TH_, float);                                                                           \
    } while (0)
    if (kind == BATCH_FLUID_THERMAL)
        CH4G_BATCH32(true, true);
    else if (kind == BATCH_FLUID)
        CH4G_BATCH32(true, false);
    else
        CH4G_BATCH32(false, true);
#undef CH4G_BATCH32
#undef CH4G_BATCH32_
#undef CH4G_BATCH32D_
#undef CH4G_BATCH32G_
}

void launch_batch2_fp32(const BatchEntry *tab, const BatchSlot *slots, int nslots, int64_t nblocks, int kind,
                        int arith64, bool ntl, hipStream_t s)
{
    dim3 grid((unsigned)nblocks), block(Tile2<4>::THREADS);
#define CH4G_B2_(FL_, TH_, A_)                                                                                        \
    do {                                                                                                              \
        if (ntl)                                                                                                      \
            hipLaunchKernelGGL((k_step2_batch32<FL_, TH_, A_, true>), grid, block, 0, s, tab, slots, nslots);         \
        else                                                                                                          \
            hipLaunchKernelGGL((k_step2_batch32<FL_, TH_, A_, false>), grid, block, 0, s, tab, slots, nslots);        \
    } while (0)
#define CH4G_B2(FL_, TH_)                                                                                             \
    do {                                                                                                              \
        if (arith64)                                                                                                  \
            CH4G_B2_(FL_, TH_, double);                                                                               \
        else                                                                                                          \
            CH4G_B2_(FL_, TH_, float);                                                                                \
    } while (0)
    if (kind == BATCH_FLUID_THERMAL)
        CH4G_B2(true, true);
    else if (kind == BATCH_FLUID)
        CH4G_B2(true, false);
    else
        CH4G_B2(false, true);
#undef CH4G_B2
#undef CH4G_B2_
}

// LB_FLUID_FORCING_NOZEROMODE on float storage (forcing.c:1123-1152): sums
// over all nodes of the float force the step kernel builds (no particle
// feedback) of the previous step, F(T_lag) with the previous HIT table, and
// of this step, F(T) with T gathered as the step kernel gathers it (STEP), or
// of the last step alone (the hydro-fields gather), in double; fixed grid and
// a tree per block, so the partial sums are the same on every run.  The
// means are formed by k_force_mean64.
namespace {
constexpr int FM32_THREADS = 256, FM32_BLOCKS = 512;

template <bool STEP, bool FIRST>
__global__ void __launch_bounds__(FM32_THREADS) k_force_sum32(KParams P, const float *__restrict__ gin,
                                                              const float *__restrict__ tlag, double *__restrict__ part)
{
    P.twoway = 0;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int64_t N = P.N, stride = (int64_t)gridDim.x * FM32_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * FM32_THREADS + threadIdx.x; i < N; i += stride) {
        const uint32_t n = (uint32_t)i;
        const int x = (int)(i % P.NX), y = (int)((i / P.NX) % P.NY), z = (int)(i / ((int64_t)P.NX * P.NY));
        const Row R = row_of(P, y);
        const float dTlag = P.thermal ? tlag[n] : 0.f;
        float F[3];
        if constexpr (STEP) {
            force_at32<true>(P, P.hit_prev, nullptr, nullptr, n, x, y, z, R.offT + dTlag, F);
            for (int c = 0; c < 3; c++) acc[c] += (double)F[c];
            float dT = 0.f;
            if (P.thermal && P.buoy) {
                float g[Q];
                gather_g32<FIRST, false>(P, R, make_nbr(x, y, z, P.NX, P.NY, P.NZ), gin, n, dTlag, g);
                if (P.arith64) {
                    double s = 0.0;
                    for (int q = 0; q < Q; q++) s += (double)g[q];
                    dT = (float)s;
                } else {
                    dT = sum19(g);
                }
            }
            force_at32<true>(P, P.hit_now, nullptr, nullptr, n, x, y, z, R.offT + dT, F);
        } else {
            force_at32<true>(P, P.hit_now, nullptr, nullptr, n, x, y, z, R.offT + dTlag, F);
        }
        for (int c = 0; c < 3; c++) acc[3 + c] += (double)F[c];
    }
    __shared__ double red[6][FM32_THREADS];
    for (int c = 0; c < 6; c++) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = FM32_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 6; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_force_mean32(int nblocks, double NX, double NY, double NZ, const double *__restrict__ part,
                               double *__restrict__ f0)
{
    const int c = threadIdx.x;
    if (c >= 6) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += part[(size_t)b * 6 + c];
    const double norm = 1.0 / (NX * NY * NZ);
    f0[c] = s * norm;
}
}  // namespace

void launch_force_mean32(const KParams &P, bool step, const float *gin, const float *tlag, double *part, double *f0,
                         hipStream_t s)
{
    const int64_t need = (P.N + FM32_THREADS - 1) / FM32_THREADS;
    const int nb = (int)(need < FM32_BLOCKS ? need : FM32_BLOCKS);
    if (!step)
        hipLaunchKernelGGL((k_force_sum32<false, false>), dim3(nb), dim3(FM32_THREADS), 0, s, P, gin, tlag, part);
    else if (P.first)
        hipLaunchKernelGGL((k_force_sum32<true, true>), dim3(nb), dim3(FM32_THREADS), 0, s, P, gin, tlag, part);
    else
        hipLaunchKernelGGL((k_force_sum32<true, false>), dim3(nb), dim3(FM32_THREADS), 0, s, P, gin, tlag, part);
    hipLaunchKernelGGL(k_force_mean32, dim3(1), dim3(64), 0, s, nb, (double)P.NX, (double)P.NY, (double)P.NZ, part,
                       f0);
}

void launch_gather_fp32(const KParams &P, const float *fin, const float *gin, const float *tlag, double *fpost,
                        double *gpost, double *dens, double *u, double *t, hipStream_t s)
{
    int bx = block_x32(P);
    dim3 grid((P.NX + bx - 1) / bx, P.NY, P.NZ);
    hipLaunchKernelGGL(k_gather_fp32, grid, dim3(bx), 0, s, P, fin, gin, tlag, fpost, gpost, dens, u, t, nullptr);
}

void launch_fields_f32(const KParams &P, const float *fin, const float *gin, const float *tlag, float4 *uT,
                       hipStream_t s)
{
    int bx = block_x32(P);
    dim3 grid((P.NX + bx - 1) / bx, P.NY, P.NZ);
    hipLaunchKernelGGL(k_gather_fp32, grid, dim3(bx), 0, s, P, fin, gin, tlag, nullptr, nullptr, nullptr, nullptr,
                       nullptr, uT);
}

}  // namespace ch4g
