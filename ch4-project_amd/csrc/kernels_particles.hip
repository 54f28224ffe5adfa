// Lagrangian point particles on the device: one thread per particle for the
// interpolation + advance + wall rules, an ordered re-injection pass that
// hands out drand48 draws in particle order by jumping the 48-bit LCG, and
// the two-way feedback as a sort of the particles by stencil cell followed
// by per-node sums in reference particle order.
//
// Compiled with -ffp-contract=off: every expression is evaluated in the
// reference's order (lagrange.c, lagrange_feedback.c, gradients.c), so
// positions, velocities and deposits are bit-identical to the reference's
// FP64 loops on the same fields.
#include <hipcub/hipcub.hpp>

#include "ch4g_internal.h"

namespace ch4g {
namespace {

constexpr int PB = 256;                         // particles per block
constexpr double FLT_EPS = 1.1920928955078125e-07;  // FLT_EPSILON (float.h) as a double
constexpr uint64_t MASK48 = (1ull << 48) - 1;

__device__ __forceinline__ bool lflag(const PartParams &P, uint32_t f) { return (P.flags & f) != 0; }

__device__ __forceinline__ int wrapi(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// wrap(x, s) = x - floor(x / s) * s (typedef.h:378)
__device__ __forceinline__ double wrapd(double x, double s) { return x - floor(x / s) * s; }

// drand48 state after k steps from x: X <- a X + c (mod 2^48) composed by
// squaring (all maps are powers of the same affine map, so they commute).
__device__ __forceinline__ uint64_t lcg_jump(uint64_t x, uint64_t k)
{
    uint64_t am = 0x5DEECE66Dull, ac = 0xBull, m = 1, c = 0;
    while (k) {
        if (k & 1) {
            m = m * am;
            c = c * am + ac;
        }
        ac = (am + 1) * ac;
        am = am * am;
        k >>= 1;
    }
    return (m * x + c) & MASK48;
}

__device__ __forceinline__ double lcg_value(uint64_t x) { return ldexp((double)x, -48); }

// Corners of the trilinear stencil in the order of the reference's sums
// (lagrange.c:1293-1319): (m,m,m) (p,m,m) (m,p,m) (m,m,p) (p,p,m) (m,p,p) (p,m,p) (p,p,p).
__device__ __forceinline__ int cdx(int c) { return (c == 1 || c == 4 || c == 6 || c == 7) ? 1 : 0; }
__device__ __forceinline__ int cdy(int c) { return (c == 2 || c == 4 || c == 5 || c == 7) ? 1 : 0; }
__device__ __forceinline__ int cdz(int c) { return (c == 3 || c == 5 || c == 6 || c == 7) ? 1 : 0; }

struct Stencil {
    int im, jm, km;   // logical cell of the (m,m,m) corner, -1 .. N-1
    double w[8];
};

// Cell lookup and volume coefficients (lagrange.c:1232-1292).  The
// reference scans the cell centres c + 1/2 for c + 1/2 <= x < c + 3/2;
// floor(x - 1/2) is that cell exactly (x - 1/2 is exact for x >= 1/4 and
// cannot cross an integer below it).
__device__ __forceinline__ bool stencil(const PartParams &P, double x, double y, double z, Stencil &st)
{
    const double px = wrapd(x, (double)P.NX), py = wrapd(y, (double)P.NY), pz = wrapd(z, (double)P.NZ);
    const double fx = floor(px - 0.5), fy = floor(py - 0.5), fz = floor(pz - 0.5);
    if (!(fx >= -1.0 && fx <= P.NX - 1.0 && fy >= -1.0 && fy <= P.NY - 1.0 && fz >= -1.0 && fz <= P.NZ - 1.0))
        return false;
    st.im = (int)fx;
    st.jm = (int)fy;
    st.km = (int)fz;
    const double dxm = px - ((double)st.im + 0.5), dxp = ((double)(st.im + 1) + 0.5) - px;
    const double dym = py - ((double)st.jm + 0.5), dyp = ((double)(st.jm + 1) + 0.5) - py;
    const double dzm = pz - ((double)st.km + 0.5), dzp = ((double)(st.km + 1) + 0.5) - pz;
    st.w[0] = dxp * dyp * dzp;
    st.w[1] = dxm * dyp * dzp;
    st.w[2] = dxp * dym * dzp;
    st.w[3] = dxp * dyp * dzm;
    st.w[4] = dxm * dym * dzp;
    st.w[5] = dxp * dym * dzm;
    st.w[6] = dxm * dyp * dzm;
    st.w[7] = dxm * dym * dzm;
    return true;
}

// Velocity at a logical node including the halo of boundary_conditions_hydro
// (boundary_conditions.c:2786-2875): periodic images, then u_g = -u on the
// first ghost plane of each no-slip y wall.
// the fields of the step at node n: double u [3N] and t [N] on FP64
// lattices; on float storage one 16-byte (u_x, u_y, u_z, T) per node
template <typename F>
__device__ __forceinline__ void field_u(const PartParams &P, size_t n, double v[3]);
template <>
__device__ __forceinline__ void field_u<double>(const PartParams &P, size_t n, double v[3])
{
    const double *u = P.u + 3 * n;
    v[0] = u[0];
    v[1] = u[1];
    v[2] = u[2];
}
template <>
__device__ __forceinline__ void field_u<float>(const PartParams &P, size_t n, double v[3])
{
    const float4 u = P.uT[n];
    v[0] = (double)u.x;
    v[1] = (double)u.y;
    v[2] = (double)u.z;
}
template <typename F>
__device__ __forceinline__ double field_t(const PartParams &P, size_t n);
template <>
__device__ __forceinline__ double field_t<double>(const PartParams &P, size_t n) { return P.t[n]; }
template <>
__device__ __forceinline__ double field_t<float>(const PartParams &P, size_t n) { return (double)P.uT[n].w; }

// node of (i, j, k) for u: periodic images, and on the first ghost plane of
// a no-slip y wall the first fluid row (ghost set: u_g = -u)
__device__ __forceinline__ size_t node_u(const PartParams &P, int i, int j, int k, bool &ghost)
{
    i = wrapi(i, P.NX);
    k = wrapi(k, P.NZ);
    ghost = false;
    if (P.wall_f && j == -1) {
        j = 0;
        ghost = true;
    } else if (P.wall_f && j == P.NY) {
        j = P.NY - 1;
        ghost = true;
    } else {
        j = wrapi(j, P.NY);
    }
    return ((size_t)k * P.NY + j) * P.NX + i;
}

// node of (i, j, k) for t, and the wall side of a ghost plane (-1 bottom, 1 top)
__device__ __forceinline__ size_t node_t(const PartParams &P, int i, int j, int k, int &side)
{
    i = wrapi(i, P.NX);
    k = wrapi(k, P.NZ);
    side = 0;
    if (P.wall_t && j == -1) {
        j = 0;
        side = -1;
    } else if (P.wall_t && j == P.NY) {
        j = P.NY - 1;
        side = 1;
    } else {
        j = wrapi(j, P.NY);
    }
    return ((size_t)k * P.NY + j) * P.NX + i;
}

// Temperature with the fixed-T ghost t_g = (2 (T_w - T_ref) / t - 1) t
// (boundary_conditions.c:2928-2945).
__device__ __forceinline__ double t_ghost(const PartParams &P, double t, int side)
{
    if (side == 0) return t;
    const double Tw = side < 0 ? P.Tw_bot : P.Tw_top;
    const double fac = 2.0 * Tw / t - 1.0;
    return fac * t;
}

template <typename F>
__device__ __forceinline__ void fetch_u(const PartParams &P, int i, int j, int k, double v[3])
{
    bool ghost;
    field_u<F>(P, node_u(P, i, j, k, ghost), v);
    if (ghost) {
        v[0] = -1.0 * v[0];
        v[1] = -1.0 * v[1];
        v[2] = -1.0 * v[2];
    }
}

template <typename F>
__device__ __forceinline__ double fetch_t(const PartParams &P, int i, int j, int k)
{
    int side;
    const double t = field_t<F>(P, node_t(P, i, j, k, side));
    return t_ghost(P, t, side);
}

// u and t of one node together, for lattices whose fluid and temperature
// walls coincide (wall_f == wall_t: both halo rules pick the same node): on
// float storage one 16-byte load
template <typename F>
__device__ __forceinline__ void fetch_ut(const PartParams &P, int i, int j, int k, double v[3], double &t)
{
    bool ghost;
    const size_t n = node_u(P, i, j, k, ghost);
    if constexpr (std::is_same<F, float>::value) {
        const float4 a = P.uT[n];
        v[0] = (double)a.x;
        v[1] = (double)a.y;
        v[2] = (double)a.z;
        t = (double)a.w;
    } else {
        field_u<F>(P, n, v);
        t = field_t<F>(P, n);
    }
    if (ghost) {
        v[0] = -1.0 * v[0];
        v[1] = -1.0 * v[1];
        v[2] = -1.0 * v[2];
    }
    t = t_ghost(P, t, ghost ? (j < 0 ? -1 : 1) : 0);
}

// second-order one-sided combinations with unit spacing: h1 = h2 = 1 gives
// a0 = -3/2, a1 = 2, a2 = -1/2 forward and the negatives backward
// (gradients.c:201-225, :372-486).
__device__ __forceinline__ void one_sided(const double a[3], const double b[3], const double c[3], bool fwd,
                                          double &X, double &Y, double &Z)
{
    const double a0 = fwd ? -1.5 : 1.5, a1 = fwd ? 2.0 : -2.0, a2 = fwd ? -0.5 : 0.5;
    X = a0 * a[0] + a1 * b[0] + a2 * c[0];
    Y = a0 * a[1] + a1 * b[1] + a2 * c[1];
    Z = a0 * a[2] + a1 * b[2] + a2 * c[2];
}

// gradient_vector with LAGRANGE on a two-node halo (gradients.c:78-499).
// G = xx, xy, xz, yx, yy, yz, zx, zy, zz (first letter: component of u,
// second: direction of the derivative).  With fluid walls (LB_FLUID_BC) the
// LAGRANGE_INTERPOLATION_FIX moves a ghost node's evaluation point to the
// first interior node (:288-368); without them it takes the one-sided
// difference at the ghost node itself (:413-488).
// the rows next to walls and the ghost corners of gradient_vector:
// one-sided differences, the interpolation fix, size-one axes zeroed
template <typename F>
__device__ void grad_u_edges(const PartParams &P, int i, int j, int k, double G[9])
{
    double a[3], b[3], c[3];
    if (P.wall_f) {
        // LB_FLUID_BC_Y: one-sided at the first and last fluid rows
        if (j == 0) {
            fetch_u<F>(P, i, 0, k, a);
            fetch_u<F>(P, i, 1, k, b);
            fetch_u<F>(P, i, 2, k, c);
            one_sided(a, b, c, true, G[1], G[4], G[7]);
        }
        if (j == P.NY - 1) {
            fetch_u<F>(P, i, P.NY - 1, k, a);
            fetch_u<F>(P, i, P.NY - 2, k, b);
            fetch_u<F>(P, i, P.NY - 3, k, c);
            one_sided(a, b, c, false, G[1], G[4], G[7]);
        }
        // LAGRANGE_INTERPOLATION_FIX: ghost nodes next to the interior take
        // the one-sided gradient of the first interior node; the moved index
        // is kept for the directions that follow
        if (i == -1) {
            i = 0;
            fetch_u<F>(P, 0, j, k, a);
            fetch_u<F>(P, 1, j, k, b);
            fetch_u<F>(P, 2, j, k, c);
            one_sided(a, b, c, true, G[0], G[3], G[6]);
        }
        if (i == P.NX) {
            i = P.NX - 1;
            fetch_u<F>(P, i, j, k, a);
            fetch_u<F>(P, i - 1, j, k, b);
            fetch_u<F>(P, i - 2, j, k, c);
            one_sided(a, b, c, false, G[0], G[3], G[6]);
        }
        if (j == -1) {
            j = 0;
            fetch_u<F>(P, i, 0, k, a);
            fetch_u<F>(P, i, 1, k, b);
            fetch_u<F>(P, i, 2, k, c);
            one_sided(a, b, c, true, G[1], G[4], G[7]);
        }
        if (j == P.NY) {
            j = P.NY - 1;
            fetch_u<F>(P, i, j, k, a);
            fetch_u<F>(P, i, j - 1, k, b);
            fetch_u<F>(P, i, j - 2, k, c);
            one_sided(a, b, c, false, G[1], G[4], G[7]);
        }
        if (k == -1) {
            k = 0;
            fetch_u<F>(P, i, j, 0, a);
            fetch_u<F>(P, i, j, 1, b);
            fetch_u<F>(P, i, j, 2, c);
            one_sided(a, b, c, true, G[2], G[5], G[8]);
        }
        if (k == P.NZ) {
            k = P.NZ - 1;
            fetch_u<F>(P, i, j, k, a);
            fetch_u<F>(P, i, j, k - 1, b);
            fetch_u<F>(P, i, j, k - 2, c);
            one_sided(a, b, c, false, G[2], G[5], G[8]);
        }
    } else {
        if (i == -1) {
            fetch_u<F>(P, -1, j, k, a);
            fetch_u<F>(P, 0, j, k, b);
            fetch_u<F>(P, 1, j, k, c);
            one_sided(a, b, c, true, G[0], G[3], G[6]);
        }
        if (i == P.NX) {
            fetch_u<F>(P, P.NX, j, k, a);
            fetch_u<F>(P, P.NX - 1, j, k, b);
            fetch_u<F>(P, P.NX - 2, j, k, c);
            one_sided(a, b, c, false, G[0], G[3], G[6]);
        }
        if (j == -1) {
            fetch_u<F>(P, i, -1, k, a);
            fetch_u<F>(P, i, 0, k, b);
            fetch_u<F>(P, i, 1, k, c);
            one_sided(a, b, c, true, G[1], G[4], G[7]);
        }
        if (j == P.NY) {
            fetch_u<F>(P, i, P.NY, k, a);
            fetch_u<F>(P, i, P.NY - 1, k, b);
            fetch_u<F>(P, i, P.NY - 2, k, c);
            one_sided(a, b, c, false, G[1], G[4], G[7]);
        }
        if (k == -1) {
            fetch_u<F>(P, i, j, -1, a);
            fetch_u<F>(P, i, j, 0, b);
            fetch_u<F>(P, i, j, 1, c);
            one_sided(a, b, c, true, G[2], G[5], G[8]);
        }
        if (k == P.NZ) {
            fetch_u<F>(P, i, j, P.NZ, a);
            fetch_u<F>(P, i, j, P.NZ - 1, b);
            fetch_u<F>(P, i, j, P.NZ - 2, c);
            one_sided(a, b, c, false, G[2], G[5], G[8]);
        }
    }
    if (P.NX == 1) G[0] = G[3] = G[6] = 0.0;
    if (P.NY == 1) G[1] = G[4] = G[7] = 0.0;
    if (P.NZ == 1) G[2] = G[5] = G[8] = 0.0;
}

template <typename F>
__device__ void grad_at(const PartParams &P, int i, int j, int k, double G[9])
{
    double a[3], b[3];
    // centred differences: both neighbours always exist for nodes -1..N
    fetch_u<F>(P, i + 1, j, k, a);
    fetch_u<F>(P, i - 1, j, k, b);
    G[0] = (a[0] - b[0]) / 2.0;
    G[3] = (a[1] - b[1]) / 2.0;
    G[6] = (a[2] - b[2]) / 2.0;
    fetch_u<F>(P, i, j + 1, k, a);
    fetch_u<F>(P, i, j - 1, k, b);
    G[1] = (a[0] - b[0]) / 2.0;
    G[4] = (a[1] - b[1]) / 2.0;
    G[7] = (a[2] - b[2]) / 2.0;
    fetch_u<F>(P, i, j, k + 1, a);
    fetch_u<F>(P, i, j, k - 1, b);
    G[2] = (a[0] - b[0]) / 2.0;
    G[5] = (a[1] - b[1]) / 2.0;
    G[8] = (a[2] - b[2]) / 2.0;
    grad_u_edges<F>(P, i, j, k, G);
}

// second-order one-sided difference of a scalar (SECOND_ORDER_GRAD with unit
// spacing: a0 = -3/2, a1 = 2, a2 = -1/2 forward, the negatives backward)
__device__ __forceinline__ double one_sided_s(double a, double b, double c, bool fwd)
{
    const double a0 = fwd ? -1.5 : 1.5, a1 = fwd ? 2.0 : -2.0, a2 = fwd ? -0.5 : 0.5;
    return a0 * a + a1 * b + a2 * c;
}

// gradient_scalar of the temperature with LAGRANGE on a two-node halo
// (gradients.c:505-827): centred differences; with temperature walls
// (LB_TEMPERATURE_BC_Y) one-sided at the first and last rows of y (:560-626)
// and the interpolation fix that moves a ghost node's evaluation point to the
// first interior node (:657-725); without them the one-sided difference at
// the ghost node itself (:762-825); components along a size-one axis zeroed.
template <typename F>
__device__ void grad_t_edges(const PartParams &P, int i, int j, int k, double g[3])
{
    auto T = [&](int a, int b, int c) { return fetch_t<F>(P, a, b, c); };
    if (P.wall_t) {
        if (j == 0) g[1] = one_sided_s(T(i, 0, k), T(i, 1, k), T(i, 2, k), true);
        if (j == P.NY - 1) g[1] = one_sided_s(T(i, P.NY - 1, k), T(i, P.NY - 2, k), T(i, P.NY - 3, k), false);
        if (i == -1) {
            i = 0;
            g[0] = one_sided_s(T(0, j, k), T(1, j, k), T(2, j, k), true);
        }
        if (i == P.NX) {
            i = P.NX - 1;
            g[0] = one_sided_s(T(i, j, k), T(i - 1, j, k), T(i - 2, j, k), false);
        }
        if (j == -1) {
            j = 0;
            g[1] = one_sided_s(T(i, 0, k), T(i, 1, k), T(i, 2, k), true);
        }
        if (j == P.NY) {
            j = P.NY - 1;
            g[1] = one_sided_s(T(i, j, k), T(i, j - 1, k), T(i, j - 2, k), false);
        }
        if (k == -1) {
            k = 0;
            g[2] = one_sided_s(T(i, j, 0), T(i, j, 1), T(i, j, 2), true);
        }
        if (k == P.NZ) {
            k = P.NZ - 1;
            g[2] = one_sided_s(T(i, j, k), T(i, j, k - 1), T(i, j, k - 2), false);
        }
    } else {
        if (i == -1) g[0] = one_sided_s(T(-1, j, k), T(0, j, k), T(1, j, k), true);
        if (i == P.NX) g[0] = one_sided_s(T(P.NX, j, k), T(P.NX - 1, j, k), T(P.NX - 2, j, k), false);
        if (j == -1) g[1] = one_sided_s(T(i, -1, k), T(i, 0, k), T(i, 1, k), true);
        if (j == P.NY) g[1] = one_sided_s(T(i, P.NY, k), T(i, P.NY - 1, k), T(i, P.NY - 2, k), false);
        if (k == -1) g[2] = one_sided_s(T(i, j, -1), T(i, j, 0), T(i, j, 1), true);
        if (k == P.NZ) g[2] = one_sided_s(T(i, j, P.NZ), T(i, j, P.NZ - 1), T(i, j, P.NZ - 2), false);
    }
    if (P.NX == 1) g[0] = 0.0;
    if (P.NY == 1) g[1] = 0.0;
    if (P.NZ == 1) g[2] = 0.0;
}

template <typename F>
__device__ void grad_t_at(const PartParams &P, int i, int j, int k, double g[3])
{
    g[0] = (fetch_t<F>(P, i + 1, j, k) - fetch_t<F>(P, i - 1, j, k)) / 2.0;
    g[1] = (fetch_t<F>(P, i, j + 1, k) - fetch_t<F>(P, i, j - 1, k)) / 2.0;
    g[2] = (fetch_t<F>(P, i, j, k + 1) - fetch_t<F>(P, i, j, k - 1)) / 2.0;
    grad_t_edges<F>(P, i, j, k, g);
}

// grad u and grad t at one node from the same six neighbour reads
template <typename F>
__device__ void grad_ut_at(const PartParams &P, int i, int j, int k, double G[9], double g[3])
{
    double a[3], b[3], ta, tb;
    fetch_ut<F>(P, i + 1, j, k, a, ta);
    fetch_ut<F>(P, i - 1, j, k, b, tb);
    G[0] = (a[0] - b[0]) / 2.0;
    G[3] = (a[1] - b[1]) / 2.0;
    G[6] = (a[2] - b[2]) / 2.0;
    g[0] = (ta - tb) / 2.0;
    fetch_ut<F>(P, i, j + 1, k, a, ta);
    fetch_ut<F>(P, i, j - 1, k, b, tb);
    G[1] = (a[0] - b[0]) / 2.0;
    G[4] = (a[1] - b[1]) / 2.0;
    G[7] = (a[2] - b[2]) / 2.0;
    g[1] = (ta - tb) / 2.0;
    fetch_ut<F>(P, i, j, k + 1, a, ta);
    fetch_ut<F>(P, i, j, k - 1, b, tb);
    G[2] = (a[0] - b[0]) / 2.0;
    G[5] = (a[1] - b[1]) / 2.0;
    G[8] = (a[2] - b[2]) / 2.0;
    g[2] = (ta - tb) / 2.0;
    grad_u_edges<F>(P, i, j, k, G);
    grad_t_edges<F>(P, i, j, k, g);
}

// sendrecv_particles on one rank (lagrange.c:4060-4116): wrap (with
// LAGRANGE_WRAP), keep the particle when its wrapped position lies in [0, S).
__device__ __forceinline__ void wrap_keep(const PartParams &P, double &x, double &y, double &z, int64_t p)
{
    const double SX = P.NX, SY = P.NY, SZ = P.NZ;
    const double px = wrapd(x, SX), py = wrapd(y, SY), pz = wrapd(z, SZ);
    if (lflag(P, CH4G_LAGRANGE_WRAP)) {
        x = wrapd(x, SX);
        z = wrapd(z, SZ);
        y = wrapd(y, SY);
    }
    if (!(px >= 0.0 && px < SX && py >= 0.0 && py < SY && pz >= 0.0 && pz < SZ)) P.alive[p] = 0;
}

// interpolate_vector_at_particles(u) [with the gradient] and
// interpolate_scalar_at_particles(t) [with its gradient] at the stencil
// (lagrange.c:1211-1593): corner values weighted in the reference's order.
template <bool GRAD, typename F>
__device__ __forceinline__ void interpolate(const PartParams &P, const Stencil &st, double u[3], double G[9],
                                            double &t, double gt[3])
{
    if (P.fluid && P.thermal && P.wall_f == P.wall_t) {
        // both fields: each corner (and gradient neighbour) read once
        double cu[3], tc;
        double v0 = 0, v1 = 0, v2 = 0, s = 0;
        for (int c = 0; c < 8; c++) {
            fetch_ut<F>(P, st.im + cdx(c), st.jm + cdy(c), st.km + cdz(c), cu, tc);
            if (c == 0) {
                v0 = cu[0] * st.w[0];
                v1 = cu[1] * st.w[0];
                v2 = cu[2] * st.w[0];
                s = tc * st.w[0];
            } else {
                v0 += cu[0] * st.w[c];
                v1 += cu[1] * st.w[c];
                v2 += cu[2] * st.w[c];
                s = s + tc * st.w[c];
            }
        }
        u[0] = v0;
        u[1] = v1;
        u[2] = v2;
        t = s;
        if constexpr (GRAD) {
            double acc[9], Gc[9], acct[3], gc[3];
            for (int c = 0; c < 8; c++) {
                grad_ut_at<F>(P, st.im + cdx(c), st.jm + cdy(c), st.km + cdz(c), Gc, gc);
                for (int m = 0; m < 9; m++) acc[m] = (c == 0) ? Gc[m] * st.w[0] : acc[m] + Gc[m] * st.w[c];
                for (int m = 0; m < 3; m++) acct[m] = (c == 0) ? gc[m] * st.w[0] : acct[m] + gc[m] * st.w[c];
            }
            for (int m = 0; m < 9; m++) G[m] = acc[m];
            for (int m = 0; m < 3; m++) gt[m] = acct[m];
        }
        return;
    }
    if (P.fluid) {
        double cu[3];
        double v0 = 0, v1 = 0, v2 = 0;
        for (int c = 0; c < 8; c++) {
            fetch_u<F>(P, st.im + cdx(c), st.jm + cdy(c), st.km + cdz(c), cu);
            if (c == 0) {
                v0 = cu[0] * st.w[0];
                v1 = cu[1] * st.w[0];
                v2 = cu[2] * st.w[0];
            } else {
                v0 += cu[0] * st.w[c];
                v1 += cu[1] * st.w[c];
                v2 += cu[2] * st.w[c];
            }
        }
        u[0] = v0;
        u[1] = v1;
        u[2] = v2;
        if constexpr (GRAD) {
            double acc[9], Gc[9];
            for (int c = 0; c < 8; c++) {
                grad_at<F>(P, st.im + cdx(c), st.jm + cdy(c), st.km + cdz(c), Gc);
                for (int m = 0; m < 9; m++) acc[m] = (c == 0) ? Gc[m] * st.w[0] : acc[m] + Gc[m] * st.w[c];
            }
            for (int m = 0; m < 9; m++) G[m] = acc[m];
        }
    }
    if (P.thermal) {
        double s = 0;
        for (int c = 0; c < 8; c++) {
            double tc = fetch_t<F>(P, st.im + cdx(c), st.jm + cdy(c), st.km + cdz(c));
            s = (c == 0) ? tc * st.w[0] : s + tc * st.w[c];
        }
        t = s;
        if constexpr (GRAD) {
            // interpolate_scalar_at_particles with LAGRANGE_GRADIENT
            // (lagrange.c:1500-1576): the corner gradients weighted in the
            // corner order of the value
            double acc[3], gc[3];
            for (int c = 0; c < 8; c++) {
                grad_t_at<F>(P, st.im + cdx(c), st.jm + cdy(c), st.km + cdz(c), gc);
                for (int m = 0; m < 3; m++) acc[m] = (c == 0) ? gc[m] * st.w[0] : acc[m] + gc[m] * st.w[c];
            }
            for (int m = 0; m < 3; m++) gt[m] = acc[m];
        }
    }
}

// The interpolation of a sub-step alone, into the particle's field rows (u,
// grad u, t, grad t; left as they are where the stencil fails, as the
// sub-step keeps them), for configurations with gradients: the corner and
// neighbour reads of the gradients then run without the registers of the
// sub-step, which follows as a stream over the rows.
template <bool GRAD, typename F>
__global__ void __launch_bounds__(PB) k_interp(PartParams P)
{
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t p = (int64_t)lb * PB + threadIdx.x;
    if (p >= P.np || !P.alive[p]) return;
    double *S = P.s;
    const int64_t ld = P.ld;
    Stencil st;
    if (!stencil(P, S[PX * ld + p], S[PY * ld + p], S[PZ * ld + p], st)) return;
    double u[3], G[9], t, gt[3];
    interpolate<GRAD, F>(P, st, u, G, t, gt);
    if (P.fluid) {
        for (int a = 0; a < 3; a++) S[(PUX + a) * ld + p] = u[a];
        if constexpr (GRAD)
            for (int m = 0; m < 9; m++) S[(PG + m) * ld + p] = G[m];
    }
    if (P.thermal) {
        S[PT * ld + p] = t;
        if constexpr (GRAD)
            for (int a = 0; a < 3; a++) S[(PGT + a) * ld + p] = gt[a];
    }
}

// The initial-state interpolations of initial_conditions_particles, on the
// FP64 fields with the halo rules of every sub-step:
// LAGRANGE_INITIAL_VELOCITY_FLUID (lagrange.c:968-980, after
// boundary_conditions_hydro): interpolate_vector_at_particles(u) (u and, with
// the gradient, grad u) and v = u unless the Stokes start replaces it; P has
// fluid on and thermal off.  LAGRANGE_INITIAL_TEMPERATURE_FLUID
// (lagrange.c:1004-1015): interpolate_scalar_at_particles(t) (t and, with the
// gradient, dx_t..dz_t) and t_p = t; P has fluid off and thermal on.
template <bool GRAD, bool VEL>
__global__ void __launch_bounds__(PB) k_finit(PartParams P, int set_v)
{
    const int64_t p = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (p >= P.np || !P.alive[p]) return;
    double *S = P.s;
    const int64_t ld = P.ld;
    Stencil st;
    if (!stencil(P, S[PX * ld + p], S[PY * ld + p], S[PZ * ld + p], st)) return;
    double u[3] = {0.0, 0.0, 0.0}, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, t = 0.0, gt[3] = {0.0, 0.0, 0.0};
    interpolate<GRAD, double>(P, st, u, G, t, gt);
    if constexpr (VEL) {
        for (int a = 0; a < 3; a++) S[(PUX + a) * ld + p] = u[a];
        if constexpr (GRAD)
            for (int m = 0; m < 9; m++) S[(PG + m) * ld + p] = G[m];
        if (set_v)
            for (int a = 0; a < 3; a++) S[(PVX + a) * ld + p] = u[a];
    } else {
        S[PT * ld + p] = t;
        if constexpr (GRAD)
            for (int a = 0; a < 3; a++) S[(PGT + a) * ld + p] = gt[a];
        S[PTP * ld + p] = t;
    }
}

// One sub-step of every particle: interpolate_vector_at_particles(u),
// interpolate_scalar_at_particles(t) (lagrange.c:1211-1593) unless ROWS (the
// fields already in the rows, k_interp), move_particles (lagrange.c:2934-4025)
// and the wrap of sendrecv_particles.  Without gradients four waves per SIMD
// (spilling) were slower than the unconstrained register count; with
// gradients the bound of three waves is kept.
template <bool GRAD, typename F, bool ROWS>
__global__ void __launch_bounds__(PB, GRAD ? 3 : 1) k_particles(PartParams P)
{
    // rows are in cell order: consecutive blocks on one XCD share corner lines
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);  // the block of particles this block moves
    const int64_t p = (int64_t)lb * PB + threadIdx.x;
    int nd = 0, pflags = 0;  // draws of the sub-step; 1: random p, 2 / 4: p_y turned at the bottom / top wall
    if (p < P.np && P.alive[p]) {
        double *S = P.s;
        const int64_t ld = P.ld;
#define R(f) S[(int64_t)(f) * ld + p]
        double x = R(PX), y = R(PY), z = R(PZ);
        // ---- interpolation at the (wrapped) position, first: its dependent
        // corner reads are the latency-bound part, and the particle's other
        // rows are read afterwards so that they do not hold registers
        // meanwhile.  Rows the sub-step overwrites are read only where it
        // does not (a failed stencil keeps u, G, t, grad t; a tau that is
        // neither 0 nor positive keeps a and u_old).
        double ux = 0.0, uy = 0.0, uz = 0.0, t = 0.0;
        double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double gt[3] = {0.0, 0.0, 0.0};
        // ROWS: k_interp has written the fields of this sub-step into the rows
        bool ok = false;
        if constexpr (!ROWS) {
            Stencil st;
            ok = stencil(P, x, y, z, st);
            if (ok) {
                double u[3];
                interpolate<GRAD, F>(P, st, u, G, t, gt);
                ux = u[0];
                uy = u[1];
                uz = u[2];
            }
        }
        if (!(ok && P.fluid)) {
            ux = R(PUX);
            uy = R(PUY);
            uz = R(PUZ);
            if constexpr (GRAD)
                for (int m = 0; m < 9; m++) G[m] = R(PG + m);
        }
        const double tau = R(PTAU);
        const bool inertial = tau > 0.0, moved = tau == 0.0 || inertial;
        // rows a configuration never changes are neither read nor written:
        // the gradient without LAGRANGE_GRADIENT, Du/Dt without added mass,
        // the temperature rows without LB_TEMPERATURE; beta is read only by
        // the switches that use it, the velocity not by tracers (which take
        // the fluid's)
        const bool am = lflag(P, CH4G_LAGRANGE_ADDEDMASS);
        const bool use_beta = am || lflag(P, CH4G_LAGRANGE_SMALLTAUD_BETA | CH4G_LAGRANGE_SMALLTAUD_FLOATER |
                                               CH4G_LAGRANGE_BC_INELASTIC_REINJECT);
        const double beta = use_beta ? R(PBETA) : 0.0;
        // LAGRANGE_GRAVITY_VARIABLE: the factor of the particle's type on g
        // (lagrange.c:3314-3335, :3824-3827, :3905-3908); 1 otherwise, which
        // leaves every product it enters unchanged
        const double gc = lflag(P, CH4G_LAGRANGE_GRAVITY_VARIABLE) ? R(PGC) : 1.0;
        double vx = 0.0, vy = 0.0, vz = 0.0;
        if (tau != 0.0) {
            vx = R(PVX);
            vy = R(PVY);
            vz = R(PVZ);
        }
        double vxo = R(PVXO), vyo = R(PVYO), vzo = R(PVZO);
        double ax = 0.0, ay = 0.0, az = 0.0, axo = 0.0, ayo = 0.0, azo = 0.0;
        double uxo = 0.0, uyo = 0.0, uzo = 0.0;
        if (!moved) {
            ax = R(PAX);
            ay = R(PAY);
            az = R(PAZ);
        }
        if (inertial) {
            axo = R(PAXO);
            ayo = R(PAYO);
            azo = R(PAZO);
        }
        if (!moved || (inertial && am && beta != 0.0)) {
            uxo = R(PUXO);
            uyo = R(PUYO);
            uzo = R(PUZO);
        }
        double dtu[3] = {0.0, 0.0, 0.0};
        if (am) {
            dtu[0] = R(PDTU);
            dtu[1] = R(PDTU + 1);
            dtu[2] = R(PDTU + 2);
        }
        double told = 0.0, dtt = 0.0;
        if (P.thermal) {
            if (!ok) {
                t = R(PT);
                if constexpr (GRAD) {
                    gt[0] = R(PGT);
                    gt[1] = R(PGT + 1);
                    gt[2] = R(PGT + 2);
                }
            }
            told = R(PTOLD);
            dtt = R(PDTT);
        }
        const bool orient = lflag(P, CH4G_LAGRANGE_ORIENTATION);
        const bool second = orient && lflag(P, CH4G_LAGRANGE_ORIENTATION_SECONDORIENTATION);
        const bool rnd = orient && (P.flags2 & CH4G_LAGRANGE_ORIENTATION_RANDOM);
        double pv[3] = {0.0, 0.0, 0.0}, fold[3] = {0.0, 0.0, 0.0};
        double nv[3] = {0.0, 0.0, 0.0}, nfold[3] = {0.0, 0.0, 0.0};
        if (orient) {
            for (int i = 0; i < 3; i++) {
                pv[i] = R(PPO + i);
                fold[i] = R(PDTP + i);
            }
        }
        if (second) {
            for (int i = 0; i < 3; i++) {
                nv[i] = R(PNO + i);
                nfold[i] = R(PDTN + i);
            }
        }

        // ---- move_particles
        const double dt = P.dt;
        const bool first = P.first;
        const double gx = P.grav[0], gy = P.grav[1], gz = P.grav[2];
        if (P.thermal) {
            if (first) {
                told = t;
            } else {
                dtt = (t - told) / dt;
                told = t;
            }
        }
        if (tau < 0.0) {  // Eulerian probe: stays put, records the fluid velocity (lagrange.c:3024-3033)
            vx = ux;
            vy = uy;
            vz = uz;
        }
        if (tau == 0.0) {  // fluid tracer (lagrange.c:3010-3227)
            vx = ux;
            vy = uy;
            vz = uz;
            if (orient && lflag(P, CH4G_LAGRANGE_ORIENTATION_ACTIVE)) {
                // swimming (lagrange.c:3134-3186): the modes in the
                // reference's #elif order, else along p on top of u
                const double sw = R(PSWIM);
                const uint32_t f2 = P.flags2;
                bool jumped = false;
                if constexpr (GRAD) {
                    if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP) {
                        // copepod jumps (lagrange.c:3048-3132): react to the
                        // shear rate sqrt(2) sqrt(...) as the reference writes
                        // it (its second term is dy_uy dz_uy), or to t
                        double sr;
                        if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_TEMPERATURE) {
                            sr = t;
                        } else {
                            const double s01 = G[1] + G[3], s02 = G[2] + G[6], s12 = G[7] + G[5];
                            sr = sqrt(2.) * sqrt(G[0] * G[0] + G[4] * G[5] + G[8] * G[8] + 0.5 * s01 * s01 +
                                                 0.5 * s02 * s02 + 0.5 * s12 * s12);
                            R(PSR) = sr;
                        }
                        const double jt = R(PJT);
                        double tfj = R(PTFJ);
                        double amp;
                        if (tfj > -jt * P.log_001) {
                            amp = 0.0;
                            const double csr = R(PCSR);
                            const bool go = (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_JUMP_HIGHPASS) ? sr < csr : sr > csr;
                            if (go) {
                                tfj = 0.0;
                                amp = sw;
                                for (int i = 0; i < 3; i++) R(PPJ + i) = pv[i];
                            }
                        } else {
                            amp = sw * exp(-tfj / jt);
                            if (jt == 0.0) amp = 0.0;
                        }
                        vx += amp * R(PPJ);
                        vy += amp * R(PPJ + 1);
                        vz += amp * R(PPJ + 2);
                        R(PTFJ) = tfj + dt;
                        jumped = true;
                    }
                }
                if (jumped) {
                    // the jump is the first of the reference's #elif modes
                } else if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_BALLISTIC) {
                    vx = sw * pv[0];
                    vy = sw * pv[1];
                    vz = sw * pv[2];
                } else if (f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_TEMPERATURE) {
                    const double react = dtt > 0.0 ? 1.0 : 0.0;
                    vx += react * sw * pv[0];
                    vy += react * sw * pv[1];
                    vz += react * sw * pv[2];
                } else if ((f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_SECONDORIENTATION) ||
                           ((f2 & CH4G_LAGRANGE_ORIENTATION_ACTIVE_SELECTIVEORIENTATION) && !(R(PASP) >= 1.0))) {
                    vx += sw * nv[0];
                    vy += sw * nv[1];
                    vz += sw * nv[2];
                } else {
                    vx += sw * pv[0];
                    vy += sw * pv[1];
                    vz += sw * pv[2];
                }
            }
            if (first) {
                vxo = vx;
                vyo = vy;
                vzo = vz;
            }
            ax = (vx - vxo) / dt;
            ay = (vy - vyo) / dt;
            az = (vz - vzo) / dt;
            if (first) {
                x += dt * vx;
                y += dt * vy;
                z += dt * vz;
            } else {
                x += dt * 0.5 * (3.0 * vx - vxo);
                y += dt * 0.5 * (3.0 * vy - vyo);
                z += dt * 0.5 * (3.0 * vz - vzo);
            }
            vxo = vx;
            vyo = vy;
            vzo = vz;
            uxo = ux;
            uyo = uy;
            uzo = uz;
        }
        if (tau > 0.0) {  // inertial particle (lagrange.c:3229-3466)
            double dux = 0.0, duy = 0.0, duz = 0.0;
            bool spheroid = false;  // the switches below need the gradient build
            if constexpr (GRAD) spheroid = (P.flags2 & CH4G_LAGRANGE_ORIENTATION_DRAG) != 0;
            if (spheroid) {
                // Stokes drag of a spheroid of aspect ratio alpha along p
                // (lagrange.c:3267-3310): M = c_perp I + (c_par - c_perp) p p,
                // rate 1 / (tau alpha); c_perp, c_par formed on the host
                const double alpha = R(PASP), cpe = R(PCPE), cpa = R(PCPA);
                const double it = 1.0 / (tau * alpha);
                const double d = cpa - cpe;
                const double m00 = cpe + d * pv[0] * pv[0], m11 = cpe + d * pv[1] * pv[1], m22 = cpe + d * pv[2] * pv[2];
                const double m01 = d * pv[0] * pv[1], m02 = d * pv[0] * pv[2], m12 = d * pv[1] * pv[2];
                const double uvx = ux - vx, uvy = uy - vy, uvz = uz - vz;
                ax = (m00 * uvx + m01 * uvy + m02 * uvz) * it;
                ay = (m01 * uvx + m11 * uvy + m12 * uvz) * it;
                az = (m02 * uvx + m12 * uvy + m22 * uvz) * it;
            } else {
                const double invtau = 1.0 / tau;
                ax = (ux - vx) * invtau;
                ay = (uy - vy) * invtau;
                az = (uz - vz) * invtau;
            }
            if (GRAD && !spheroid && (P.flags2 & CH4G_LAGRANGE_ADDEDMASS_WAKEDRAG)) {
                // Schiller-Naumann correction (lagrange.c:3242-3264): diameter
                // from tau = r^2 / (3 beta nu), particle Reynolds number of the
                // slip velocity, extra rate 0.15 Re_p^0.687 / tau
                const double diameter_p = 2.0 * sqrt(tau * 3.0 * beta * P.nu);
                const double re_p = sqrt((ux - vx) * (ux - vx) + (uy - vy) * (uy - vy) + (uz - vz) * (uz - vz)) *
                                    diameter_p / P.nu;
                const double wtau = 0.15 * pow(re_p, 0.687) / tau;
                ax += (ux - vx) * wtau;
                ay += (uy - vy) * wtau;
                az += (uz - vz) * wtau;
            }
            const bool grav = lflag(P, CH4G_LAGRANGE_GRAVITY);
            if (grav) {
                ax -= gc * gx;
                ay -= gc * gy;
                az -= gc * gz;
            }
            if (lflag(P, CH4G_LAGRANGE_ADDEDMASS) && beta != 0.0) {
                if (grav) {
                    ax -= (-beta) * gc * gx;
                    ay -= (-beta) * gc * gy;
                    az -= (-beta) * gc * gz;
                }
                if (first) {
                    uxo = ux;
                    uyo = uy;
                    uzo = uz;
                }
                dux = (ux - uxo) / dt + (ux - vx) * G[0] + (uy - vy) * G[1] + (uz - vz) * G[2];
                duy = (uy - uyo) / dt + (ux - vx) * G[3] + (uy - vy) * G[4] + (uz - vz) * G[5];
                duz = (uz - uzo) / dt + (ux - vx) * G[6] + (uy - vy) * G[7] + (uz - vz) * G[8];
                ax += dux * beta;
                ay += duy * beta;
                az += duz * beta;
                dtu[0] = dux;
                dtu[1] = duy;
                dtu[2] = duz;
                if (GRAD && (P.flags2 & CH4G_LAGRANGE_ADDEDMASS_LIFT)) {
                    // lift with C_L = 1/2 (lagrange.c:3365-3380): (beta/3) (u - v) x omega
                    const double ox = G[7] - G[5], oy = G[2] - G[6], oz = G[3] - G[1];
                    const double lc = beta / 3.0;
                    ax += lc * ((uy - vy) * oz - (uz - vz) * oy);
                    ay += lc * ((uz - vz) * ox - (ux - vx) * oz);
                    az += lc * ((ux - vx) * oy - (uy - vy) * ox);
                }
            }
            if (first) {
                x += dt * vx;
                y += dt * vy;
                z += dt * vz;
            } else {
                x += dt * 0.5 * (3.0 * vx - vxo);
                y += dt * 0.5 * (3.0 * vy - vyo);
                z += dt * 0.5 * (3.0 * vz - vzo);
            }
            vxo = vx;
            vyo = vy;
            vzo = vz;
            uxo = ux;
            uyo = uy;
            uzo = uz;
            if (first) {
                vx += dt * ax;
                vy += dt * ay;
                vz += dt * az;
            } else {
                vx += dt * 0.5 * (3.0 * ax - axo);
                vy += dt * 0.5 * (3.0 * ay - ayo);
                vz += dt * 0.5 * (3.0 * az - azo);
            }
            axo = ax;
            ayo = ay;
            azo = az;
            if (lflag(P, CH4G_LAGRANGE_SMALLTAUD_BETA) && tau < 10.0 && tau > 0.0) {
                vx = ux;
                vy = uy;
                vz = uz;
                vx -= gx * (1.0 - beta) * tau;
                vy -= gy * (1.0 - beta) * tau;
                vz -= gz * (1.0 - beta) * tau;
                vx -= dux * (1.0 - beta) * tau;
                vy -= duy * (1.0 - beta) * tau;
                vz -= duz * (1.0 - beta) * tau;
            }
            if (lflag(P, CH4G_LAGRANGE_SMALLTAUD_FLOATER) && tau > 0.0) {
                vx = ux;
                vy = uy;
                vz = uz;
                const double f1 = tau * (1.0 - beta + 2. / 3. * P.beta_t * (t - 0.4));
                vx += f1 * (-gx - dtu[0]);
                vy += f1 * (-gy - dtu[1]);
                vz += f1 * (-gz - dtu[2]);
            }
        }

        // ---- orientation (lagrange.c:3470-3752)
        if (orient) {
            double rhs[3] = {0.0, 0.0, 0.0};
            if (lflag(P, CH4G_LAGRANGE_ORIENTATION_JEFFREY)) {
                // Jeffery's equation dp/dt = W p + S' p - (p.S' p) p, S' =
                // f(alpha) S, with the gradient matrix A[i][j] = d_j u_i (G)
                const double alpha = R(PASP);
                const double fa = (alpha * alpha - 1.0) / (1.0 + alpha * alpha);
                double Sm[3][3], Wm[3][3];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        Sm[i][j] = 0.5 * (G[3 * i + j] + G[3 * j + i]);
                        Wm[i][j] = 0.5 * (G[3 * i + j] - G[3 * j + i]);
                    }
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) Sm[i][j] *= fa;
                if (lflag(P, CH4G_LAGRANGE_ORIENTATION_JEFFREY_GYROTAXIS)) {
                    const double gv = R(PGYV);
                    if (gv != 0) {
                        // -1/(2 v0) g_i p_j enters the first column of S'
                        // only: the reference adds the other six products
                        // to the gradient matrix after S' and W are formed
                        // (lagrange.c:3553-3562), where they have no effect
                        const double gyro = -0.5 / gv;
                        double va[3];
                        if (lflag(P, CH4G_LAGRANGE_GRAVITY)) {
                            va[0] = gyro * (-gx + ax);
                            va[1] = gyro * (-gy + ay);
                            va[2] = gyro * (-gz + az);
                        } else {
                            va[0] = gyro * 0.0;
                            va[1] = gyro * 0.0;
                            va[2] = gyro * 1.0;
                        }
                        for (int i = 0; i < 3; i++) Sm[i][0] += va[i] * pv[0];
                    }
                }
                double tmp[3] = {0.0, 0.0, 0.0};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) tmp[i] += Sm[i][j] * pv[j];
                double oso = 0.0;
                for (int i = 0; i < 3; i++) oso += pv[i] * tmp[i];
                for (int i = 0; i < 3; i++) {
                    rhs[i] = 0.0;
                    for (int j = 0; j < 3; j++) rhs[i] += Wm[i][j] * pv[j] + (Sm[i][j] * pv[j]);
                    rhs[i] -= pv[i] * oso;
                }
                double rhsn[3] = {0.0, 0.0, 0.0};
                if (second) {
                    // n: W n - (p . S' n) p, with p before its update
                    // (lagrange.c:3586-3616)
                    double tn[3] = {0.0, 0.0, 0.0};
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) tn[i] += Sm[i][j] * nv[j];
                    double osn = 0.0;
                    for (int i = 0; i < 3; i++) osn += pv[i] * tn[i];
                    for (int i = 0; i < 3; i++) {
                        for (int j = 0; j < 3; j++) rhsn[i] += Wm[i][j] * nv[j];
                        rhsn[i] -= pv[i] * osn;
                    }
                }
                for (int i = 0; i < 3; i++) {
                    if (first)
                        pv[i] = pv[i] + dt * rhs[i];
                    else
                        pv[i] = pv[i] + 0.5 * dt * (3. * rhs[i] - fold[i]);
                    fold[i] = rhs[i];
                }
                if (second)
                    for (int i = 0; i < 3; i++) {
                        if (first)
                            nv[i] = nv[i] + dt * rhsn[i];
                        else
                            nv[i] = nv[i] + 0.5 * dt * (3. * rhsn[i] - nfold[i]);
                        nfold[i] = rhsn[i];
                    }
            }
            double norm = 0.0;
            for (int i = 0; i < 3; i++) norm += pv[i] * pv[i];
            for (int i = 0; i < 3; i++) pv[i] /= sqrt(norm);
            if (second) {
                double nn = 0.0;
                for (int i = 0; i < 3; i++) nn += nv[i] * nv[i];
                for (int i = 0; i < 3; i++) nv[i] /= sqrt(nn);
            }
            if (rnd) {
                // LAGRANGE_ORIENTATION_RANDOM (lagrange.c:3698-3713): p is drawn
                // anew from the particle's first two draws of the sub-step, in
                // the draw pass; the row keeps the old p that dp/dt is taken from
                nd += 2;
                pflags |= 1;
            }
        }

        // ---- wall rules at y (lagrange.c:3754-3993); x and z of a random
        // re-injection are filled in by the ordered draw pass.  A re-injected
        // particle takes the temperature of the wall it re-enters at (the
        // later of the two rules wins, as in the reference's sequence)
        int rewall = 0;  // 1: re-entered at the top (T_top), 2: at the bottom (T_bot)
        if (P.wall_f) {
            const bool inel = lflag(P, CH4G_LAGRANGE_BC_INELASTIC);
            const bool reinj = lflag(P, CH4G_LAGRANGE_BC_INELASTIC_REINJECT);
            const bool norand = lflag(P, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_NORANDOM);
            const bool vzero = lflag(P, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_ZERO);
            const double fac1 = inel ? 0.0 : -1.0, fac2 = inel ? 0.0 : 1.0;
            // LAGRANGE_BC_RADIUSandDENSITY: the walls act at the particle's
            // radius sqrt(3 nu beta tau) (lagrange.c:3777-3788); 0 otherwise
            const double radius =
                (P.flags2 & CH4G_LAGRANGE_BC_RADIUSANDDENSITY) ? sqrt(tau * P.nu * 3.0 * beta) : 0.0;
            const double SY = P.NY;
            if (y < radius) {
                if (!reinj) {
                    y = fac1 * (y - radius) + radius;
                    vx *= fac2;
                    if (vy < 0.0) vy *= fac1;
                    vz *= fac2;
                }
                if (inel && reinj) {
                    y += SY;
                    if (!norand) nd += 2;
                    if (vzero) {
                        vx = 0.0;
                        vy = 0.0;
                        vz = 0.0;
                    } else {
                        vx = -gx * gc * (1.0 - beta) * tau;
                        vy = -gy * gc * (1.0 - beta) * tau;
                        vz = -gz * gc * (1.0 - beta) * tau;
                    }
                    if (lflag(P, CH4G_LAGRANGE_BC_INELASTIC_REINJECT_VELOCITY_STOKES_HALF)) {
                        vx /= 2.0;
                        vy /= 2.0;
                        vz /= 2.0;
                    }
                    rewall = 1;  // lagrange.c:3839-3842
                }
                // LAGRANGE_BC_INELASTIC_SEDIMENT: one more deposition (lagrange.c:3845-3863)
                if (inel && (P.flags2 & CH4G_LAGRANGE_BC_INELASTIC_SEDIMENT)) R(PSED) += 1.0;
                if (lflag(P, CH4G_LAGRANGE_ORIENTATION_BC)) {
                    if (rnd)
                        pflags |= 2;  // applied to the drawn p by the draw pass
                    else if (pv[1] < 0.0)
                        pv[1] *= -1.0;
                }
            }
            if (y >= SY - radius) {
                if (!reinj) {
                    y = fac1 * (y - (SY - radius)) + (SY - radius);
                    vx *= fac2;
                    if (vy > 0.0) vy *= fac1;
                    vz *= fac2;
                }
                if (inel && reinj) {
                    y -= (SY - FLT_EPS);
                    if (!norand) nd += 2;
                    if (vzero) {
                        vx = 0.0;
                        vy = 0.0;
                        vz = 0.0;
                    } else {
                        vx = -gx * gc * (1.0 - beta) * tau;
                        vy = -gy * gc * (1.0 - beta) * tau;
                        vz = -gz * gc * (1.0 - beta) * tau;
                    }
                    rewall = 2;  // lagrange.c:3915-3918
                }
                if (lflag(P, CH4G_LAGRANGE_ORIENTATION_BC)) {
                    if (rnd)
                        pflags |= 4;
                    else if (pv[1] > 0.0)
                        pv[1] *= -1.0;
                }
            }
        }
        if (nd == 0) wrap_keep(P, x, y, z, p);

        // ---- polymer conformation tensor along the trajectory
        // (polymer.c:150-240, after the wall rules, lagrange.c:3998-4002):
        // dc/dt = A^T c + c A + (I - c) / tau_p with A[i][j] = d_j u_i, Adams-
        // Bashforth over the lattice step time_dt = 1 (the reference's Euler
        // branch, itime == 0, is never taken inside move_particles)
        if constexpr (GRAD) {
            if (P.flags2 & CH4G_LAGRANGE_POLYMER) {
                double C[3][3], Fc[3][3];
                const double cd[6] = {R(PCT), R(PCT + 1), R(PCT + 2), R(PCT + 3), R(PCT + 4), R(PCT + 5)};
                C[0][0] = cd[0];
                C[1][1] = cd[1];
                C[2][2] = cd[2];
                C[0][1] = C[1][0] = cd[3];
                C[1][2] = C[2][1] = cd[4];
                C[0][2] = C[2][0] = cd[5];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        double f = 0.0;
                        for (int k = 0; k < 3; k++) f += G[3 * k + i] * C[k][j] + C[i][k] * G[3 * k + j];
                        f += ((i == j ? 1.0 : 0.0) - C[i][j]) * P.inv_tau_polymer;
                        Fc[i][j] = f;
                    }
                const int ui[6] = {0, 1, 2, 0, 1, 0}, uj[6] = {0, 1, 2, 1, 2, 2};
                for (int m = 0; m < 6; m++) {
                    const int i = ui[m], j = uj[m];
                    R(PCT + m) = C[i][j] + 0.5 * 1.0 * (3. * Fc[i][j] - R(PDCT + m));
                    R(PDCT + m) = Fc[i][j];
                }
            }
        }

        // ---- particle temperature (lagrange.c:4007-4020): relaxation to the
        // fluid's t over the heat response time, T_p <- (T_p + fac1 t)
        // exp(-fac1); fac1 and exp(-fac1) are per-particle constants formed
        // on the host when the rows are loaded
        if (P.thermal && lflag(P, CH4G_LAGRANGE_TEMPERATURE)) {
            double tp = rewall == 1 ? P.T_top : rewall == 2 ? P.T_bot : R(PTP);
            const double fac1 = R(PTFAC);
            tp = (tp + fac1 * t) * R(PTEXP);
            R(PTP) = tp;
        }

        R(PX) = x;
        R(PY) = y;
        R(PZ) = z;
        R(PVX) = vx;
        R(PVY) = vy;
        R(PVZ) = vz;
        R(PVXO) = vxo;
        R(PVYO) = vyo;
        R(PVZO) = vzo;
        R(PAX) = ax;
        R(PAY) = ay;
        R(PAZ) = az;
        if (inertial) {
            R(PAXO) = axo;
            R(PAYO) = ayo;
            R(PAZO) = azo;
        }
        if constexpr (!ROWS) {
            R(PUX) = ux;
            R(PUY) = uy;
            R(PUZ) = uz;
        }
        R(PUXO) = uxo;
        R(PUYO) = uyo;
        R(PUZO) = uzo;
        if constexpr (GRAD && !ROWS)
            for (int m = 0; m < 9; m++) R(PG + m) = G[m];
        if (am) {
            R(PDTU) = dtu[0];
            R(PDTU + 1) = dtu[1];
            R(PDTU + 2) = dtu[2];
        }
        if (orient && !rnd) {
            for (int i = 0; i < 3; i++) {
                R(PPO + i) = pv[i];
                R(PDTP + i) = fold[i];
            }
        }
        if (second) {
            for (int i = 0; i < 3; i++) {
                R(PNO + i) = nv[i];
                R(PDTN + i) = nfold[i];
            }
        }
        if (P.thermal) {
            if constexpr (!ROWS) R(PT) = t;
            R(PTOLD) = told;
            R(PDTT) = dtt;
            if constexpr (GRAD && !ROWS) {
                R(PGT) = gt[0];
                R(PGT + 1) = gt[1];
                R(PGT + 2) = gt[2];
            }
        }
#undef R
    }
    if (P.bsum) {
        // draws of this block, for the ordered re-injection pass
        if (p < P.np) P.ndraw[p] = nd | (pflags << 16);
        __shared__ int red[PB / 64];
        int v = nd;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int w = 0; w < PB / 64; w++) s += red[w];
            P.bsum[lb] = s;
        }
    }
}

// Exclusive scan of the per-block draw counts (one block), then the stream
// state advances past every draw of the sub-step.  rng[1] keeps the state
// the draws start from.
__global__ void __launch_bounds__(1024) k_draw_offsets(int *bsum, int nb, uint64_t *rng)
{
    __shared__ int part[1024];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int a = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nb) bsum[i] = (int)(carry + part[threadIdx.x] - v);
        __syncthreads();
        if (threadIdx.x == 0) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rng[1] = rng[0];
        rng[0] = lcg_jump(rng[0], (uint64_t)carry);
    }
}

// The draws of each particle in order: with LAGRANGE_ORIENTATION_RANDOM the
// first two give p = random_vector (random.c:139-150) and dp/dt = (p - p_old)
// / dt, then p is normalised and turned at the walls as recorded; a
// re-injected particle takes x = SX r, z = SZ r from its last pair (the last
// wall crossed wins); then sendrecv's wrap.
__global__ void __launch_bounds__(PB) k_draw_apply(PartParams P)
{
    const int64_t p = (int64_t)blockIdx.x * PB + threadIdx.x;
    const int raw = p < P.np ? P.ndraw[p] : 0;
    const int nd = raw & 0xFFFF, pf = raw >> 16;
    __shared__ int sc[PB];
    sc[threadIdx.x] = nd;
    __syncthreads();
    for (int o = 1; o < PB; o <<= 1) {
        int a = threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
        __syncthreads();
        sc[threadIdx.x] += a;
        __syncthreads();
    }
    if (nd == 0) return;
    const uint64_t off = (uint64_t)P.bsum[blockIdx.x] + (uint64_t)(sc[threadIdx.x] - nd);
    const uint64_t base = P.rng[1];
    double *S = P.s;
    const int64_t ld = P.ld;
    if (pf & 1) {
        const double theta = acos(2.0 * lcg_value(lcg_jump(base, off + 1)) - 1.0);
        const double phi = 2 * 3.14159265359 * lcg_value(lcg_jump(base, off + 2));
        double st, ct, sp, cp;
        sincos(theta, &st, &ct);
        sincos(phi, &sp, &cp);
        double v[3] = {st * sp, st * cp, ct};
        double norm = 0.0;
        for (int i = 0; i < 3; i++) {
            S[(PDTP + i) * ld + p] = (v[i] - S[(PPO + i) * ld + p]) / P.dt;
            norm += v[i] * v[i];
        }
        for (int i = 0; i < 3; i++) v[i] /= sqrt(norm);
        if ((pf & 2) && v[1] < 0.0) v[1] *= -1.0;
        if ((pf & 4) && v[1] > 0.0) v[1] *= -1.0;
        for (int i = 0; i < 3; i++) S[(PPO + i) * ld + p] = v[i];
    }
    double x = S[(int64_t)PX * ld + p], y = S[(int64_t)PY * ld + p], z = S[(int64_t)PZ * ld + p];
    if (nd > ((pf & 1) ? 2 : 0)) {
        x = P.NX * lcg_value(lcg_jump(base, off + nd - 1));
        z = P.NZ * lcg_value(lcg_jump(base, off + nd));
    }
    wrap_keep(P, x, y, z, p);
    S[(int64_t)PX * P.ld + p] = x;
    S[(int64_t)PY * P.ld + p] = y;
    S[(int64_t)PZ * P.ld + p] = z;
}

// ---------------------------------------------------------------------------
// Two-way feedback (lagrange_feedback.c:11-256).
//
// The reference loops over its particles in order and adds each particle's
// fac * f_p * w_c into the force of the eight corners of its stencil cell
// (lagrange_feedback.c:149-176), from force_twoway = 0.  Here every particle
// writes its stencil cell under its reference index (fac * f_p, position and
// the index itself under its row); a stable sort by cell leaves each cell's
// particles in reference order; and one
// thread per node merges the lists of the eight cells that have the node as
// a corner, smallest reference index first, so each node sums exactly the
// reference's terms in the reference's order.  Deposits into ghost nodes
// never reach the force (add_sendrecv_borders_vector only adds into ghosts,
// lagrange_feedback.c:525-836): only interior nodes are summed.

// stencil cell (im, jm, km) in -1 .. N-1, shifted by one: (NX+1)(NY+1)(NZ+1) cells
__device__ __forceinline__ uint32_t fb_cell(const PartParams &P, int im, int jm, int km)
{
    return (uint32_t)((((int64_t)km + 1) * (P.NY + 1) + (jm + 1)) * (P.NX + 1) + (im + 1));
}

__global__ void __launch_bounds__(PB) k_fb_cells(PartParams P, uint32_t sent, uint32_t *kin, uint32_t *vin,
                                                 double *rec, double g0, double g1, double g2)
{
    const int64_t p = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (p >= P.np) return;
    const uint32_t o = (uint32_t)P.orig[p];
    uint32_t key = sent;
    const double *S = P.s;
    const int64_t ld = P.ld;
    const double fac = S[PVOL * ld + p];
    const double x = S[PX * ld + p], y = S[PY * ld + p], z = S[PZ * ld + p];
    Stencil st;
    // tracers (zero volume) deposit exact zeros: left out
    if (P.alive[p] && fac != 0.0 && stencil(P, x, y, z, st)) {
        const double dr = S[PDRAT * ld + p];
        double fpx, fpy, fpz;
        if (lflag(P, CH4G_LAGRANGE_TWOWAY_SMALLTAUD)) {
            fpx = (S[PDTU * ld + p] - g0) * (1 - dr);
            fpy = (S[(PDTU + 1) * ld + p] - g1) * (1 - dr);
            fpz = (S[(PDTU + 2) * ld + p] - g2) * (1 - dr);
        } else {
            fpx = S[PDTU * ld + p] - g0 + dr * (g0 - S[PAX * ld + p]);
            fpy = S[(PDTU + 1) * ld + p] - g1 + dr * (g1 - S[PAY * ld + p]);
            fpz = S[(PDTU + 2) * ld + p] - g2 + dr * (g2 - S[PAZ * ld + p]);
        }
        key = fb_cell(P, st.im, st.jm, st.km);
        // by row: rows are kept in cell order, so neighbouring nodes read
        // neighbouring records
        double *r = rec + 8 * p;
        r[0] = fac * fpx;
        r[1] = fac * fpy;
        r[2] = fac * fpz;
        // the wrapped position the stencil weights are formed from
        r[3] = wrapd(x, (double)P.NX);
        r[4] = wrapd(y, (double)P.NY);
        r[5] = wrapd(z, (double)P.NZ);
        r[6] = (double)o;  // the reference index orders the sums
    }
    // input in reference order: the stable sort keeps that order per cell
    kin[o] = key;
    vin[o] = (uint32_t)p;
}

// head[cell] = first sorted position of the cell (set), or -1 (clear)
__global__ void __launch_bounds__(PB) k_fb_heads(const uint32_t *keys, int64_t M, uint32_t sent, int *head, int set)
{
    const int64_t t = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (t >= M) return;
    const uint32_t key = keys[t];
    if (key == sent) return;
    if (!set)
        head[key] = -1;
    else if (t == 0 || keys[t - 1] != key)
        head[key] = (int)t;
}

// The deposits of every sorted particle, corner-major: dep[c][t][3] =
// (fac f_p) * w_c with the stencil's weights (lagrange.c:1271-1292), w_c =
// (X * Y) * Z, X = x - (im + 1/2) on the +x corners and (im + 3/2) - x on the
// -x corners, likewise Y and Z; ord[t] the particle's reference index.
__global__ void __launch_bounds__(PB) k_fb_deps(PartParams P, const uint32_t *__restrict__ keys,
                                                const uint32_t *__restrict__ vals, const double *__restrict__ rec,
                                                uint32_t sent, int64_t M, double *__restrict__ dep,
                                                uint32_t *__restrict__ ord)
{
    const int64_t t = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (t >= M || keys[t] == sent) return;
    const double *q = rec + 8 * (int64_t)vals[t];
    const double px = q[3], py = q[4], pz = q[5];
    const double im = floor(px - 0.5), jm = floor(py - 0.5), km = floor(pz - 0.5);
    const double X[2] = {(im + 1.0 + 0.5) - px, px - (im + 0.5)};
    const double Y[2] = {(jm + 1.0 + 0.5) - py, py - (jm + 0.5)};
    const double Z[2] = {(km + 1.0 + 0.5) - pz, pz - (km + 0.5)};
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const double w = X[cdx(c)] * Y[cdy(c)] * Z[cdz(c)];
        double *d = dep + 3 * ((int64_t)c * M + t);
        d[0] = q[0] * w;
        d[1] = q[1] * w;
        d[2] = q[2] * w;
    }
    ord[t] = (uint32_t)q[6];
}

// One thread per node, 64 x 4 nodes per block (x, y), one z plane per grid
// layer: the lists of the eight cells with this node as a corner, merged by
// reference index.  Most nodes have no term or one.
constexpr int FBX = 64, FBY = 4;

template <typename T>
__global__ void __launch_bounds__(FBX * FBY) k_fb_gather(PartParams P, const uint32_t *__restrict__ keys,
                                                         const uint32_t *__restrict__ ord,
                                                         const int *__restrict__ head,
                                                         const double *__restrict__ dep, int64_t M, T *__restrict__ fb)
{
    const int i = blockIdx.x * FBX + threadIdx.x, j = blockIdx.y * FBY + threadIdx.y, k = blockIdx.z;
    if (i >= P.NX || j >= P.NY) return;
    const int64_t N = P.N;
    const int64_t n = ((int64_t)k * P.NY + j) * P.NX + i;
    // cursor into the sorted list of the cell whose corner c is this node
    uint32_t cell[8];
    int cur[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        cell[c] = fb_cell(P, i - cdx(c), j - cdy(c), k - cdz(c));
        cur[c] = head[cell[c]];
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (;;) {
        int nact = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) nact += cur[c] >= 0;
        if (nact == 0) break;
        // the next term in reference order: the smallest index at the cursors
        int bc = -1;
        if (nact == 1) {
#pragma unroll
            for (int c = 0; c < 8; c++)
                if (cur[c] >= 0) bc = c;
        } else {
            uint32_t best = 0xFFFFFFFFu;
#pragma unroll
            for (int c = 0; c < 8; c++)
                if (cur[c] >= 0) {
                    const uint32_t o = ord[cur[c]];
                    if (o < best) {
                        best = o;
                        bc = c;
                    }
                }
        }
        int t = 0;
        uint32_t cb = 0;
#pragma unroll
        for (int c = 0; c < 8; c++)
            if (c == bc) {
                t = cur[c];
                cb = cell[c];
            }
        const double *d = dep + 3 * ((int64_t)bc * M + t);
        const bool more = t + 1 < M && keys[t + 1] == cb;
        sx += d[0];
        sy += d[1];
        sz += d[2];
#pragma unroll
        for (int c = 0; c < 8; c++)
            if (c == bc) cur[c] = more ? t + 1 : -1;
    }
    fb[n] = (T)sx;
    fb[N + n] = (T)sy;
    fb[2 * N + n] = (T)sz;
}

// Cell key of each particle: the node of its wrapped position (z, y, x
// order); dropped particles sort last.
__global__ void __launch_bounds__(PB) k_cell_keys(PartParams P, uint32_t *key, uint32_t *val)
{
    const int64_t p = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (p >= P.np) return;
    uint32_t k = (uint32_t)P.N;
    if (P.alive[p]) {
        const double *S = P.s;
        const double x = wrapd(S[PX * P.ld + p], (double)P.NX), y = wrapd(S[PY * P.ld + p], (double)P.NY),
                     z = wrapd(S[PZ * P.ld + p], (double)P.NZ);
        if (x == x && y == y && z == z) {
            const int i = min(max((int)x, 0), P.NX - 1), j = min(max((int)y, 0), P.NY - 1),
                      l = min(max((int)z, 0), P.NZ - 1);
            k = (uint32_t)(((int64_t)l * P.NY + j) * P.NX + i);
        }
    }
    key[p] = k;
    val[p] = (uint32_t)p;
}

// Row f of every particle into its new place (row-major grid: y = row).
__global__ void __launch_bounds__(PB) k_permute(PartParams P, const uint32_t *src, double *s_out, int *alive_out,
                                                int *orig_out, PermRows rows)
{
    const int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (i >= P.np) return;
    const int64_t from = src[i];
    const int f = rows.r[blockIdx.y];
    s_out[(int64_t)f * P.ld + i] = P.s[(int64_t)f * P.ld + from];
    if (blockIdx.y == 0) {
        alive_out[i] = P.alive[from];
        orig_out[i] = P.orig[from];
    }
}

int blocks_of(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

int launch_particles(const PartParams &P, hipStream_t s)
{
    if (P.np == 0) return 0;
    const int nb = blocks_of(P.np, PB);
    const bool grad = (P.flags & CH4G_LAGRANGE_GRADIENT) != 0;
    // with gradients the interpolation runs as its own kernel and the
    // sub-step streams the interpolated rows (18.8 against 19.1-19.4 ms per
    // step of the 512^3 two-way configuration); without, one kernel (0.25
    // against 0.28 ms per step for the 2e6 particles of HIT 256^3)
    if (P.uT) {
        if (grad) {
            hipLaunchKernelGGL((k_interp<true, float>), dim3(nb), dim3(PB), 0, s, P);
            hipLaunchKernelGGL((k_particles<true, float, true>), dim3(nb), dim3(PB), 0, s, P);
        } else {
            hipLaunchKernelGGL((k_particles<false, float, false>), dim3(nb), dim3(PB), 0, s, P);
        }
    } else {
        if (grad) {
            hipLaunchKernelGGL((k_interp<true, double>), dim3(nb), dim3(PB), 0, s, P);
            hipLaunchKernelGGL((k_particles<true, double, true>), dim3(nb), dim3(PB), 0, s, P);
        } else {
            hipLaunchKernelGGL((k_particles<false, double, false>), dim3(nb), dim3(PB), 0, s, P);
        }
    }
    if (P.bsum) {
        hipLaunchKernelGGL(k_draw_offsets, dim3(1), dim3(1024), 0, s, P.bsum, nb, P.rng);
        hipLaunchKernelGGL(k_draw_apply, dim3(nb), dim3(PB), 0, s, P);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_particle_finit(const PartParams &P, bool velocity, bool set_v, hipStream_t s)
{
    if (P.np == 0) return 0;
    if (velocity ? !P.u : !P.t) return -1;
    PartParams Q = P;
    Q.fluid = velocity ? 1 : 0;
    Q.thermal = velocity ? 0 : 1;
    Q.uT = nullptr;
    const int nb = blocks_of(P.np, PB);
    const bool grad = (P.flags & CH4G_LAGRANGE_GRADIENT) != 0;
    const int sv = set_v ? 1 : 0;
    if (velocity && grad)
        hipLaunchKernelGGL((k_finit<true, true>), dim3(nb), dim3(PB), 0, s, Q, sv);
    else if (velocity)
        hipLaunchKernelGGL((k_finit<false, true>), dim3(nb), dim3(PB), 0, s, Q, sv);
    else if (grad)
        hipLaunchKernelGGL((k_finit<true, false>), dim3(nb), dim3(PB), 0, s, Q, sv);
    else
        hipLaunchKernelGGL((k_finit<false, false>), dim3(nb), dim3(PB), 0, s, Q, sv);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_reorder(const PartParams &P, const ReorderBufs &B, hipStream_t s)
{
    if (P.np == 0) return 0;
    const int nb = blocks_of(P.np, PB);
    hipLaunchKernelGGL(k_cell_keys, dim3(nb), dim3(PB), 0, s, P, B.kin, B.vin);
    size_t bytes = B.tmp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(B.tmp, bytes, B.kin, B.kout, B.vin, B.vout, (int)P.np, 0, B.key_bits, s) !=
        hipSuccess)
        return -2;
    if (B.rows.n < 1 || B.rows.n > PF_COUNT) return -1;
    hipLaunchKernelGGL(k_permute, dim3(nb, B.rows.n), dim3(PB), 0, s, P, (const uint32_t *)B.vout, B.s_out,
                       B.alive_out, B.orig_out, B.rows);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t feedback_sort_bytes(int64_t nitems, int key_bits)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)nitems, 0,
                                             key_bits);
    return bytes;
}

int launch_feedback(const PartParams &P, const FeedbackBufs &B, hipStream_t s)
{
    const uint32_t sent = B.key_bits >= 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << B.key_bits) - 1);
    if (B.old_n > 0)
        hipLaunchKernelGGL(k_fb_heads, dim3(blocks_of(B.old_n, PB)), dim3(PB), 0, s, B.old_keys, B.old_n, sent, B.head,
                           0);
    const int64_t M = P.np;
    if (M > 0) {
        hipLaunchKernelGGL(k_fb_cells, dim3(blocks_of(M, PB)), dim3(PB), 0, s, P, sent, B.kin, B.vin, B.rec, B.gv[0],
                           B.gv[1], B.gv[2]);
        size_t bytes = B.tmp_bytes;
        if (hipcub::DeviceRadixSort::SortPairs(B.tmp, bytes, B.kin, B.kout, B.vin, B.vout, (int)M, 0, B.key_bits,
                                               s) != hipSuccess)
            return -2;
        hipLaunchKernelGGL(k_fb_heads, dim3(blocks_of(M, PB)), dim3(PB), 0, s, (const uint32_t *)B.kout, M, sent,
                           B.head, 1);
        hipLaunchKernelGGL(k_fb_deps, dim3(blocks_of(M, PB)), dim3(PB), 0, s, P, (const uint32_t *)B.kout,
                           (const uint32_t *)B.vout, (const double *)B.rec, sent, M, B.dep, B.ord);
    }
    const dim3 grid(blocks_of(P.NX, FBX), blocks_of(P.NY, FBY), P.NZ), block(FBX, FBY);
    if (B.fb_float)
        hipLaunchKernelGGL(k_fb_gather<float>, grid, block, 0, s, P, (const uint32_t *)B.kout, (const uint32_t *)B.ord,
                           (const int *)B.head, (const double *)B.dep, M, (float *)B.fb);
    else
        hipLaunchKernelGGL(k_fb_gather<double>, grid, block, 0, s, P, (const uint32_t *)B.kout, (const uint32_t *)B.ord,
                           (const int *)B.head, (const double *)B.dep, M, (double *)B.fb);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace ch4g
