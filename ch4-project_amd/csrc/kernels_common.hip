// Initialisation, population load, and the FP64 reductions of dump_averages.
// Compiled with -ffp-contract=off: the FP64 init arithmetic (g = w t,
// p = w, t = sum g) must equal the reference's bit for bit.
#include "ch4g_internal.h"

namespace ch4g {

__device__ __forceinline__ int wrapc(int a, int L) { return a < 0 ? a + L : (a >= L ? a - L : a); }

__device__ __forceinline__ double msum19(const double *p)
{
    return p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8] + p[9] + p[10] + p[11] + p[12] + p[13] +
           p[14] + p[15] + p[16] + p[17] + p[18];
}

// initial_conditions on the device from a per-row fluid profile prow[y][19]
// and a temperature field t0[n] built on the host (initial_conditions.c:45,
// :72-84, :256-279, :303-308, :421-424, :486-487); tlag <- sum g (the t of
// the first hydro_fields call, main.c:29-31).
template <typename R>
__global__ void k_init(KParams P, const double *__restrict__ prow, const double *__restrict__ t0,
                       R *__restrict__ f, R *__restrict__ g, R *__restrict__ tlag, int dev)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    const int y = (int)((n / P.NX) % P.NY);
    // FP32-deviation: subtract the per-row base state (rho_b, T_b)
    const double rb = dev ? P.row64[2 * y] : 0.0, tb = dev ? P.row64[2 * y + 1] : 0.0;
    if (P.fluid) {
        for (int q = 0; q < Q; q++) {
            double v = prow[y * Q + q];
            f[q * P.S + n] = dev ? (R)(v - wq(q) * rb) : (R)v;
        }
    }
    if (P.thermal) {
        double gq[Q];
        double t = t0[n];
        for (int q = 0; q < Q; q++) gq[q] = wq(q) * t;
        for (int q = 0; q < Q; q++) g[q * P.S + n] = dev ? (R)(gq[q] - wq(q) * tb) : (R)gq[q];
        double T = msum19(gq);
        tlag[n] = dev ? (R)(T - tb) : (R)T;
    }
}

// Load populations from FP64 SoA staging (restart), tlag <- sum g.
template <typename R>
__global__ void k_load(KParams P, const double *__restrict__ fs, const double *__restrict__ gs, R *__restrict__ f,
                       R *__restrict__ g, R *__restrict__ tlag, int dev)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.N) return;
    const int y = (int)((n / P.NX) % P.NY);
    const double rb = dev ? P.row64[2 * y] : 0.0, tb = dev ? P.row64[2 * y + 1] : 0.0;
    if (P.fluid && fs)
        for (int q = 0; q < Q; q++) {
            double v = fs[q * P.N + n];
            f[q * P.S + n] = dev ? (R)(v - wq(q) * rb) : (R)v;
        }
    if (P.thermal && gs) {
        double gq[Q];
        for (int q = 0; q < Q; q++) gq[q] = gs[q * P.N + n];
        for (int q = 0; q < Q; q++) g[q * P.S + n] = dev ? (R)(gq[q] - wq(q) * tb) : (R)gq[q];
        double T = msum19(gq);
        tlag[n] = dev ? (R)(T - tb) : (R)T;
    }
}

void launch_init(const KParams &P, const double *prow, const double *t0, void *f, void *g, void *tlag, int dev,
                 hipStream_t s)
{
    int64_t nb = (P.N + 255) / 256;
    if (dev)
        hipLaunchKernelGGL(k_init<float>, dim3((unsigned)nb), dim3(256), 0, s, P, prow, t0, (float *)f, (float *)g,
                           (float *)tlag, 1);
    else
        hipLaunchKernelGGL(k_init<double>, dim3((unsigned)nb), dim3(256), 0, s, P, prow, t0, (double *)f,
                           (double *)g, (double *)tlag, 0);
}

void launch_load(const KParams &P, const double *fs, const double *gs, void *f, void *g, void *tlag, int dev,
                 hipStream_t s)
{
    int64_t nb = (P.N + 255) / 256;
    if (dev)
        hipLaunchKernelGGL(k_load<float>, dim3((unsigned)nb), dim3(256), 0, s, P, fs, gs, (float *)f, (float *)g,
                           (float *)tlag, 1);
    else
        hipLaunchKernelGGL(k_load<double>, dim3((unsigned)nb), dim3(256), 0, s, P, fs, gs, (double *)f,
                           (double *)g, (double *)tlag, 0);
}

// ---------------------------------------------------------------------------
// dump_averages (output.c:183-491): per-node FP64 contributions, centred
// differences with periodic wrap, one-sided second order at y walls
// (gradients.c:78-499, :505-626), block partials in LDS tree order, then one
// block sums the partials in block order.  Output order = ch4g_output.

constexpr int RED_THREADS = 256;

int reduce_blocks(const KParams &P)
{
    int64_t nb = (P.N + RED_THREADS - 1) / RED_THREADS;
    return (int)(nb < 2048 ? nb : 2048);
}

// Per-node terms of dump_averages (output.c:183-491) at node n = (x,y,z),
// added into a[] in the order of ch4g_output.  u is [N][3], fields of the
// current step; periodic images, one-sided differences at y walls.
__device__ __forceinline__ void node_terms(const KParams &P, const double *__restrict__ dens,
                                           const double *__restrict__ u, const double *__restrict__ t, double nu,
                                           double kappa, int64_t n, int x, int y, int z, double a[NOUT])
{
    const int NX = P.NX, NY = P.NY, NZ = P.NZ;
    auto id = [&](int i, int j, int k) -> int64_t { return ((int64_t)k * NY + j) * NX + i; };
    const int xp = wrapc(x + 1, NX), xm = wrapc(x - 1, NX);
    const int yp = wrapc(y + 1, NY), ym = wrapc(y - 1, NY);
    const int zp = wrapc(z + 1, NZ), zm = wrapc(z - 1, NZ);
    if (P.fluid) {
        const double ux = u[3 * n], uy = u[3 * n + 1], uz = u[3 * n + 2];
        a[0] += (double)x + 0.5;
        a[1] += (double)y + 0.5;
        a[2] += (double)z + 0.5;
        a[3] += ux;
        a[4] += uy;
        a[5] += uz;
        a[6] += ux * ux;
        a[7] += uy * uy;
        a[8] += uz * uz;
        a[9] += dens[n];
        a[10] += 0.5 * (ux * ux + uy * uy + uz * uz);
        a[12] += dens[n] * dens[n];
        double G[3][3];  // G[c][d] = d u_c / d x_d
        for (int c = 0; c < 3; c++) {
            G[c][0] = (u[3 * id(xp, y, z) + c] - u[3 * id(xm, y, z) + c]) / 2.0;
            G[c][1] = (u[3 * id(x, yp, z) + c] - u[3 * id(x, ym, z) + c]) / 2.0;
            G[c][2] = (u[3 * id(x, y, zp) + c] - u[3 * id(x, y, zm) + c]) / 2.0;
            if (P.wall_f && NY >= 3) {
                if (y == 0)
                    G[c][1] = -1.5 * u[3 * n + c] + 2.0 * u[3 * id(x, 1, z) + c] - 0.5 * u[3 * id(x, 2, z) + c];
                if (y == NY - 1)
                    G[c][1] = 1.5 * u[3 * n + c] - 2.0 * u[3 * id(x, NY - 2, z) + c] +
                              0.5 * u[3 * id(x, NY - 3, z) + c];
            }
            if (NX == 1) G[c][0] = 0.0;
            if (NY == 1) G[c][1] = 0.0;
            if (NZ == 1) G[c][2] = 0.0;
        }
        double e = 0.0;
        for (int c = 0; c < 3; c++)
            for (int d = 0; d < 3; d++) {
                double s = G[c][d] + G[d][c];
                e += s * s;
            }
        a[11] += e * 0.5 * nu;
    }
    if (P.thermal) {
        const double tn = t[n];
        double gx = (t[id(xp, y, z)] - t[id(xm, y, z)]) / 2.0;
        double gy = (t[id(x, yp, z)] - t[id(x, ym, z)]) / 2.0;
        double gz = (t[id(x, y, zp)] - t[id(x, y, zm)]) / 2.0;
        if (P.wall_t && NY >= 3) {
            if (y == 0) gy = -1.5 * tn + 2.0 * t[id(x, 1, z)] - 0.5 * t[id(x, 2, z)];
            if (y == NY - 1) gy = 1.5 * tn - 2.0 * t[id(x, NY - 2, z)] + 0.5 * t[id(x, NY - 3, z)];
        }
        double ux = 0.0, uy = 0.0, uz = 0.0;
        if (P.fluid) {
            ux = u[3 * n];
            uy = u[3 * n + 1];
            uz = u[3 * n + 2];
        }
        a[13] += tn;
        a[14] += tn * tn;
        a[15] += kappa * (gx * gx + gy * gy + gz * gz);
        a[16] += gx;
        a[17] += gy;
        a[18] += gz;
        a[19] += ux * tn;
        a[20] += uy * tn;
        a[21] += uz * tn;
        a[22] += -kappa * gx + ux * tn;
        a[23] += -kappa * gy + uy * tn;
        a[24] += -kappa * gz + uz * tn;
    }
}

__global__ void __launch_bounds__(RED_THREADS) k_averages(KParams P, const double *__restrict__ dens,
                                                          const double *__restrict__ u,
                                                          const double *__restrict__ t, double nu, double kappa,
                                                          double *__restrict__ partials)
{
    double a[NOUT];
    for (int k = 0; k < NOUT; k++) a[k] = 0.0;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < P.N;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(n % P.NX);
        const int y = (int)((n / P.NX) % P.NY);
        const int z = (int)(n / ((int64_t)P.NX * P.NY));
        node_terms(P, dens, u, t, nu, kappa, n, x, y, z, a);
    }
    __shared__ double sh[RED_THREADS];
    for (int k = 0; k < NOUT; k++) {
        sh[threadIdx.x] = a[k];
        __syncthreads();
        for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * NOUT + k] = sh[0];
        __syncthreads();
    }
}

__global__ void k_final(const double *__restrict__ partials, int nb, double *__restrict__ out)
{
    const int k = threadIdx.x;
    if (k >= NOUT) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += partials[(int64_t)b * NOUT + k];
    out[k] = s;
}

// dump_averages rulers (output.c:212-491): block p sums the per-node terms of
// plane p along AXIS (0: x = p, 1: y = p, 2: z = p) over its nodes, LDS tree
// order, into out[p][NOUT].
template <int AXIS>
__global__ void __launch_bounds__(RED_THREADS) k_rulers(KParams P, const double *__restrict__ dens,
                                                        const double *__restrict__ u, const double *__restrict__ t,
                                                        double nu, double kappa, double *__restrict__ out)
{
    double a[NOUT];
    for (int k = 0; k < NOUT; k++) a[k] = 0.0;
    const int p = blockIdx.x;
    const int NX = P.NX, NY = P.NY, NZ = P.NZ;
    const int64_t M = AXIS == 0 ? (int64_t)NY * NZ : (AXIS == 1 ? (int64_t)NX * NZ : (int64_t)NX * NY);
    for (int64_t m = threadIdx.x; m < M; m += RED_THREADS) {
        int x, y, z;
        if (AXIS == 0) {
            x = p;
            y = (int)(m % NY);
            z = (int)(m / NY);
        } else if (AXIS == 1) {
            x = (int)(m % NX);
            y = p;
            z = (int)(m / NX);
        } else {
            x = (int)(m % NX);
            y = (int)(m / NX);
            z = p;
        }
        const int64_t n = ((int64_t)z * NY + y) * NX + x;
        node_terms(P, dens, u, t, nu, kappa, n, x, y, z, a);
    }
    __shared__ double sh[RED_THREADS];
    for (int k = 0; k < NOUT; k++) {
        sh[threadIdx.x] = a[k];
        __syncthreads();
        for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(int64_t)p * NOUT + k] = sh[0];
        __syncthreads();
    }
}

void launch_rulers(const KParams &P, const double *dens, const double *u, const double *t, double nu, double kappa,
                   double *rx, double *ry, double *rz, hipStream_t s)
{
    if (rx) hipLaunchKernelGGL(k_rulers<0>, dim3(P.NX), dim3(RED_THREADS), 0, s, P, dens, u, t, nu, kappa, rx);
    if (ry) hipLaunchKernelGGL(k_rulers<1>, dim3(P.NY), dim3(RED_THREADS), 0, s, P, dens, u, t, nu, kappa, ry);
    if (rz) hipLaunchKernelGGL(k_rulers<2>, dim3(P.NZ), dim3(RED_THREADS), 0, s, P, dens, u, t, nu, kappa, rz);
}

void launch_averages(const KParams &P, const double *dens, const double *u, const double *t, double nu, double kappa,
                     double *partials, double *result, hipStream_t s)
{
    int nb = reduce_blocks(P);
    hipLaunchKernelGGL(k_averages, dim3(nb), dim3(RED_THREADS), 0, s, P, dens, u, t, nu, kappa, partials);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, s, partials, nb, result);
}

}  // namespace ch4g
