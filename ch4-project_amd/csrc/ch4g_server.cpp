// Serving of many lattices on one GPU (include/ch4g.h ch4g_step_batch,
// ch4g_serve): batched steps over a tile -> lattice table, and the
// continuous-batching request loop (admission by device footprint, chunks of
// steps replayed from captured graphs, retirement with the final averages).
//
// Compiled with -ffp-contract=off (host arithmetic included).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "ch4g_lattice.h"

using namespace ch4g;

namespace {

// A set of lattices stepped together, grouped by kernel variant; the device
// copy of the table is owned by the group set and reused across calls.
struct Batch {
    struct Group {
        int fluid, thermal, dev32, arith64;
        bool ntl;
        int first, count;   // range in the host table
        int64_t ntiles;
    };
    std::vector<BatchEntry> host;
    std::vector<Group> groups;
    BatchEntry *dev = nullptr;
    size_t dev_cap = 0;     // entries
    hipStream_t stream = nullptr;
    std::vector<ch4g_lattice *> lats;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int graph_parity0 = 0;
    hipEvent_t ev = nullptr;

    ~Batch() { release(); }
    void drop_graph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_steps = 0;
    }
    void release()
    {
        drop_graph();
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        dev_cap = 0;
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

int check_member(const ch4g_lattice *L)
{
    if (!L) return fail(-1, "null lattice in batch");
    if (L->first || L->itime < 1) return fail(-1, "batched lattices must be past their first step (ch4g_step(lat, 1))");
    if (L->base.hit) return fail(-1, "HIT-forced lattices need per-step host tables and are not batched");
    if (L->part) return fail(-1, "lattices with particles are not batched");
    return 0;
}

// Table of the lattices, grouped by variant, tiles numbered per group.
int build(Batch &B, const std::vector<ch4g_lattice *> &lats)
{
    B.drop_graph();
    B.lats = lats;
    B.host.clear();
    B.groups.clear();
    for (ch4g_lattice *L : lats) {
        int rc = check_member(L);
        if (rc) return rc;
    }
    std::vector<ch4g_lattice *> order = lats;
    auto key = [](const ch4g_lattice *L) {
        return (L->base.fluid << 3) | (L->base.thermal << 2) | (L->dev32 << 1) | L->base.arith64;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](const ch4g_lattice *a, const ch4g_lattice *b) { return key(a) < key(b); });
    for (size_t i = 0; i < order.size();) {
        Batch::Group G;
        const ch4g_lattice *L0 = order[i];
        G.fluid = L0->base.fluid;
        G.thermal = L0->base.thermal;
        G.dev32 = L0->dev32;
        G.arith64 = L0->base.arith64;
        G.first = (int)B.host.size();
        G.count = 0;
        G.ntiles = 0;
        int64_t bytes = 0;
        for (; i < order.size() && key(order[i]) == key(L0); i++) {
            ch4g_lattice *L = order[i];
            BatchEntry E;
            memset(&E, 0, sizeof E);
            E.P = params_for(L, 0, (int)(L->itime + 1));
            for (int b = 0; b < 2; b++) {
                E.f[b] = L->f[b];
                E.g[b] = L->g[b];
            }
            E.tlag = L->tlag;
            E.taules = L->taules;
            E.cur = L->cur;
            E.tile0 = G.ntiles;
            G.ntiles += (L->N + 255) / 256;
            bytes += 2 * (int64_t)Q * L->N * 4 * ((L->base.fluid ? 1 : 0) + (L->base.thermal ? 1 : 0));
            B.host.push_back(E);
            G.count++;
        }
        // non-temporal loads when the group's lattices fit the last-level cache
        G.ntl = bytes <= ((int64_t)192 << 20);
        B.groups.push_back(G);
    }
    if (!B.stream) HIPCHK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    if (!B.ev) HIPCHK(hipEventCreateWithFlags(&B.ev, hipEventDisableTiming));
    // the previous table may still be read by queued kernels
    HIPCHK(hipStreamSynchronize(B.stream));
    if (B.host.size() > B.dev_cap) {
        if (B.dev) (void)hipFree(B.dev);
        B.dev = nullptr;
        const size_t cap = std::max<size_t>(B.host.size(), 64);
        HIPCHK(hipMalloc((void **)&B.dev, cap * sizeof(BatchEntry)));
        B.dev_cap = cap;
    }
    if (!B.host.empty()) {
        HIPCHK(hipMemcpyAsync(B.dev, B.host.data(), B.host.size() * sizeof(BatchEntry), hipMemcpyHostToDevice,
                              B.stream));
        HIPCHK(hipStreamSynchronize(B.stream));
    }
    return 0;
}

void launch_steps(Batch &B, int nsteps, int parity0)
{
    for (int k = 0; k < nsteps; k++) {
        const int parity = (parity0 + k) & 1;
        for (const Batch::Group &G : B.groups) {
            const BatchEntry *tab = B.dev + G.first;
            if (G.dev32)
                launch_batch_fp32(tab, G.count, G.ntiles, G.fluid, G.thermal, G.arith64, G.ntl, parity, B.stream);
            else
                launch_batch_fp64(tab, G.count, G.ntiles, G.fluid, G.thermal, parity, B.stream);
        }
    }
}

// Advance every lattice of the table by nsteps (ordered after each lattice's
// own queued work, and before any work queued on it later).
int advance(Batch &B, int nsteps, int parity0, bool use_graph)
{
    if (B.lats.empty() || nsteps <= 0) return 0;
    for (ch4g_lattice *L : B.lats) {
        HIPCHK(hipEventRecord(B.ev, L->stream));
        HIPCHK(hipStreamWaitEvent(B.stream, B.ev, 0));
    }
    if (use_graph && nsteps % 2 == 0) {
        if (!B.graph || B.graph_steps != nsteps || B.graph_parity0 != parity0) {
            B.drop_graph();
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            launch_steps(B, nsteps, parity0);
            hipError_t e = hipStreamEndCapture(B.stream, &g);
            if (e != hipSuccess) return fail(-2, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                B.graph = nullptr;
                return fail(-2, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            }
            B.graph_steps = nsteps;
            B.graph_parity0 = parity0;
        }
        HIPCHK(hipGraphLaunch(B.graph, B.stream));
    } else {
        launch_steps(B, nsteps, parity0);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(B.ev, B.stream));
    for (ch4g_lattice *L : B.lats) HIPCHK(hipStreamWaitEvent(L->stream, B.ev, 0));
    return 0;
}

// The lattices' own state after nsteps batched steps.
void commit(Batch &B, int nsteps)
{
    for (ch4g_lattice *L : B.lats) {
        if (nsteps & 1) L->cur ^= 1;
        L->itime += nsteps;
    }
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// device bytes a job holds: populations and lags, the FP64 fields and
// partial sums of its final dump_averages, the per-row base tables
int64_t job_bytes(const ch4g_config &c)
{
    const int64_t N = (int64_t)c.NX * c.NY * c.NZ;
    return ch4g_footprint_bytes(&c) + N * 40 + ((int64_t)1 << 20);
}

}  // namespace

extern "C" {

int ch4g_step_batch(ch4g_lattice *const *lats, int n, int nsteps)
{
    if (n < 0 || nsteps < 0 || (n > 0 && !lats)) return fail(-1, "bad arguments");
    if (n == 0 || nsteps == 0) return 0;
    std::vector<ch4g_lattice *> v(lats, lats + n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++)
            if (v[i] == v[j]) return fail(-1, "a lattice appears twice in the batch");
    Batch B;
    int rc = build(B, v);
    if (rc) return rc;
    // parity 0 = each lattice's current buffer (E.cur)
    rc = advance(B, nsteps, 0, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(B.stream));
    commit(B, nsteps);
    return 0;
}

int ch4g_serve(const ch4g_serve_params *sp, ch4g_job *jobs, int32_t njobs)
{
    if (!sp || (njobs > 0 && !jobs) || njobs < 0) return fail(-1, "bad arguments");
    const int chunk = sp->chunk_steps > 0 ? sp->chunk_steps : 32;
    const bool batched = sp->mode == CH4G_SERVE_BATCHED;
    int64_t budget = sp->hbm_budget;
    if (budget <= 0) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        budget = (int64_t)((double)fr * 0.9);
    }
    // release order: arrival time, then index
    std::vector<int> order((size_t)njobs);
    for (int i = 0; i < njobs; i++) {
        order[(size_t)i] = i;
        jobs[i].status = 1; // pending
        jobs[i].admit_s = jobs[i].finish_s = -1.0;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return jobs[a].arrival_s < jobs[b].arrival_s; });

    struct Active {
        int job;
        ch4g_lattice *L;
        int64_t remaining;
        int64_t bytes;
    };
    std::vector<Active> active;
    Batch B;
    bool table_dirty = true;
    int64_t used = 0;
    size_t next = 0;
    int done = 0;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = 0;

    auto retire = [&](size_t a) -> int {
        Active &A = active[a];
        ch4g_job &J = jobs[A.job];
        int r = ch4g_dump_averages(A.L, &J.out);
        J.finish_s = seconds_since(t0);
        J.status = r;
        ch4g_destroy(A.L);
        used -= A.bytes;
        active.erase(active.begin() + (long)a);
        done++;
        table_dirty = true;
        return 0;
    };

    while (done < njobs) {
        // ---- admission, in arrival order, while the footprint fits
        const double now = seconds_since(t0);
        while (next < order.size() && jobs[order[next]].arrival_s <= now) {
            ch4g_job &J = jobs[order[next]];
            const int64_t nb = job_bytes(J.cfg);
            if (nb > budget) { // can never fit
                J.status = fail(-3, "job needs %lld bytes, budget %lld", (long long)nb, (long long)budget);
                next++;
                done++;
                continue;
            }
            if (!batched && !active.empty()) break;
            if (used + nb > budget || (sp->max_active > 0 && (int)active.size() >= sp->max_active)) break;
            ch4g_lattice *L = nullptr;
            int r = ch4g_create(&J.cfg, &L);
            if (!r) r = ch4g_initial_conditions(L);
            if (!r && batched && L->base.hit) r = fail(-1, "HIT-forced jobs are not served");
            if (!r && J.nsteps > 0) r = ch4g_step(L, 1); // first step: reads the initial populations
            if (r) {
                if (L) ch4g_destroy(L);
                J.status = r;
                next++;
                done++;
                continue;
            }
            J.admit_s = seconds_since(t0);
            active.push_back({order[next], L, J.nsteps > 0 ? J.nsteps - 1 : 0, nb});
            used += nb;
            next++;
            table_dirty = true;
        }
        // jobs with nothing (left) to run retire at once
        for (size_t a = 0; a < active.size();) {
            if (active[a].remaining == 0)
                retire(a);
            else
                a++;
        }
        if (active.empty()) {
            if (next >= order.size()) break;
            const double wait = jobs[order[next]].arrival_s - seconds_since(t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            continue;
        }
        // ---- one chunk for every active lattice (no job overshoots); the
        // plain mode runs its single job to the end in one call
        int64_t m = batched ? chunk : active[0].remaining;
        for (const Active &A : active) m = std::min(m, A.remaining);
        if (batched) {
            if (table_dirty) {
                std::vector<ch4g_lattice *> v;
                for (const Active &A : active) v.push_back(A.L);
                if ((rc = build(B, v))) break;
                table_dirty = false;
            }
            if ((rc = advance(B, (int)m, 0, sp->use_graphs != 0))) break;
            commit(B, (int)m);
            // the table holds each lattice's current buffer: after an odd
            // chunk it is rebuilt
            table_dirty = (m & 1) != 0;
            HIPCHK(hipStreamSynchronize(B.stream));
        } else {
            if ((rc = ch4g_step(active[0].L, (int)m))) break;
            if ((rc = ch4g_synchronize(active[0].L))) break;
        }
        for (Active &A : active) A.remaining -= m;
        for (size_t a = 0; a < active.size();) {
            if (active[a].remaining == 0)
                retire(a);
            else
                a++;
        }
    }
    for (Active &A : active) ch4g_destroy(A.L);
    for (int i = 0; i < njobs; i++)
        if (jobs[i].status == 1) jobs[i].status = rc ? rc : -1;
    return rc;
}

}  // extern "C"
