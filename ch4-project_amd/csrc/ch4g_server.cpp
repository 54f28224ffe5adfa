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

// A set of lattices stepped together.  A chunk is `chunk` launches per
// storage kind; lattice i takes steps[i] of them, spread evenly over the
// chunk, so every launch carries a similar share of the work.  Each launch
// has its own short slot list (lattice, parity of its step, first block).
// 2-D float-storage lattices (two_step_eligible) form groups of their own
// whose launches advance two steps each: steps[i] takes steps[i] / 2 such
// launches and, for an odd share, one single step at the end.
struct Batch {
    struct Group {
        int dev32, arith64, kind;
        bool dense;                       // lanes decode their own rows (dense_rows)
        bool gen;                         // LES lattices (the kernel instance with those paths)
        bool ntl;
        bool two;                         // two steps per launch (k_step2_batch32)
        int first, count;                 // range in the entry table
        int nlaunch;                      // launches of the group per chunk
        std::vector<int64_t> off;         // per launch: first slot in the slot table
        std::vector<int> nslot;           // per launch: slots (without the sentinel)
        std::vector<int64_t> nblk;        // per launch: wave tiles (64 x 4 tiles for two-step groups)
    };
    std::vector<BatchEntry> host;
    std::vector<BatchSlot> slots;
    std::vector<Group> groups;
    BatchEntry *dev = nullptr;
    size_t dev_cap = 0;       // entries
    BatchSlot *dev_slots = nullptr;
    size_t slots_cap = 0;
    hipStream_t stream = nullptr;
    // one stream per group after the first: the groups' launches are
    // independent (different lattices) and overlap, each group's tail
    // filled by the others' blocks
    std::vector<hipStream_t> gstream;
    std::vector<hipEvent_t> gdone;
    hipEvent_t fork = nullptr;
    std::vector<ch4g_lattice *> lats;
    std::vector<int> steps;   // steps of each lattice per chunk
    std::vector<char> two;    // lattice i goes two steps per launch
    std::vector<char> flips;  // a chunk changes lattice i's buffer or lag parity
    // 3-D lattices large enough for two steps per launch alone run their own
    // launches (k_step2z_fp32) as a branch of the chunk on a stream of their
    // own instead of joining a batched group
    std::vector<int> solo;    // lattice indices
    std::vector<hipStream_t> sstream;
    std::vector<hipEvent_t> sdone;
    int chunk = 0;            // launches per chunk = max steps
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev = nullptr;

    ~Batch() { release(); }
    void drop_graph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    void release()
    {
        drop_graph();
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        dev_cap = 0;
        if (dev_slots) (void)hipFree(dev_slots);
        dev_slots = nullptr;
        slots_cap = 0;
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
        if (fork) (void)hipEventDestroy(fork);
        fork = nullptr;
        for (hipEvent_t e : gdone) (void)hipEventDestroy(e);
        gdone.clear();
        for (hipStream_t t : gstream) (void)hipStreamDestroy(t);
        gstream.clear();
        for (hipEvent_t e : sdone) (void)hipEventDestroy(e);
        sdone.clear();
        for (hipStream_t t : sstream) (void)hipStreamDestroy(t);
        sstream.clear();
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
    if (L->base.les_fd) return fail(-1, "finite-difference LES lattices need a velocity pass and are not batched");
    if (L->d_f0) return fail(-1, "NOZEROMODE lattices need a force-mean pass and are not batched");
    return 0;
}

// population sets of a lattice (BATCH_FLUID, BATCH_THERMAL or both)
int kind_of(const ch4g_lattice *L) { return (L->base.fluid ? BATCH_FLUID : 0) | (L->base.thermal ? BATCH_THERMAL : 0); }

// Entry and block tables of the lattices, grouped by storage kind and
// population sets (one kernel instance per group).
int build(Batch &B, const std::vector<ch4g_lattice *> &lats, const std::vector<int> &steps)
{
    B.drop_graph();
    B.lats = lats;
    B.steps = steps;
    B.host.clear();
    B.slots.clear();
    B.groups.clear();
    B.chunk = 0;
    B.two.assign(lats.size(), 0);
    B.flips.assign(lats.size(), 0);
    B.solo.clear();
    std::vector<char> solo(lats.size(), 0);
    for (size_t i = 0; i < lats.size(); i++) {
        int rc = check_member(lats[i]);
        if (rc) return rc;
        if (steps[i] < 1) return fail(-1, "steps per chunk must be >= 1");
        B.chunk = std::max(B.chunk, steps[i]);
        ch4g_lattice *L = lats[i];
        // two steps per launch: 2-D lattices in the batched two-step kernel,
        // large 3-D lattices on their own branch
        B.two[i] = two_step_eligible(L);
        solo[i] = B.two[i] && L->cfg.NZ > 1;
        if (solo[i]) B.solo.push_back((int)i);
        if (B.two[i] && (rc = ensure_tlag_alt(L))) return rc;
        // two-step lattices: the population buffer flips with every launch,
        // the lag array with every two-step launch
        const int pairs = steps[i] / 2, launches = pairs + steps[i] % 2;
        B.flips[i] = B.two[i] ? ((launches & 1) || (pairs & 1)) : (steps[i] & 1);
    }
    std::vector<int> order;
    for (size_t i = 0; i < lats.size(); i++)
        if (!solo[i]) order.push_back((int)i);
    auto key = [&](int i) {
        const ch4g_lattice *L = lats[(size_t)i];
        const int two = B.two[(size_t)i];
        return (L->base.les << 6) | (two << 5) | (L->dev32 << 4) | (L->base.arith64 << 3) |
               ((int)(!two && dense_rows(L->cfg.NX)) << 2) | kind_of(L);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
    for (size_t i = 0; i < order.size();) {
        Batch::Group G;
        const int k0 = key(order[i]);
        const ch4g_lattice *L0 = lats[(size_t)order[i]];
        G.dev32 = L0->dev32;
        G.arith64 = L0->base.arith64;
        G.kind = kind_of(L0);
        G.two = B.two[(size_t)order[i]] != 0;
        G.dense = !G.two && dense_rows(L0->cfg.NX);
        G.gen = L0->base.les != 0;
        G.first = (int)B.host.size();
        G.count = 0;
        int64_t bytes = 0;
        // members of the group and their launches per chunk (two-step
        // groups: pairs, then the single step of an odd share)
        size_t i1 = i;
        while (i1 < order.size() && key(order[i1]) == k0) i1++;
        int KG = 0;
        for (size_t j = i; j < i1; j++) {
            const int st = steps[(size_t)order[j]];
            KG = std::max(KG, G.two ? st / 2 + st % 2 : st);
        }
        G.nlaunch = KG;
        // launch index of each member at each launch (-1: not in that launch)
        std::vector<std::vector<int>> at;
        std::vector<int> npair;
        std::vector<int64_t> ntile;
        for (; i < i1; i++) {
            ch4g_lattice *L = lats[(size_t)order[i]];
            const int st = steps[(size_t)order[i]];
            BatchEntry E;
            memset(&E, 0, sizeof E);
            E.P = params_for(L, 0, (int)(L->itime + 1));
            for (int b = 0; b < 2; b++) {
                E.f[b] = L->f[b];
                E.g[b] = L->g[b];
            }
            E.tlag = L->tlag;
            E.tlag2 = G.two ? L->tlag_alt : nullptr;
            E.taules = L->taules;
            E.cur = L->cur;
            B.host.push_back(E);
            const int nl = G.two ? st / 2 + st % 2 : st;
            std::vector<int> a((size_t)KG, -1);
            for (int s_ = 0; s_ < nl; s_++) a[(size_t)(((int64_t)s_ * KG) / nl)] = s_;
            at.push_back(a);
            npair.push_back(G.two ? st / 2 : 0);
            ntile.push_back(G.two ? batch2_tiles(L->cfg.NX, L->cfg.NY) : batch_tiles(L->cfg.NX, L->cfg.NY, L->cfg.NZ));
            bytes += 2 * (int64_t)Q * L->N * (L->dev32 ? 4 : 8) * ((L->base.fluid ? 1 : 0) + (L->base.thermal ? 1 : 0));
            G.count++;
        }
        G.off.assign((size_t)KG, 0);
        G.nslot.assign((size_t)KG, 0);
        G.nblk.assign((size_t)KG, 0);
        for (int k = 0; k < KG; k++) {
            G.off[(size_t)k] = (int64_t)B.slots.size();
            int64_t t0 = 0;
            for (int e = 0; e < G.count; e++) {
                const int s_ = at[(size_t)e][(size_t)k];
                if (s_ < 0) continue;
                uint32_t ent = (uint32_t)e | ((uint32_t)(s_ & 1) << 31);
                if (G.two) {
                    const int np = npair[(size_t)e];
                    if (s_ >= np) ent |= BATCH_SINGLE;
                    if (std::min(s_, np) & 1) ent |= BATCH_LAG_PARITY;
                }
                B.slots.push_back({ent, (uint32_t)t0});
                t0 += ntile[(size_t)e];
            }
            G.nslot[(size_t)k] = (int)((int64_t)B.slots.size() - G.off[(size_t)k]);
            G.nblk[(size_t)k] = t0;
            B.slots.push_back({0u, (uint32_t)t0}); // sentinel
        }
        // non-temporal loads when the group's lattices fit the last-level cache
        G.ntl = bytes <= ((int64_t)192 << 20);
        B.groups.push_back(G);
    }
    if (!B.stream) HIPCHK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    if (!B.ev) HIPCHK(hipEventCreateWithFlags(&B.ev, hipEventDisableTiming));
    if (!B.fork) HIPCHK(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
    while (B.gstream.size() + 1 < B.groups.size()) {
        hipStream_t t = nullptr;
        hipEvent_t e = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        B.gstream.push_back(t);
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        B.gdone.push_back(e);
    }
    while (B.sstream.size() < B.solo.size()) {
        hipStream_t t = nullptr;
        hipEvent_t e = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        B.sstream.push_back(t);
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        B.sdone.push_back(e);
    }
    // the previous tables may still be read by queued kernels
    HIPCHK(hipStreamSynchronize(B.stream));
    if (B.host.size() > B.dev_cap) {
        if (B.dev) (void)hipFree(B.dev);
        B.dev = nullptr;
        const size_t cap = std::max<size_t>(B.host.size(), 64);
        HIPCHK(hipMalloc((void **)&B.dev, cap * sizeof(BatchEntry)));
        B.dev_cap = cap;
    }
    if (B.slots.size() > B.slots_cap) {
        if (B.dev_slots) (void)hipFree(B.dev_slots);
        B.dev_slots = nullptr;
        const size_t cap = std::max<size_t>(B.slots.size() + B.slots.size() / 2, 1024);
        HIPCHK(hipMalloc((void **)&B.dev_slots, cap * sizeof(BatchSlot)));
        B.slots_cap = cap;
    }
    if (!B.host.empty()) {
        HIPCHK(hipMemcpyAsync(B.dev, B.host.data(), B.host.size() * sizeof(BatchEntry), hipMemcpyHostToDevice,
                              B.stream));
        HIPCHK(hipMemcpyAsync(B.dev_slots, B.slots.data(), B.slots.size() * sizeof(BatchSlot),
                              hipMemcpyHostToDevice, B.stream));
        HIPCHK(hipStreamSynchronize(B.stream));
    }
    return 0;
}

// The chunk's launches: group 0 on the batch stream, group g > 0 on its own
// stream forked from and joined back into the batch stream (under capture
// this makes parallel branches of one graph).
int launch_steps(Batch &B)
{
    const size_t ng = B.groups.size(), ns = B.solo.size();
    if (ng > 1 || ns > 0) {
        HIPCHK(hipEventRecord(B.fork, B.stream));
        for (size_t g = 1; g < ng; g++) HIPCHK(hipStreamWaitEvent(B.gstream[g - 1], B.fork, 0));
        for (size_t k = 0; k < ns; k++) HIPCHK(hipStreamWaitEvent(B.sstream[k], B.fork, 0));
    }
    // solo 3-D lattices: their share as two-step launches (plus one single
    // step for an odd share) from the buffers and lag arrays of the chunk's
    // start, exactly as ch4g_step advances them
    for (size_t k = 0; k < ns; k++) {
        ch4g_lattice *L = B.lats[(size_t)B.solo[k]];
        const int st = B.steps[(size_t)B.solo[k]];
        hipStream_t t = B.sstream[k];
        const KParams P = params_for(L, 0, (int)L->itime + 1);
        int cur = L->cur;
        void *tl = L->tlag, *ta = L->tlag_alt;
        for (int s_ = 0; s_ + 1 < st; s_ += 2) {
            launch_two_steps_fp32(P, (const float *)L->f[cur], (const float *)L->g[cur], (float *)L->f[cur ^ 1],
                                  (float *)L->g[cur ^ 1], (const float *)tl, (float *)ta, t);
            cur ^= 1;
            if (L->base.thermal) std::swap(tl, ta);
        }
        if (st & 1)
            launch_step_fp32(P, (const float *)L->f[cur], (const float *)L->g[cur], (float *)L->f[cur ^ 1],
                             (float *)L->g[cur ^ 1], (float *)tl, (float *)L->taules, t);
        HIPCHK(hipEventRecord(B.sdone[k], t));
    }
    for (size_t g = 0; g < ng; g++) {
        const Batch::Group &G = B.groups[g];
        hipStream_t st = g == 0 ? B.stream : B.gstream[g - 1];
        for (int k = 0; k < G.nlaunch; k++) {
            const int64_t nt = G.nblk[(size_t)k];
            if (nt == 0) continue;
            const int64_t nb = (nt + BATCH_WAVES_PER_BLOCK - 1) / BATCH_WAVES_PER_BLOCK;
            const BatchEntry *tab = B.dev + G.first;
            const BatchSlot *sl = B.dev_slots + G.off[(size_t)k];
            const int ns = G.nslot[(size_t)k];
            if (G.two)
                launch_batch2_fp32(tab, sl, ns, nt, G.kind, G.arith64, G.ntl, st);
            else if (G.dev32)
                launch_batch_fp32(tab, sl, ns, nb, G.kind, G.dense, G.gen, G.arith64, G.ntl, st);
            else
                launch_batch_fp64(tab, sl, ns, nb, G.kind, G.dense, st);
        }
    }
    for (size_t g = 1; g < ng; g++) {
        HIPCHK(hipEventRecord(B.gdone[g - 1], B.gstream[g - 1]));
        HIPCHK(hipStreamWaitEvent(B.stream, B.gdone[g - 1], 0));
    }
    for (size_t k = 0; k < ns; k++) HIPCHK(hipStreamWaitEvent(B.stream, B.sdone[k], 0));
    return 0;
}

// Advance every lattice of the table by its steps of one chunk (ordered
// after each lattice's own queued work, and before any work queued on it
// later); the launches are replayed from a captured graph when asked.
int advance(Batch &B, bool use_graph)
{
    if (B.lats.empty() || B.chunk <= 0) return 0;
    for (ch4g_lattice *L : B.lats) {
        HIPCHK(hipEventRecord(B.ev, L->stream));
        HIPCHK(hipStreamWaitEvent(B.stream, B.ev, 0));
    }
    if (use_graph) {
        if (!B.graph) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            int lrc = launch_steps(B);
            hipError_t e = hipStreamEndCapture(B.stream, &g);
            if (lrc) {
                if (g) (void)hipGraphDestroy(g);
                return lrc;
            }
            if (e != hipSuccess) return fail(-2, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                B.graph = nullptr;
                return fail(-2, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            }
        }
        HIPCHK(hipGraphLaunch(B.graph, B.stream));
    } else {
        int lrc = launch_steps(B);
        if (lrc) return lrc;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(B.ev, B.stream));
    for (ch4g_lattice *L : B.lats) HIPCHK(hipStreamWaitEvent(L->stream, B.ev, 0));
    return 0;
}

// The lattices' own state after a chunk.
void commit(Batch &B)
{
    for (size_t i = 0; i < B.lats.size(); i++) {
        ch4g_lattice *L = B.lats[i];
        const int st = B.steps[i];
        if (B.two[i]) {
            const int pairs = st / 2;
            if ((pairs + st % 2) & 1) L->cur ^= 1;
            if ((pairs & 1) && L->base.thermal) {
                std::swap(L->tlag, L->tlag_alt);
                L->tpar ^= 1;
            }
        } else if (st & 1) {
            L->cur ^= 1;
        }
        L->itime += st;
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
    // the batch context (stream, device tables) is reused by later calls
    // from the same thread
    thread_local Batch B;
    int rc = build(B, v, std::vector<int>((size_t)n, nsteps));
    if (rc) return rc;
    rc = advance(B, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(B.stream));
    commit(B);
    return 0;
}

int ch4g_serve(const ch4g_serve_params *sp, ch4g_job *jobs, int32_t njobs)
{
    if (!sp || (njobs > 0 && !jobs) || njobs < 0) return fail(-1, "bad arguments");
    // 128 launches between admission points: the mixed ensemble ran 23.05
    // (32), 23.60 (128), 23.72 (512) GLUPS batched against 23.33 one job at a
    // time, with p50 latency unchanged (2.6-2.7 s)
    const int chunk = sp->chunk_steps > 0 ? sp->chunk_steps : 128;
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
    int chunks_since_share = 0;
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

    // ---- admission, in arrival order, while the footprint fits
    auto admit_ready = [&]() {
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
            if (!r && batched && (L->base.hit || L->base.les_fd || L->d_f0))
                r = fail(-1, "HIT-forced, finite-difference LES and NOZEROMODE jobs are not batched (serve them "
                             "with CH4G_SERVE_SEQUENTIAL)");
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
    };

    while (done < njobs) {
        admit_ready();
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
        // ---- one chunk.  Batched: every resident job takes part, with all
        // the chunk's steps (lock-step) or a share proportional to its
        // remaining steps, spread over the chunk's launches (even shares keep
        // each lattice's buffer parity, so the tables and their captured
        // graph stay valid until the set changes; a job that can finish
        // does); shares are refreshed every 16 chunks.  Jobs that arrived
        // while the chunk ran are set up after it and join the next one.
        // Plain: the single job runs to its end in one call.
        if (batched) {
            if (table_dirty || ++chunks_since_share >= 16) {
                int64_t rmax = 1;
                for (const Active &A : active) rmax = std::max(rmax, A.remaining);
                std::vector<ch4g_lattice *> v;
                std::vector<int> st;
                for (const Active &A : active) {
                    int64_t s_ = chunk;
                    if (sp->schedule == CH4G_SCHEDULE_PROPORTIONAL) {
                        s_ = (chunk * A.remaining + rmax - 1) / rmax;
                        if (s_ & 1) s_ += 1;
                    }
                    s_ = std::max<int64_t>(1, std::min<int64_t>(s_, A.remaining));
                    v.push_back(A.L);
                    st.push_back((int)s_);
                }
                if ((rc = build(B, v, st))) break;
                table_dirty = false;
                chunks_since_share = 0;
            }
            if ((rc = advance(B, sp->use_graphs != 0))) break;
            const size_t in_chunk = B.lats.size(); // active[0 .. in_chunk) take part
            // The chunk (a graph whose branches run on several streams)
            // completes before any other runtime call: admission allocates,
            // frees and launches, and doing that while a chunk was in flight
            // ended the process in the runtime now and then.
            HIPCHK(hipStreamSynchronize(B.stream));
            commit(B);
            admit_ready();
            for (size_t a = 0; a < in_chunk; a++) {
                active[a].remaining -= B.steps[a];
                // a share that flips the lattice's buffer (or lag array): the table follows
                if (B.flips[a]) table_dirty = true;
                if (active[a].remaining < B.steps[a]) table_dirty = true;
            }
        } else {
            const int64_t m = active[0].remaining;
            if ((rc = ch4g_step(active[0].L, (int)m))) break;
            if ((rc = ch4g_synchronize(active[0].L))) break;
            active[0].remaining -= m;
        }
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
