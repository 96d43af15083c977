// zkg_batch.hip -- batched commits (hipGraph-captured pipelines over many
// traces) and the continuous-batching serving loop.
//
// Reference counterparts: BatchProver::prove_all and recommend_job_count
// (zkf-metal/src/batch_prover.rs:95-360), where each job runs the whole
// pipeline on its own thread and jobs serialise on the GPU queue, and the
// per-node synchronous GPU dispatch of zkf-runtime (scheduler.rs:247-464).
#include "../../include/zkg.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <algorithm>
#include <chrono>
#include <cstring>

using namespace zkg;

namespace zkg {

static uint64_t esz(int field) { return field == FIELD_GL ? 8 : 4; }
static int ilog2b(uint64_t n)
{
    int l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

// One prepared job: tree storage and LDE scratch allocated, tables resident.
struct Prepared {
    zkg_job job;
    zkg_tree *tree = nullptr;
    void *tmp = nullptr;
    uint64_t tmp_bytes = 0;
};

static int validate_job(zkg_ctx *ctx, const zkg_job &j)
{
    if (!j.trace || !j.lde_out) return set_err(ZKG_EINVAL, "job without buffers");
    if (j.field != FIELD_GL && j.field != FIELD_BB) return set_err(ZKG_EINVAL, "job field %d", j.field);
    if (!j.height || (j.height & (j.height - 1)) || ilog2b(j.height) > 22)
        return set_err(ZKG_EINVAL, "job height %llu", (unsigned long long)j.height);
    if (j.log_blowup > 2 || j.width == 0) return set_err(ZKG_EINVAL, "job blowup/width");
    if (!get_p2(ctx, j.field, j.p2_width)) return set_err(ZKG_ESTATE, "Poseidon2 constants for job not loaded");
    uint64_t es = esz(j.field);
    if (span_bytes(j.height, j.width, es) > j.trace->bytes || span_bytes(j.height << j.log_blowup, j.width, es) > j.lde_out->bytes)
        return set_err(ZKG_EINVAL, "job buffer too small");
    return ZKG_OK;
}

// Tree + scratch bytes a job needs beyond its caller-owned buffers.
static uint64_t job_extra_bytes(const zkg_job &j)
{
    uint64_t es = esz(j.field), big = j.height << j.log_blowup;
    return 2 * big * (j.p2_width / 2) * es + 256 * 32 + coset_lde_tmp_bytes(j.field, j.height, j.width, j.log_blowup);
}

static void release_prepared(zkg_ctx *ctx, Prepared &p, bool keep_tree)
{
    dev_free(ctx, p.tmp);
    p.tmp = nullptr;
    if (!keep_tree && p.tree) zkg_tree_free(p.tree);
    p.tree = nullptr;
}

static int prepare_job(zkg_ctx *ctx, const zkg_job &j, Prepared &p)
{
    p.job = j;
    int rc = validate_job(ctx, j);
    if (rc) return rc;
    const CosetSet *cs;
    if ((rc = get_coset(ctx, j.field, ilog2b(j.height), j.log_blowup, j.field == FIELD_GL ? 7 : 31, &cs))) return rc;
    p.tmp_bytes = coset_lde_tmp_bytes(j.field, j.height, j.width, j.log_blowup);
    if (p.tmp_bytes && (rc = dev_alloc(ctx, p.tmp_bytes, &p.tmp))) return rc;
    if ((rc = tree_alloc(ctx, j.field, j.p2_width, j.height << j.log_blowup, j.lde_out->ptr, j.width, &p.tree))) {
        dev_free(ctx, p.tmp);
        p.tmp = nullptr;
        return rc;
    }
    return ZKG_OK;
}

// ---------------------------------------------------------------- shared launches
// A job set runs as:
//   * the coset LDE of every job, dealt (largest first, grouped by field and
//     Poseidon2 width) round-robin over K branch streams, so small transforms
//     overlap large ones;
//   * per (field, width) group, on a stream of its own once the group's LDEs
//     are done: ONE persistent row-hash launch over a table of every member's
//     rows, then one Merkle launch per round, where round r holds the r-th
//     launch of every member's tree (k_merkle_segs maps its blocks to the
//     trees by prefix offsets).
// Under stream capture the branches and group streams become parallel paths
// of the graph.
constexpr int MAX_BRANCHES = 8, MAX_GROUPS = 3;

struct Fork {
    int k = 0;
    hipStream_t s[MAX_BRANCHES] = {};  // LDE branches
    hipStream_t gs[MAX_GROUPS] = {};   // row hash + Merkle of one group
    hipEvent_t start = nullptr;
    hipEvent_t join[MAX_BRANCHES] = {}, gjoin[MAX_GROUPS] = {};
    hipEvent_t lde_done[MAX_GROUPS][MAX_BRANCHES] = {};
    // timed runs: end of each branch's LDEs; per group, row hash start / end
    // and Merkle end
    hipEvent_t t_lde[MAX_BRANCHES] = {}, t_grp[MAX_GROUPS][3] = {};
};

static int fork_init(Fork &f, int k)
{
    f.k = k;
    hipError_t e = hipEventCreateWithFlags(&f.start, hipEventDisableTiming);
    for (int i = 0; i < k && e == hipSuccess; i++) {
        e = hipStreamCreateWithFlags(&f.s[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f.join[i], hipEventDisableTiming);
        for (int g = 0; g < MAX_GROUPS && e == hipSuccess; g++)
            e = hipEventCreateWithFlags(&f.lde_done[g][i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreate(&f.t_lde[i]);
    }
    for (int g = 0; g < MAX_GROUPS && e == hipSuccess; g++) {
        e = hipStreamCreateWithFlags(&f.gs[g], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f.gjoin[g], hipEventDisableTiming);
        for (int t = 0; t < 3 && e == hipSuccess; t++) e = hipEventCreate(&f.t_grp[g][t]);
    }
    return e == hipSuccess ? ZKG_OK : hip_err(e, "fork streams");
}

static void fork_destroy(Fork &f)
{
    for (int i = 0; i < f.k; i++) {
        if (f.s[i]) {
            (void)hipStreamSynchronize(f.s[i]);
            (void)hipStreamDestroy(f.s[i]);
        }
        if (f.join[i]) (void)hipEventDestroy(f.join[i]);
        for (int g = 0; g < MAX_GROUPS; g++)
            if (f.lde_done[g][i]) (void)hipEventDestroy(f.lde_done[g][i]);
        if (f.t_lde[i]) (void)hipEventDestroy(f.t_lde[i]);
    }
    for (int g = 0; g < MAX_GROUPS; g++) {
        if (f.gs[g]) {
            (void)hipStreamSynchronize(f.gs[g]);
            (void)hipStreamDestroy(f.gs[g]);
        }
        if (f.gjoin[g]) (void)hipEventDestroy(f.gjoin[g]);
        for (int t = 0; t < 3; t++)
            if (f.t_grp[g][t]) (void)hipEventDestroy(f.t_grp[g][t]);
    }
    if (f.start) (void)hipEventDestroy(f.start);
    f = Fork();
}

static uint64_t job_cells(const zkg_job &j) { return (j.height << j.log_blowup) * j.width * (j.field == FIELD_GL ? 4 : 1); }

// The job tables of one job set, laid out in one block of device memory:
// per group its HashJob table, its counter pair and one MerkleSeg table per
// round.  `host` is the image uploaded to `dev` (the counters zero).
struct GroupPlan {
    int field = 0, width = 0;
    std::vector<size_t> members;  // indices into the prepared jobs, LDE order
    int njobs = 0;
    uint32_t ntiles = 0;
    uint64_t hash_off = 0, ctr_off = 0;
    std::vector<uint64_t> round_off;
    std::vector<int> round_segs;
    std::vector<uint32_t> round_blocks;
};

struct SharedPlan {
    std::vector<GroupPlan> groups;
    std::vector<uint8_t> host;
    void *dev = nullptr;
};

static void plan_release(zkg_ctx *ctx, SharedPlan &sp)
{
    dev_free(ctx, sp.dev);
    sp = SharedPlan();
}

static uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

static int plan_build(zkg_ctx *ctx, std::vector<Prepared> &ps, SharedPlan &sp)
{
    sp = SharedPlan();
    for (size_t i = 0; i < ps.size(); i++) {
        const zkg_job &j = ps[i].job;
        GroupPlan *g = nullptr;
        for (auto &x : sp.groups)
            if (x.field == j.field && x.width == j.p2_width) g = &x;
        if (!g) {
            if (sp.groups.size() == MAX_GROUPS) return set_err(ZKG_EINVAL, "more than %d Poseidon2 instances", MAX_GROUPS);
            sp.groups.push_back(GroupPlan());
            g = &sp.groups.back();
            g->field = j.field;
            g->width = j.p2_width;
        }
        g->members.push_back(i);
    }
    uint64_t off = 0;
    std::vector<std::vector<HashJob>> hash(sp.groups.size());
    std::vector<std::vector<std::vector<MerkleSeg>>> rounds(sp.groups.size());
    for (size_t gi = 0; gi < sp.groups.size(); gi++) {
        GroupPlan &g = sp.groups[gi];
        // LDEs largest first; row-hash tiles heaviest (most permutations per
        // row) first, so the persistent grid takes the long tiles early
        std::sort(g.members.begin(), g.members.end(),
                  [&](size_t a, size_t b) { return job_cells(ps[a].job) > job_cells(ps[b].job); });
        std::vector<size_t> by_cost = g.members;
        const uint32_t rate = (uint32_t)g.width / 2;
        std::stable_sort(by_cost.begin(), by_cost.end(), [&](size_t a, size_t b) {
            return (ps[a].job.width + rate - 1) / rate > (ps[b].job.width + rate - 1) / rate;
        });
        uint64_t tiles = 0;
        for (size_t m : by_cost) {
            const zkg_tree *t = ps[m].tree;
            HashJob h{t->mat, t->layers[0], t->n_leaves, ps[m].job.width, (uint32_t)tiles};
            hash[gi].push_back(h);
            tiles += (t->n_leaves + HASH_TILE_ROWS - 1) / HASH_TILE_ROWS;
        }
        if (tiles > UINT32_MAX / 2) return set_err(ZKG_EINVAL, "job set too large for one launch");
        g.ntiles = (uint32_t)tiles;
        g.njobs = (int)hash[gi].size();
        g.hash_off = off = align_up(off, 256);
        off += hash[gi].size() * sizeof(HashJob);
        g.ctr_off = off = align_up(off, 256);
        off += 2 * sizeof(uint32_t);
        for (size_t m : g.members) {
            zkg_tree *t = ps[m].tree;
            std::vector<MerkleSeg> segs;
            if (merkle_plan(g.field, g.width, t->layers.data(), t->nlayers, t->n_leaves, segs))
                return set_err(ZKG_EINVAL, "Poseidon2 field %d width %d", g.field, g.width);
            for (size_t r = 0; r < segs.size(); r++) {
                if (rounds[gi].size() <= r) rounds[gi].emplace_back();
                rounds[gi][r].push_back(segs[r]);
            }
        }
        for (auto &rd : rounds[gi]) {
            uint64_t first = 0;
            for (auto &sg : rd) {
                sg.first = (uint32_t)first;
                first += sg.blocks;
            }
            if (first > UINT32_MAX / 2) return set_err(ZKG_EINVAL, "job set too large for one launch");
            g.round_off.push_back(off = align_up(off, 256));
            g.round_segs.push_back((int)rd.size());
            g.round_blocks.push_back((uint32_t)first);
            off += rd.size() * sizeof(MerkleSeg);
        }
    }
    sp.host.assign(align_up(off, 256), 0);
    for (size_t gi = 0; gi < sp.groups.size(); gi++) {
        const GroupPlan &g = sp.groups[gi];
        memcpy(sp.host.data() + g.hash_off, hash[gi].data(), hash[gi].size() * sizeof(HashJob));
        for (size_t r = 0; r < rounds[gi].size(); r++)
            memcpy(sp.host.data() + g.round_off[r], rounds[gi][r].data(), rounds[gi][r].size() * sizeof(MerkleSeg));
    }
    return dev_alloc(ctx, sp.host.size(), &sp.dev);
}

// timed: also records the Fork's timing events (never under stream capture).
static int enqueue_shared(zkg_ctx *ctx, hipStream_t st, std::vector<Prepared> &ps, Fork &f, const SharedPlan &sp,
                          bool timed = false)
{
    const int k = std::max(1, std::min<int>(f.k, (int)ps.size()));
    hipError_t e = hipEventRecord(f.start, st);
    for (int i = 0; i < k && e == hipSuccess; i++) e = hipStreamWaitEvent(f.s[i], f.start, 0);
    if (e != hipSuccess) return hip_err(e, "fork");
    int rc = ZKG_OK;
    // LDEs, group by group, dealt over the branches
    int next = 0;
    for (size_t gi = 0; gi < sp.groups.size() && !rc; gi++) {
        for (size_t m : sp.groups[gi].members) {
            Prepared &p = ps[m];
            rc = run_lde(ctx, f.s[next], p.job.field, p.job.trace->ptr, p.job.height, p.job.width, p.job.log_blowup,
                         p.job.field == FIELD_GL ? 7 : 31, p.job.lde_out->ptr, p.tmp);
            next = (next + 1) % k;
            if (rc) break;
        }
        for (int i = 0; i < k && !rc; i++) {
            e = hipEventRecord(f.lde_done[gi][i], f.s[i]);
            if (e != hipSuccess) rc = hip_err(e, "lde event");
        }
    }
    for (int i = 0; i < k && timed && !rc; i++) (void)hipEventRecord(f.t_lde[i], f.s[i]);
    // per group: the shared row hash, then the Merkle rounds
    uint8_t *dev = (uint8_t *)sp.dev;
    for (size_t gi = 0; gi < sp.groups.size() && !rc; gi++) {
        const GroupPlan &g = sp.groups[gi];
        hipStream_t gs = f.gs[gi];
        for (int i = 0; i < k && e == hipSuccess; i++) e = hipStreamWaitEvent(gs, f.lde_done[gi][i], 0);
        const P2Set *p2 = get_p2(ctx, g.field, g.width);
        if (timed && e == hipSuccess) e = hipEventRecord(f.t_grp[gi][0], gs);
        if (e == hipSuccess)
            e = launch_hash_jobs(g.field, g.width, *p2, (const HashJob *)(dev + g.hash_off), g.njobs, g.ntiles,
                                 (uint32_t *)(dev + g.ctr_off), gs);
        if (timed && e == hipSuccess) e = hipEventRecord(f.t_grp[gi][1], gs);
        for (size_t r = 0; r < g.round_off.size() && e == hipSuccess; r++)
            e = launch_merkle_segs(g.field, g.width, *p2, (const MerkleSeg *)(dev + g.round_off[r]), g.round_segs[r],
                                   g.round_blocks[r], gs);
        if (timed && e == hipSuccess) e = hipEventRecord(f.t_grp[gi][2], gs);
        if (e != hipSuccess) rc = hip_err(e, "shared launches");
    }
    // always join, so the origin stream (or capture) is left consistent
    for (int i = 0; i < k; i++) {
        hipError_t e2 = hipEventRecord(f.join[i], f.s[i]);
        if (e2 == hipSuccess) e2 = hipStreamWaitEvent(st, f.join[i], 0);
        if (e2 != hipSuccess && !rc) rc = hip_err(e2, "join");
    }
    for (size_t gi = 0; gi < sp.groups.size(); gi++) {
        hipError_t e2 = hipEventRecord(f.gjoin[gi], f.gs[gi]);
        if (e2 == hipSuccess) e2 = hipStreamWaitEvent(st, f.gjoin[gi], 0);
        if (e2 != hipSuccess && !rc) rc = hip_err(e2, "join");
    }
    return rc;
}

// Cached executable graph for one job-set (same shapes, buffers and trees).
struct BatchState {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    std::vector<Prepared> prepared;
    SharedPlan plan;
    Fork fork;
};

static std::vector<uint64_t> batch_key(const zkg_job *jobs, uint32_t n)
{
    std::vector<uint64_t> k;
    for (uint32_t i = 0; i < n; i++) {
        const zkg_job &j = jobs[i];
        k.insert(k.end(), {(uint64_t)j.field, (uint64_t)j.p2_width, (uint64_t)(uintptr_t)j.trace->ptr, j.height,
                           j.width, j.log_blowup, (uint64_t)(uintptr_t)j.lde_out->ptr});
    }
    return k;
}

void batch_release(zkg_ctx *ctx)
{
    BatchState *b = ctx->batch;
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    for (auto &p : b->prepared) release_prepared(ctx, p, false);
    plan_release(ctx, b->plan);
    fork_destroy(b->fork);
    delete b;
    ctx->batch = nullptr;
}

}  // namespace zkg

// Trees returned to the caller are copies of the cached graph's trees (the
// graph writes into its own tree storage, so a replay never touches a tree the
// caller holds).
static int clone_tree(zkg_ctx *ctx, hipStream_t st, const zkg_tree *src, zkg_tree **out)
{
    zkg_tree *t;
    int rc = tree_alloc(ctx, src->field, src->p2_width, src->n_leaves, src->mat, src->mat_width, &t);
    if (rc) return rc;
    uint64_t bytes = src->layer_off.back() + src->digest * esz(src->field);
    hipError_t e = hipMemcpyAsync(t->base, src->base, bytes, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) {
        zkg_tree_free(t);
        return hip_err(e, "clone tree");
    }
    *out = t;
    return ZKG_OK;
}

extern "C" int zkg_commit_batch(zkg_ctx *ctx, const zkg_job *jobs, uint32_t n, zkg_tree **trees_out)
{
    if (!ctx || (!jobs && n) || (!trees_out && n)) return set_err(ZKG_EINVAL, "null argument");
    if (!n) return ZKG_OK;
    for (uint32_t i = 0; i < n; i++)
        if (int rc = validate_job(ctx, jobs[i])) return rc;
    hipStream_t st = ctx->stream;
    std::vector<uint64_t> key = batch_key(jobs, n);
    if (!ctx->batch || ctx->batch->key != key) {
        (void)hipStreamSynchronize(st);
        batch_release(ctx);
        BatchState *b = new BatchState();
        b->key = key;
        b->prepared.resize(n);
        int rc = ZKG_OK;
        for (uint32_t i = 0; i < n && !rc; i++) rc = prepare_job(ctx, jobs[i], b->prepared[i]);
        if (rc) {
            for (auto &p : b->prepared) release_prepared(ctx, p, false);
            delete b;
            return rc;
        }
        if (!rc) rc = fork_init(b->fork, 4);
        // the job tables (counters zero) are uploaded once, before capture:
        // the graph holds kernel launches only
        if (!rc) rc = plan_build(ctx, b->prepared, b->plan);
        if (!rc) {
            hipError_t e = hipMemcpy(b->plan.dev, b->plan.host.data(), b->plan.host.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = hip_err(e, "upload job tables");
        }
        hipGraph_t g = nullptr;
        if (!rc) {
            hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) rc = hip_err(e, "begin capture");
            if (!rc) {
                rc = enqueue_shared(ctx, st, b->prepared, b->fork, b->plan);
                e = hipStreamEndCapture(st, &g);
                if (!rc && e != hipSuccess) rc = hip_err(e, "end capture");
            }
        }
        if (!rc) {
            hipError_t e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = hip_err(e, "graph instantiate");
        }
        if (g) (void)hipGraphDestroy(g);
        if (rc) {
            for (auto &p : b->prepared) release_prepared(ctx, p, false);
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            plan_release(ctx, b->plan);
            fork_destroy(b->fork);
            delete b;
            return rc;
        }
        ctx->batch = b;
    }
    // With timing on, the same launches run straight from the streams (not
    // from the graph) so that events can bracket the stages on the streams
    // the kernels run on.
    BatchState *b = ctx->batch;
    stage_mark(ctx, st, 0);
    hipError_t e = hipSuccess;
    if (ctx->timing) {
        if (int rc = enqueue_shared(ctx, st, b->prepared, b->fork, b->plan, true)) return rc;
    } else {
        e = hipGraphLaunch(b->exec, st);
        if (e != hipSuccess) return hip_err(e, "graph launch");
    }
    stage_mark(ctx, st, 1);
    for (uint32_t i = 0; i < n; i++) {
        int rc = clone_tree(ctx, st, b->prepared[i].tree, &trees_out[i]);
        if (rc) {
            for (uint32_t k = 0; k < i; k++) zkg_tree_free(trees_out[k]);
            return rc;
        }
    }
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_err(e, "batch sync");
    if (ctx->timing) {
        // stages: 0 LDE (start to the last branch's end), 1 row hash and 2
        // Merkle (summed over the Poseidon2 groups), 3 the whole batch
        for (int i = 0; i < 4; i++) ctx->stage_ms[i] = 0;
        float ms = 0;
        const int k = std::max(1, std::min<int>(b->fork.k, (int)b->prepared.size()));
        for (int i = 0; i < k; i++)
            if (hipEventElapsedTime(&ms, ctx->ev[0], b->fork.t_lde[i]) == hipSuccess)
                ctx->stage_ms[0] = std::max<double>(ctx->stage_ms[0], ms);
        for (size_t g = 0; g < b->plan.groups.size(); g++) {
            if (hipEventElapsedTime(&ms, b->fork.t_grp[g][0], b->fork.t_grp[g][1]) == hipSuccess) ctx->stage_ms[1] += ms;
            if (hipEventElapsedTime(&ms, b->fork.t_grp[g][1], b->fork.t_grp[g][2]) == hipSuccess) ctx->stage_ms[2] += ms;
        }
        if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->stage_ms[3] = ms;
    }
    return ZKG_OK;
}

// ======================================================================
// Serving loop
// ======================================================================
namespace zkg {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket;
    zkg_job job;
    Clock::time_point arrival;
    Clock::time_point finish;
    zkg_tree *tree = nullptr;
    int status = ZKG_EBUSY;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Fork fork;
    bool busy = false;
    std::vector<Prepared> prepared;
    SharedPlan plan;  // its host image stays alive until the slot's upload has run
    std::vector<uint64_t> tickets;
};

struct ServeState {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<Slot> slots;
    uint32_t max_batch = 16;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
};

static void complete_slot(zkg_ctx *ctx, ServeState *s, Slot &sl)
{
    Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> g(s->mu);
    for (size_t i = 0; i < sl.prepared.size(); i++) {
        Request &r = s->reqs[sl.tickets[i]];
        r.tree = sl.prepared[i].tree;
        r.status = ZKG_OK;
        r.finish = now;
        release_prepared(ctx, sl.prepared[i], true);
    }
    plan_release(ctx, sl.plan);
    sl.prepared.clear();
    sl.tickets.clear();
    sl.busy = false;
}

static void serve_loop(zkg_ctx *ctx, ServeState *s)
{
    (void)hipSetDevice(ctx->device);
    for (;;) {
        bool progressed = false;
        for (Slot &sl : s->slots) {
            if (sl.busy && hipEventQuery(sl.done) == hipSuccess) {
                complete_slot(ctx, s, sl);
                progressed = true;
            }
        }
        for (Slot &sl : s->slots) {
            if (sl.busy) continue;
            // admit FIFO requests while the HBM budget allows
            std::vector<uint64_t> take;
            std::vector<zkg_job> jobs;
            {
                std::lock_guard<std::mutex> g(s->mu);
                while (!s->queue.empty() && take.size() < s->max_batch) {
                    uint64_t t = s->queue.front();
                    const zkg_job &j = s->reqs[t].job;
                    if (hbm_in_use(ctx) + job_extra_bytes(j) > ctx->hbm_budget && !take.empty()) break;
                    take.push_back(t);
                    jobs.push_back(j);
                    s->queue.pop_front();
                }
            }
            if (take.empty()) break;
            std::vector<Prepared> ps(take.size());
            size_t ok = 0;
            int rc = ZKG_OK;
            for (; ok < take.size(); ok++) {
                rc = prepare_job(ctx, jobs[ok], ps[ok]);
                if (rc) break;
            }
            if (ok < take.size()) {
                // A request that cannot get memory waits at the queue head for
                // a slot to drain; with nothing in flight it can never fit and
                // fails.  Any other error fails that request only.
                bool others_busy = false;
                for (Slot &o : s->slots) others_busy |= o.busy;
                bool wait = rc == ZKG_ENOMEM && (ok > 0 || others_busy);
                std::lock_guard<std::mutex> g(s->mu);
                size_t first_requeue = ok;
                if (!wait) {
                    s->reqs[take[ok]].status = rc;
                    s->reqs[take[ok]].finish = Clock::now();
                    first_requeue = ok + 1;
                }
                for (size_t k = take.size(); k-- > first_requeue;) s->queue.push_front(take[k]);
                ps.resize(ok);
                take.resize(ok);
                if (ok == 0) {
                    if (wait) break;
                    continue;
                }
            }
            // the admitted requests share launches: one row hash per
            // (field, width) group and one Merkle launch per round
            rc = plan_build(ctx, ps, sl.plan);
            if (rc == ZKG_OK) {
                hipError_t e = hipMemcpyAsync(sl.plan.dev, sl.plan.host.data(), sl.plan.host.size(),
                                              hipMemcpyHostToDevice, sl.stream);
                if (e != hipSuccess) rc = hip_err(e, "upload job tables");
            }
            if (rc == ZKG_OK) rc = enqueue_shared(ctx, sl.stream, ps, sl.fork, sl.plan);
            if (rc == ZKG_OK) {
                hipError_t e = hipEventRecord(sl.done, sl.stream);
                if (e != hipSuccess) rc = hip_err(e, "event record");
            }
            if (rc != ZKG_OK) {
                (void)hipStreamSynchronize(sl.stream);
                for (int i = 0; i < sl.fork.k; i++) (void)hipStreamSynchronize(sl.fork.s[i]);
                for (int g = 0; g < MAX_GROUPS; g++) (void)hipStreamSynchronize(sl.fork.gs[g]);
                plan_release(ctx, sl.plan);
                std::lock_guard<std::mutex> g(s->mu);
                for (size_t i = 0; i < ps.size(); i++) {
                    release_prepared(ctx, ps[i], false);
                    s->reqs[take[i]].status = rc;
                    s->reqs[take[i]].finish = Clock::now();
                }
                continue;
            }
            sl.prepared = std::move(ps);
            sl.tickets = take;
            sl.busy = true;
            progressed = true;
        }
        bool any_busy = false;
        for (Slot &sl : s->slots) any_busy |= sl.busy;
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop && !any_busy && s->queue.empty()) break;
        if (!progressed) {
            if (any_busy)
                s->cv.wait_for(lk, std::chrono::microseconds(50));
            else
                s->cv.wait_for(lk, std::chrono::milliseconds(2));
        }
    }
}

void serve_shutdown(zkg_ctx *ctx)
{
    ServeState *s = ctx->serve;
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Slot &sl : s->slots) {
        (void)hipStreamSynchronize(sl.stream);
        fork_destroy(sl.fork);
        (void)hipEventDestroy(sl.done);
        (void)hipStreamDestroy(sl.stream);
    }
    for (auto &kv : s->reqs)
        if (kv.second.tree) zkg_tree_free(kv.second.tree);
    delete s;
    ctx->serve = nullptr;
}

}  // namespace zkg

extern "C" int zkg_serve_start(zkg_ctx *ctx, uint32_t slots, uint32_t max_batch)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    if (ctx->serve) return set_err(ZKG_ESTATE, "serving loop already running");
    if (slots == 0 || slots > 8 || max_batch == 0) return set_err(ZKG_EINVAL, "slots 1..8, max_batch >= 1");
    ServeState *s = new ServeState();
    s->max_batch = max_batch;
    s->slots.resize(slots);
    for (Slot &sl : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        int rc = e == hipSuccess ? fork_init(sl.fork, 2) : hip_err(e, "serve stream");
        if (rc != ZKG_OK) {
            for (Slot &x : s->slots) {
                fork_destroy(x.fork);
                if (x.done) (void)hipEventDestroy(x.done);
                if (x.stream) (void)hipStreamDestroy(x.stream);
            }
            delete s;
            return rc;
        }
    }
    ctx->serve = s;
    s->worker = std::thread(serve_loop, ctx, s);
    return ZKG_OK;
}

extern "C" int zkg_serve_submit(zkg_ctx *ctx, const zkg_job *job, uint64_t *ticket)
{
    if (!ctx || !job || !ticket) return set_err(ZKG_EINVAL, "null argument");
    ServeState *s = ctx->serve;
    if (!s) return set_err(ZKG_ESTATE, "serving loop not running");
    int rc = validate_job(ctx, *job);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return set_err(ZKG_ESTATE, "serving loop stopping");
        Request r;
        r.ticket = s->next_ticket++;
        r.job = *job;
        r.arrival = Clock::now();
        s->reqs[r.ticket] = r;
        s->queue.push_back(r.ticket);
        *ticket = r.ticket;
    }
    s->cv.notify_all();
    return ZKG_OK;
}

extern "C" int zkg_serve_poll(zkg_ctx *ctx, uint64_t ticket, zkg_tree **tree, double *latency_ms)
{
    if (!ctx || !tree) return set_err(ZKG_EINVAL, "null argument");
    ServeState *s = ctx->serve;
    if (!s) return set_err(ZKG_ESTATE, "serving loop not running");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_err(ZKG_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    Request &r = it->second;
    if (r.status == ZKG_EBUSY) return ZKG_EBUSY;
    int st = r.status;
    if (latency_ms) latency_ms[0] = std::chrono::duration<double, std::milli>(r.finish - r.arrival).count();
    *tree = r.tree;
    s->reqs.erase(it);
    if (st != ZKG_OK) return set_err(st, "request %llu failed", (unsigned long long)ticket);
    return ZKG_OK;
}

extern "C" int zkg_serve_stop(zkg_ctx *ctx)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    if (!ctx->serve) return ZKG_OK;
    serve_shutdown(ctx);
    return ZKG_OK;
}
