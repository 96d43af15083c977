// zkg_tuning.hip -- per-shape choice of the transform tile plan, and the
// work-form splits of the row hash and Merkle levels.
//
// Each setting resolves through three layers, first hit wins, as the
// reference's device tuning does (zkf-metal/src/tuning.rs:276-365: runtime
// override > learned > static profile; learned thresholds come from
// zkf-runtime/src/adaptive_tuning.rs):
//   1. override: set by the caller (zkg_lde_plan_set / zkg_split_set with
//      ZKG_TUNE_OVERRIDE; the ZKG_MERKLE_LANES_LOG2, ZKG_MERKLE_CHUNKS_LOG2,
//      ZKG_HASH_CHUNKS_LOG2 and ZKG_MERKLE_FUSED_K environment variables seed
//      the split overrides at first use);
//   2. learned: measured on the live device (zkg_lde_plan_learn) or installed
//      by the caller (ZKG_TUNE_LEARNED);
//   3. static: the table ziros_amd/tune.py measured on gfx950 and wrote into
//      zkg_tuned.h, matched to the nearest measured shape; then the built-in
//      defaults (plan 0; the split levels below).
// Override and learned entries match a shape exactly, with -1 fields
// matching anything.  All of it is process-wide and thread-safe: plans are
// looked up on every transform launch, including from the serving thread.
#include "../../include/zkg.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#if __has_include("zkg_tuned.h")
#include "zkg_tuned.h"
#endif
#ifndef ZKG_TUNED_LDE_TABLE
#define ZKG_TUNED_LDE_TABLE {}
#endif
#ifndef ZKG_TUNED_SPLIT_TABLE
#define ZKG_TUNED_SPLIT_TABLE {}
#endif

using namespace zkg;

namespace {

// Built-in split levels (log2 of the level / matrix sizes), measured by
// sweeps on gfx950: Merkle levels of at most 2^lanes digests use W lanes per
// compression; Goldilocks levels of at most 2^chunks a 4-chunk of the state
// per lane; Goldilocks row hashes of at most 2^hash_chunks rows a 4-chunk per
// lane; fused_k levels per one-compression-per-lane launch.
struct SplitDefault {
    int field, width, lanes, chunks, hash_chunks, fused;
};
constexpr SplitDefault SPLIT_BUILTIN[] = {
    {ZKG_BABYBEAR, 16, 15, 0, 0, 3},
    {ZKG_GOLDILOCKS, 16, 13, 16, 15, 3},
    {ZKG_GOLDILOCKS, 8, 14, 16, 15, 3},
};

std::mutex g_mu;
std::vector<zkg_plan_entry> g_plan[2];    // override, learned
std::vector<zkg_split_entry> g_split[2];  // override, learned
const zkg_plan_entry g_plan_static[] = ZKG_TUNED_LDE_TABLE;
const zkg_split_entry g_split_static[] = ZKG_TUNED_SPLIT_TABLE;
constexpr size_t N_PLAN_STATIC = sizeof(g_plan_static) / sizeof(zkg_plan_entry);
constexpr size_t N_SPLIT_STATIC = sizeof(g_split_static) / sizeof(zkg_split_entry);
bool g_env_read = false;

int width_class(uint32_t width)
{
    int c = 0;
    while ((1ull << c) < width) c++;
    return c;
}

bool matches(int32_t want, int32_t have) { return want < 0 || want == have; }

int env_int(const char *name, int lo, int hi)
{
    const char *e = getenv(name);
    int v = 0;
    if (!e || sscanf(e, "%d", &v) != 1 || v < lo || v > hi) return -1;
    return v;
}

// The environment knobs become override entries for every instance (under g_mu).
void read_env_locked()
{
    if (g_env_read) return;
    g_env_read = true;
    zkg_split_entry e{-1, -1, env_int("ZKG_MERKLE_LANES_LOG2", 0, 30), env_int("ZKG_MERKLE_CHUNKS_LOG2", 0, 30),
                      env_int("ZKG_HASH_CHUNKS_LOG2", 0, 30), env_int("ZKG_MERKLE_FUSED_K", 1, 9)};
    if (e.merkle_lanes_log2 >= 0 || e.merkle_chunks_log2 >= 0 || e.hash_chunks_log2 >= 0 || e.fused_levels >= 0)
        g_split[ZKG_TUNE_OVERRIDE].insert(g_split[ZKG_TUNE_OVERRIDE].begin(), e);
}

int plan_lookup_locked(int field, int log_n, int wc, int lb, int *source)
{
    for (int layer = 0; layer < 2; layer++)
        for (const zkg_plan_entry &e : g_plan[layer])
            if (matches(e.field, field) && matches(e.log_height, log_n) && matches(e.width_class, wc) &&
                matches(e.log_blowup, lb)) {
                if (source) *source = layer;
                return e.plan;
            }
    // static: the nearest measured shape of the same field
    int best = -1, best_d = 1 << 30;
    for (size_t i = 0; i < N_PLAN_STATIC; i++) {
        const zkg_plan_entry &e = g_plan_static[i];
        if (e.field != field) continue;
        const int d = 2 * abs(e.log_height - log_n) + abs(e.width_class - wc) + 2 * abs(e.log_blowup - lb);
        if (d < best_d) {
            best_d = d;
            best = e.plan;
        }
    }
    if (best >= 0) {
        if (source) *source = 2;
        return best;
    }
    if (source) *source = 3;
    return 0;
}

int split_field(const zkg_split_entry &e, int which)
{
    switch (which) {
    case SPLIT_LANES: return e.merkle_lanes_log2;
    case SPLIT_CHUNKS: return e.merkle_chunks_log2;
    case SPLIT_HASH_CHUNKS: return e.hash_chunks_log2;
    default: return e.fused_levels;
    }
}

int builtin_split(int field, int width, int which)
{
    for (const SplitDefault &d : SPLIT_BUILTIN)
        if (d.field == field && d.width == width) {
            switch (which) {
            case SPLIT_LANES: return d.lanes;
            case SPLIT_CHUNKS: return d.chunks;
            case SPLIT_HASH_CHUNKS: return d.hash_chunks;
            default: return d.fused;
            }
        }
    return which == SPLIT_FUSED ? 3 : 0;
}

int split_valid(int which, int v) { return which == SPLIT_FUSED ? (v >= 1 && v <= 9) : (v >= 0 && v <= 30); }

}  // namespace

int zkg::lde_plan_for(int field, int log_n, uint32_t width, uint32_t log_blowup)
{
    std::lock_guard<std::mutex> g(g_mu);
    int p = plan_lookup_locked(field, log_n, width_class(width), (int)log_blowup, nullptr);
    return p >= 0 && p < LDE_PLANS ? p : 0;
}

int zkg::split_value(int field, int width, int which)
{
    std::lock_guard<std::mutex> g(g_mu);
    read_env_locked();
    for (int layer = 0; layer < 2; layer++)
        for (const zkg_split_entry &e : g_split[layer])
            if (matches(e.field, field) && matches(e.p2_width, width)) {
                const int v = split_field(e, which);
                if (split_valid(which, v)) return v;
            }
    for (size_t i = 0; i < N_SPLIT_STATIC; i++) {
        const zkg_split_entry &e = g_split_static[i];
        if (e.field == field && e.p2_width == width) {
            const int v = split_field(e, which);
            if (split_valid(which, v)) return v;
        }
    }
    return builtin_split(field, width, which);
}

extern "C" int zkg_lde_plan_count(void) { return LDE_PLANS; }

extern "C" const char *zkg_lde_plan_name(int plan)
{
    return plan >= 0 && plan < LDE_PLANS ? LDE_PLAN_TABLE[plan].name : nullptr;
}

extern "C" int zkg_lde_plan_set(int layer, const zkg_plan_entry *entries, uint32_t n)
{
    if (layer != ZKG_TUNE_OVERRIDE && layer != ZKG_TUNE_LEARNED) return set_err(ZKG_EINVAL, "layer %d", layer);
    if (n && !entries) return set_err(ZKG_EINVAL, "null entries");
    for (uint32_t i = 0; i < n; i++)
        if (entries[i].plan < 0 || entries[i].plan >= LDE_PLANS)
            return set_err(ZKG_EINVAL, "entry %u: plan %d not in 0..%d", i, entries[i].plan, LDE_PLANS - 1);
    std::lock_guard<std::mutex> g(g_mu);
    g_plan[layer].assign(entries, entries + n);
    return ZKG_OK;
}

extern "C" int zkg_lde_plan_for(int field, uint64_t height, uint32_t width, uint32_t log_blowup, int *plan,
                                int *source)
{
    if (!plan) return set_err(ZKG_EINVAL, "null argument");
    if (field != ZKG_GOLDILOCKS && field != ZKG_BABYBEAR) return set_err(ZKG_EINVAL, "field %d", field);
    int log_n = 0;
    while ((1ull << log_n) < height) log_n++;
    std::lock_guard<std::mutex> g(g_mu);
    *plan = plan_lookup_locked(field, log_n, width_class(width), (int)log_blowup, source);
    return ZKG_OK;
}

extern "C" int zkg_split_set(int layer, const zkg_split_entry *entries, uint32_t n)
{
    if (layer != ZKG_TUNE_OVERRIDE && layer != ZKG_TUNE_LEARNED) return set_err(ZKG_EINVAL, "layer %d", layer);
    if (n && !entries) return set_err(ZKG_EINVAL, "null entries");
    for (uint32_t i = 0; i < n; i++) {
        const zkg_split_entry &e = entries[i];
        if ((e.merkle_lanes_log2 >= 0 && e.merkle_lanes_log2 > 30) || (e.merkle_chunks_log2 >= 0 && e.merkle_chunks_log2 > 30) ||
            (e.hash_chunks_log2 >= 0 && e.hash_chunks_log2 > 30) || (e.fused_levels >= 0 && (e.fused_levels < 1 || e.fused_levels > 9)))
            return set_err(ZKG_EINVAL, "entry %u out of range (log2 levels 0..30, fused levels 1..9)", i);
    }
    std::lock_guard<std::mutex> g(g_mu);
    read_env_locked();  // so that a later first use does not add them back
    g_split[layer].assign(entries, entries + n);  // replaces the layer (the environment's entries too)
    return ZKG_OK;
}

extern "C" int zkg_split_get(int field, int p2_width, zkg_split_entry *out)
{
    if (!out) return set_err(ZKG_EINVAL, "null argument");
    out->field = field;
    out->p2_width = p2_width;
    out->merkle_lanes_log2 = split_value(field, p2_width, SPLIT_LANES);
    out->merkle_chunks_log2 = split_value(field, p2_width, SPLIT_CHUNKS);
    out->hash_chunks_log2 = split_value(field, p2_width, SPLIT_HASH_CHUNKS);
    out->fused_levels = split_value(field, p2_width, SPLIT_FUSED);
    return ZKG_OK;
}

// Times every plan on this shape (random input on the context's device, the
// median of 3 runs after a warm-up) and records the fastest in the learned
// layer, replacing an entry for the same shape.
extern "C" int zkg_lde_plan_learn(zkg_ctx *ctx, int field, uint64_t height, uint32_t width, uint32_t log_blowup,
                                  int *plan_out)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    if (field != ZKG_GOLDILOCKS && field != ZKG_BABYBEAR) return set_err(ZKG_EINVAL, "field %d", field);
    if (!height || (height & (height - 1)) || height > (1ull << 22) || log_blowup > 2 || !width)
        return set_err(ZKG_EINVAL, "shape: height a power of two <= 2^22, blowup <= 4, width >= 1");
    int log_n = 0;
    while ((1ull << log_n) < height) log_n++;
    if (log_n + (int)log_blowup > LOG_TW_TOP) return set_err(ZKG_EINVAL, "LDE beyond 2^%d rows", LOG_TW_TOP);
    const uint64_t es = field == ZKG_GOLDILOCKS ? 8 : 4;
    const uint64_t in_b = span_bytes(height, width, es), out_b = span_bytes(height << log_blowup, width, es);
    if (in_b > ctx->hbm_budget || out_b > ctx->hbm_budget) return set_err(ZKG_EINVAL, "shape exceeds the HBM budget");
    const uint64_t tmp_b = coset_lde_tmp_bytes(field, height, width, log_blowup);
    void *in = nullptr, *out = nullptr, *tmp = nullptr;
    int rc = dev_alloc(ctx, in_b, &in);
    if (!rc) rc = dev_alloc(ctx, out_b, &out);
    if (!rc && tmp_b) rc = dev_alloc(ctx, tmp_b, &tmp);
    const CosetSet *cs = nullptr;
    if (!rc) rc = get_coset(ctx, field, log_n, (int)log_blowup, field == ZKG_GOLDILOCKS ? 7 : 31, &cs);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best_ms = 0;
    int best = -1;
    if (!rc) {
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = launch_gen(field, in, 0x7E57ull + (uint64_t)log_n, 0, height * width, ctx->stream);
        for (int p = 0; p < LDE_PLANS && e == hipSuccess; p++) {
            float t[3];
            for (int r = 0; r < 4 && e == hipSuccess; r++) {
                e = hipEventRecord(e0, ctx->stream);
                if (e == hipSuccess)
                    e = launch_coset_lde_plan(p, field, ctx->tw[field], *cs, in, out, tmp, height, width, log_blowup,
                                              ctx->stream);
                if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                if (e == hipSuccess && r > 0) e = hipEventElapsedTime(&t[r - 1], e0, e1);
            }
            if (e != hipSuccess) break;
            float lo = t[0] < t[1] ? t[0] : t[1], hi = t[0] < t[1] ? t[1] : t[0];
            const float med = t[2] < lo ? lo : (t[2] > hi ? hi : t[2]);
            if (best < 0 || med < best_ms) {
                best_ms = med;
                best = p;
            }
        }
        if (e != hipSuccess) rc = hip_err(e, "plan timing");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(ctx, in);
    dev_free(ctx, out);
    dev_free(ctx, tmp);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> g(g_mu);
        auto &L = g_plan[ZKG_TUNE_LEARNED];
        const zkg_plan_entry ent{field, log_n, width_class(width), (int32_t)log_blowup, best};
        bool replaced = false;
        for (auto &x : L)
            if (x.field == ent.field && x.log_height == ent.log_height && x.width_class == ent.width_class &&
                x.log_blowup == ent.log_blowup) {
                x.plan = best;
                replaced = true;
            }
        if (!replaced) L.push_back(ent);
    }
    if (plan_out) *plan_out = best;
    return ZKG_OK;
}
