// zkg_runtime.hip -- host runtime behind include/zkg.h: contexts, HBM
// accounting, resident constant tables, trees, the commit pipeline, FRI commit
// phase, batched commits and the serving loop.
//
// Reference counterparts: MetalContext (zkf-metal/src/device.rs:250+,
// 1469-1480), GpuBufferAllocator (zkf-runtime/src/buffer_bridge.rs:77-84),
// MetalMerkleTreeMmcs (zkf-metal/src/merkle.rs:32-443), MetalFri
// (zkf-metal/src/fri.rs:17-147), BatchProver (zkf-metal/src/batch_prover.rs:
// 95-360).  Unlike the reference, twiddles and round constants stay resident
// in HBM for the context's lifetime, and a commit is one stream-ordered
// sequence of launches with no host round trip.
#include "../../include/zkg.h"
#include "zkg_field.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace zkg;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

int zkg::set_err(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int zkg::hip_err(hipError_t e, const char *what)
{
    return set_err(e == hipErrorOutOfMemory ? ZKG_ENOMEM : ZKG_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                       \
    do {                                          \
        hipError_t _e = (expr);                   \
        if (_e != hipSuccess) return hip_err(_e, what); \
    } while (0)

extern "C" const char *zkg_last_error(void) { return g_err.c_str(); }
extern "C" int zkg_abi_version(void) { return ZKG_ABI_VERSION; }

// ------------------------------------------------------------------ host field helpers
namespace {

template <class F>
typename F::T hpow(typename F::T a, uint64_t e)
{
    typename F::T r = F::one();
    while (e) {
        if (e & 1) r = F::mul(r, a);
        a = F::mul(a, a);
        e >>= 1;
    }
    return r;
}

// omega of order 2^log_n in register form (p3 two_adic_generator: the order
// 2^adicity root squared down).
template <class F>
typename F::T root_of_unity(int log_n)
{
    if constexpr (sizeof(typename F::T) == 8) {
        uint64_t w = 0x185629dcda58878cULL;
        for (int i = log_n; i < 32; i++) w = Gl::mul(w, w);
        return w;
    } else {
        uint32_t w = Bb::from_canon(0x1a427a41u);
        for (int i = log_n; i < 27; i++) w = Bb::mul(w, w);
        return w;
    }
}

template <class F>
typename F::T finv(typename F::T a)
{
    return hpow<F>(a, (uint64_t)F::P - 2);
}

int ilog2u(uint64_t n)
{
    int l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

bool is_pow2(uint64_t n) { return n && !(n & (n - 1)); }

uint32_t rev_bits(uint32_t x, int bits)
{
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// rand 0.9.2 SmallRng (Xoshiro256++, SplitMix64 seeding) -- the generator the
// reference feeds to Poseidon2::new_from_rng_128 (zkf-backends/src/plonky3.rs:
// 606-620).
struct SmallRng {
    uint64_t s[4];
    explicit SmallRng(uint64_t st)
    {
        for (int i = 0; i < 4; i++) {
            st += 0x9e3779b97f4a7c15ULL;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next_u64()
    {
        uint64_t out = rotl(s[0] + s[3], 23) + s[0];
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return out;
    }
    uint32_t next_u32() { return (uint32_t)(next_u64() >> 32); }
};

// p3 StandardUniform sampling, returned in canonical form.
//   Goldilocks: next_u64 rejected at >= p.
//   BabyBear (MontyField31): next_u32 >> 1 rejected at >= p, taken as the
//   Montgomery form, so the canonical value is v * 2^-32.
uint64_t sample_canon(SmallRng &r, int field)
{
    if (field == FIELD_GL) {
        for (;;) {
            uint64_t v = r.next_u64();
            if (v < Gl::P) return v;
        }
    }
    for (;;) {
        uint32_t v = r.next_u32() >> 1;
        if (v < Bb::P) return Bb::to_canon(v);  // REDC(v) = v * 2^-32
    }
}

const uint64_t GL_DIAG16[16] = {
    0xde9b91a467d6afc0ULL, 0xc5f16b9c76a9be17ULL, 0x0ab0fef2d540ac55ULL, 0x3001d27009d05773ULL,
    0xed23b1f906d3d9ebULL, 0x5ce73743cba97054ULL, 0x1c3bab944af4ba24ULL, 0x2faa105854dbafaeULL,
    0x53ffb3ae6d421a10ULL, 0xbcda9df8884ba396ULL, 0xfc1273e4a31807bbULL, 0xc77952573d5142c0ULL,
    0x56683339a819b85eULL, 0x328fcbd8f0ddc8ebULL, 0xb5101e303fce9cb7ULL, 0x774487b8c40089bbULL,
};
// p3-goldilocks MATRIX_DIAG_8_GOLDILOCKS (not in the reference tree).
const uint64_t GL_DIAG8[8] = {
    0xa98811a1fed4e3a5ULL, 0x1cc48b54f377e2a0ULL, 0xe40cd4f6c5609a26ULL, 0x11de79ebca97a4a3ULL,
    0x9177c73d8b7e929cULL, 0x2a6fe8085797e791ULL, 0x3de6e93329f8d5adULL, 0x3f7af9125da962feULL,
};

int p2_index(int field, int width)
{
    if (field == FIELD_GL && width == 16) return 0;
    if (field == FIELD_GL && width == 8) return 1;
    if (field == FIELD_BB && width == 16) return 2;
    return -1;
}

int rounds_p_of(int field) { return field == FIELD_GL ? 22 : 13; }

}  // namespace

// ------------------------------------------------------------------ context internals
// Device memory: a per-context cache of freed blocks (sizes rounded to 4 KiB,
// 2 MiB above 1 MiB) so steady-state commits allocate nothing from HIP.  A
// block is only cached after the work that used it has completed (every free
// site follows a stream synchronisation or a completed event).  Live plus
// cached bytes stay within the HBM budget: cached blocks are released first
// when a new allocation would exceed it.
static uint64_t round_alloc(uint64_t b)
{
    if (b == 0) b = 256;
    uint64_t g = b >= (1ull << 20) ? (2ull << 20) : 4096;
    return (b + g - 1) / g * g;
}

static void cache_evict_locked(zkg_ctx *ctx, uint64_t need)
{
    while (!ctx->cache.empty() && ctx->hbm_used + ctx->cached_bytes + need > ctx->hbm_budget) {
        auto it = std::prev(ctx->cache.end());  // largest first
        ctx->cached_bytes -= it->first;
        (void)hipFree(it->second);
        ctx->cache.erase(it);
    }
}

int zkg::dev_alloc(zkg_ctx *ctx, uint64_t bytes, void **out)
{
    std::lock_guard<std::mutex> g(ctx->mem_mu);
    const uint64_t sz = round_alloc(bytes);
    auto it = ctx->cache.lower_bound(sz);
    if (it != ctx->cache.end() && it->first <= sz + sz / 4) {
        void *p = it->second;
        uint64_t have = it->first;
        ctx->cached_bytes -= have;
        ctx->cache.erase(it);
        ctx->hbm_used += have;
        ctx->alloc_sizes[p] = have;
        *out = p;
        return ZKG_OK;
    }
    if (ctx->hbm_used + sz > ctx->hbm_budget)
        return set_err(ZKG_ENOMEM, "HBM budget exceeded: %llu + %llu > %llu", (unsigned long long)ctx->hbm_used,
                       (unsigned long long)sz, (unsigned long long)ctx->hbm_budget);
    cache_evict_locked(ctx, sz);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e == hipErrorOutOfMemory && !ctx->cache.empty()) {
        (void)hipGetLastError();
        for (auto &kv : ctx->cache) (void)hipFree(kv.second);
        ctx->cache.clear();
        ctx->cached_bytes = 0;
        e = hipMalloc(&p, sz);
    }
    if (e != hipSuccess) return hip_err(e, "hipMalloc");
    ctx->hbm_used += sz;
    ctx->alloc_sizes[p] = sz;
    *out = p;
    return ZKG_OK;
}

void zkg::dev_free(zkg_ctx *ctx, void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> g(ctx->mem_mu);
    auto it = ctx->alloc_sizes.find(p);
    if (it == ctx->alloc_sizes.end()) {
        (void)hipFree(p);
        return;
    }
    uint64_t sz = it->second;
    ctx->hbm_used -= sz;
    ctx->alloc_sizes.erase(it);
    ctx->cache.emplace(sz, p);
    ctx->cached_bytes += sz;
}

uint64_t zkg::hbm_in_use(const zkg_ctx *ctx)
{
    std::lock_guard<std::mutex> g(ctx->mem_mu);
    return ctx->hbm_used;
}

void zkg::cache_release(zkg_ctx *ctx)
{
    std::lock_guard<std::mutex> g(ctx->mem_mu);
    for (auto &kv : ctx->cache) (void)hipFree(kv.second);
    ctx->cache.clear();
    ctx->cached_bytes = 0;
}

namespace {

template <class F>
int upload_vec(zkg_ctx *ctx, const std::vector<typename F::T> &v, void **out)
{
    int rc = dev_alloc(ctx, v.size() * sizeof(typename F::T), out);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(*out, v.data(), v.size() * sizeof(typename F::T), hipMemcpyHostToDevice), "upload table");
    return ZKG_OK;
}

template <class F>
int build_twiddles(zkg_ctx *ctx, TwiddleSet &ts)
{
    using T = typename F::T;
    const int half_tile = 1 << (LOG_TW_TILE - 1);
    T w = root_of_unity<F>(LOG_TW_TILE), wi = finv<F>(w);
    std::vector<T> tf(half_tile), ti(half_tile);
    T a = F::one(), b = F::one();
    for (int j = 0; j < half_tile; j++) {
        tf[j] = a;
        ti[j] = b;
        a = F::mul(a, w);
        b = F::mul(b, wi);
    }
    T top = root_of_unity<F>(LOG_TW_TOP), topi = finv<F>(top);
    const int nlo = 1 << TW_SPLIT, nhi = 1 << (LOG_TW_TOP - TW_SPLIT);
    std::vector<T> lf(nlo), li(nlo), hf(nhi), hi(nhi);
    T x = F::one(), y = F::one();
    for (int i = 0; i < nlo; i++) {
        lf[i] = x;
        li[i] = y;
        x = F::mul(x, top);
        y = F::mul(y, topi);
    }
    // x = top^4096 now
    T step = x, stepi = y;
    x = F::one();
    y = F::one();
    for (int i = 0; i < nhi; i++) {
        hf[i] = x;
        hi[i] = y;
        x = F::mul(x, step);
        y = F::mul(y, stepi);
    }
    int rc;
    for (int plan = 0; plan < LDE_PLANS; plan++) {  // the inter-round twiddles depend on the plan's radix split
        std::vector<T> rf(RTW_ENTRIES), ri(RTW_ENTRIES);
        fill_round_twiddles(plan, sizeof(T) == 8 ? FIELD_GL : FIELD_BB, (uint64_t)w, rf.data());
        fill_round_twiddles(plan, sizeof(T) == 8 ? FIELD_GL : FIELD_BB, (uint64_t)wi, ri.data());
        if ((rc = upload_vec<F>(ctx, rf, &ts.round_fwd[plan]))) return rc;
        if ((rc = upload_vec<F>(ctx, ri, &ts.round_inv[plan]))) return rc;
    }
    if ((rc = upload_vec<F>(ctx, tf, &ts.tile_fwd))) return rc;
    if ((rc = upload_vec<F>(ctx, ti, &ts.tile_inv))) return rc;
    if ((rc = upload_vec<F>(ctx, lf, &ts.top_lo_fwd))) return rc;
    if ((rc = upload_vec<F>(ctx, hf, &ts.top_hi_fwd))) return rc;
    if ((rc = upload_vec<F>(ctx, li, &ts.top_lo_inv))) return rc;
    if ((rc = upload_vec<F>(ctx, hi, &ts.top_hi_inv))) return rc;
    return ZKG_OK;
}

template <class F>
int build_coset(zkg_ctx *ctx, int log_n, int log_b, uint64_t shift_canon, CosetSet &cs)
{
    using T = typename F::T;
    const int B = 1 << log_b, nlo = 1 << TW_SPLIT, nhi = 1 << (LOG_TW_TOP - TW_SPLIT - 2);
    std::vector<T> lo((size_t)B * nlo), hi((size_t)B * nhi);
    T shift = F::from_canon((T)shift_canon);
    T wbig = root_of_unity<F>(log_n + log_b);
    // 1/N
    T ninv = finv<F>(F::from_canon((T)(1ull << log_n)));
    for (int b = 0; b < B; b++) {
        T sb = F::mul(shift, hpow<F>(wbig, rev_bits((uint32_t)b, log_b)));
        T acc = ninv;
        for (int i = 0; i < nlo; i++) {
            lo[(size_t)b * nlo + i] = acc;
            acc = F::mul(acc, sb);
        }
        T step = hpow<F>(sb, (uint64_t)nlo), h = F::one();
        for (int i = 0; i < nhi; i++) {
            hi[(size_t)b * nhi + i] = h;
            h = F::mul(h, step);
        }
    }
    int rc;
    if ((rc = upload_vec<F>(ctx, lo, &cs.lo))) return rc;
    if ((rc = upload_vec<F>(ctx, hi, &cs.hi))) return rc;
    cs.log_n = log_n;
    cs.log_blowup = log_b;
    return ZKG_OK;
}

}  // namespace

int zkg::get_coset(zkg_ctx *ctx, int field, int log_n, int log_b, uint64_t shift, const CosetSet **out)
{
    std::lock_guard<std::mutex> g(ctx->table_mu);
    CosetKey key{field, log_n, log_b, shift};
    auto it = ctx->cosets.find(key);
    if (it == ctx->cosets.end()) {
        CosetSet cs;
        int rc = field == FIELD_GL ? build_coset<Gl>(ctx, log_n, log_b, shift, cs)
                                   : build_coset<Bb>(ctx, log_n, log_b, shift, cs);
        if (rc) return rc;
        it = ctx->cosets.emplace(key, cs).first;
    }
    *out = &it->second;
    return ZKG_OK;
}

const P2Set *zkg::get_p2(zkg_ctx *ctx, int field, int width)
{
    int i = p2_index(field, width);
    if (i < 0 || !ctx->p2[i].loaded) return nullptr;
    return &ctx->p2[i];
}

namespace {

// Uploads one constant set (register form; BabyBear also as canonical doubles
// for the FP64 permutation) into `p`, which the caller owns.
template <class F>
int p2_build(zkg_ctx *ctx, int width, const std::vector<uint64_t> &flat_canon, const uint64_t *diag_canon, P2Set &p)
{
    using T = typename F::T;
    std::vector<T> rc(flat_canon.size()), dg(width);
    for (size_t i = 0; i < flat_canon.size(); i++) rc[i] = F::from_canon((T)flat_canon[i]);
    for (int i = 0; i < width; i++) dg[i] = F::from_canon((T)diag_canon[i]);
    p = P2Set();
    int r;
    if ((r = upload_vec<F>(ctx, rc, &p.rc))) return r;
    if ((r = upload_vec<F>(ctx, dg, &p.diag))) return r;
    if constexpr (sizeof(typename F::T) == 4) {
        std::vector<double> rf(flat_canon.size());
        for (size_t i = 0; i < flat_canon.size(); i++) rf[i] = (double)flat_canon[i];
        if ((r = dev_alloc(ctx, rf.size() * sizeof(double), &p.rc_f64))) return r;
        HIP_TRY(hipMemcpy(p.rc_f64, rf.data(), rf.size() * sizeof(double), hipMemcpyHostToDevice), "upload rc");
    }
    p.loaded = true;
    return ZKG_OK;
}

void p2_release(zkg_ctx *ctx, P2Set &p)
{
    dev_free(ctx, p.rc);
    dev_free(ctx, p.diag);
    dev_free(ctx, p.rc_f64);
    p = P2Set();
}

void bb_diag16(uint64_t *d)
{
    const uint64_t p = Bb::P;
    auto canon_pow_inv2 = [](int e) -> uint64_t {
        uint32_t inv2 = Bb::from_canon((Bb::P + 1) / 2);
        return Bb::to_canon(hpow<Bb>(inv2, e));
    };
    uint64_t i1 = canon_pow_inv2(1);
    uint64_t v[16] = {p - 2, 1, 2, i1, 3, 4, p - i1, p - 3, p - 4, canon_pow_inv2(8), canon_pow_inv2(2),
                      canon_pow_inv2(3), canon_pow_inv2(27), p - canon_pow_inv2(8), p - canon_pow_inv2(4),
                      p - canon_pow_inv2(27)};
    memcpy(d, v, sizeof v);
}

// A constant set for (field, width) from a flat table, diagonal per instance.
int p2_build_flat(zkg_ctx *ctx, int field, int width, const std::vector<uint64_t> &flat, P2Set &out)
{
    uint64_t diag[16];
    if (field == FIELD_GL)
        memcpy(diag, width == 16 ? GL_DIAG16 : GL_DIAG8, width * sizeof(uint64_t));
    else
        bb_diag16(diag);
    int rc = field == FIELD_GL ? p2_build<Gl>(ctx, width, flat, diag, out) : p2_build<Bb>(ctx, width, flat, diag, out);
    if (rc) p2_release(ctx, out);
    return rc;
}

// Replaces the context's resident set for (field, width).  The new tables are
// uploaded first; the old ones are freed only once nothing can still read
// them: every entry point has drained its streams when it returns, and the
// cached batch graph, whose kernels hold the old table addresses, is dropped.
// Refused while the serving loop runs (its slots hold launches that read the
// resident tables).
int p2_replace(zkg_ctx *ctx, int field, int width, const std::vector<uint64_t> &flat)
{
    if (ctx->serve) return set_err(ZKG_ESTATE, "Poseidon2 constants cannot change while the serving loop runs");
    P2Set fresh;
    if (int rc = p2_build_flat(ctx, field, width, flat, fresh)) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipStreamSynchronize(ctx->side[i]);
    if (e != hipSuccess) {
        p2_release(ctx, fresh);
        return hip_err(e, "sync before constant reload");
    }
    batch_release(ctx);
    P2Set &cur = ctx->p2[p2_index(field, width)];
    p2_release(ctx, cur);
    cur = fresh;
    ctx->p2_epoch++;
    return ZKG_OK;
}

}  // namespace

// ------------------------------------------------------------------ context API
extern "C" int zkg_ctx_create(int device, uint64_t budget, zkg_ctx **out)
{
    if (!out) return set_err(ZKG_EINVAL, "null out");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return set_err(ZKG_EINVAL, "device %d not present (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    if (int rc = noalloc_setup(device)) return rc;
    zkg_ctx *c = new zkg_ctx();
    c->device = device;
    c->hbm_budget = budget ? budget : (288ull << 30);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_err(e, "hipStreamCreate");
    }
    for (int i = 0; i < 8; i++) (void)hipEventCreate(&c->ev[i]);
    for (int i = 0; i < 2; i++) {
        (void)hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking);
        (void)hipEventCreateWithFlags(&c->side_join[i], hipEventDisableTiming);
    }
    for (int i = 0; i < 32; i++) (void)hipEventCreateWithFlags(&c->fri_ev[i], hipEventDisableTiming);
    int rc = build_twiddles<Gl>(c, c->tw[FIELD_GL]);
    if (!rc) rc = build_twiddles<Bb>(c, c->tw[FIELD_BB]);
    if (rc) {
        zkg_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return ZKG_OK;
}

extern "C" int zkg_ctx_destroy(zkg_ctx *ctx)
{
    if (!ctx) return ZKG_OK;
    serve_shutdown(ctx);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ce_cache_release(ctx);
    batch_release(ctx);
    std::vector<void *> ptrs;
    for (auto &kv : ctx->alloc_sizes) ptrs.push_back(kv.first);
    for (void *p : ptrs) dev_free(ctx, p);
    cache_release(ctx);
    for (int i = 0; i < 8; i++) (void)hipEventDestroy(ctx->ev[i]);
    for (int i = 0; i < 2; i++) {
        if (ctx->side[i]) {
            (void)hipStreamSynchronize(ctx->side[i]);
            (void)hipStreamDestroy(ctx->side[i]);
        }
        if (ctx->side_join[i]) (void)hipEventDestroy(ctx->side_join[i]);
    }
    for (int i = 0; i < 32; i++)
        if (ctx->fri_ev[i]) (void)hipEventDestroy(ctx->fri_ev[i]);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ZKG_OK;
}

extern "C" int zkg_ctx_synchronize(zkg_ctx *ctx)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return ZKG_OK;
}

extern "C" int zkg_ctx_hbm_used(const zkg_ctx *ctx, uint64_t *bytes)
{
    if (!ctx || !bytes) return set_err(ZKG_EINVAL, "null argument");
    *bytes = hbm_in_use(ctx);
    return ZKG_OK;
}

extern "C" int zkg_ctx_set_timing(zkg_ctx *ctx, int enable)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    ctx->timing = enable != 0;
    return ZKG_OK;
}

extern "C" int zkg_ctx_stage_ms(const zkg_ctx *ctx, double *ms4)
{
    if (!ctx || !ms4) return set_err(ZKG_EINVAL, "null argument");
    for (int i = 0; i < 4; i++) ms4[i] = ctx->stage_ms[i];
    return ZKG_OK;
}

// ------------------------------------------------------------------ buffers
extern "C" int zkg_buf_alloc(zkg_ctx *ctx, uint64_t bytes, zkg_buf **out)
{
    if (!ctx || !out) return set_err(ZKG_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    void *p;
    int rc = dev_alloc(ctx, bytes, &p);
    if (rc) return rc;
    zkg_buf *b = new zkg_buf{ctx, p, bytes};
    *out = b;
    return ZKG_OK;
}

extern "C" int zkg_buf_free(zkg_buf *b)
{
    if (!b) return ZKG_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    dev_free(b->ctx, b->ptr);
    delete b;
    return ZKG_OK;
}

extern "C" int zkg_buf_upload(zkg_buf *b, uint64_t off, const void *src, uint64_t bytes)
{
    if (!b || (!src && bytes)) return set_err(ZKG_EINVAL, "null argument");
    if (!range_fits(off, bytes, b->bytes)) return set_err(ZKG_EINVAL, "upload of %llu bytes at %llu exceeds buffer of %llu",
                                               (unsigned long long)bytes, (unsigned long long)off,
                                               (unsigned long long)b->bytes);
    HIP_TRY(hipStreamSynchronize(b->ctx->stream), "sync");
    HIP_TRY(hipMemcpy((char *)b->ptr + off, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return ZKG_OK;
}

extern "C" int zkg_buf_download(const zkg_buf *b, uint64_t off, void *dst, uint64_t bytes)
{
    if (!b || (!dst && bytes)) return set_err(ZKG_EINVAL, "null argument");
    if (!range_fits(off, bytes, b->bytes)) return set_err(ZKG_EINVAL, "download exceeds buffer");
    HIP_TRY(hipStreamSynchronize(b->ctx->stream), "sync");
    HIP_TRY(hipMemcpy(dst, (const char *)b->ptr + off, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return ZKG_OK;
}

extern "C" int zkg_buf_size(const zkg_buf *b, uint64_t *bytes)
{
    if (!b || !bytes) return set_err(ZKG_EINVAL, "null argument");
    *bytes = b->bytes;
    return ZKG_OK;
}

static uint64_t esize(int field) { return field == FIELD_GL ? 8 : 4; }

static int check_field(int field)
{
    if (field != FIELD_GL && field != FIELD_BB) return set_err(ZKG_EINVAL, "unknown field %d", field);
    return ZKG_OK;
}

extern "C" int zkg_fill_random(zkg_ctx *ctx, int field, zkg_buf *buf, uint64_t seed, uint64_t start, uint64_t n)
{
    if (!ctx || !buf) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_field(field)) return rc;
    if (span_bytes(n, esize(field)) > buf->bytes) return set_err(ZKG_EINVAL, "buffer too small");
    HIP_TRY(launch_gen(field, buf->ptr, seed, start, n, ctx->stream), "gen");
    return ZKG_OK;
}

// ------------------------------------------------------------------ constants
// RNG order of Poseidon2::new_from_rng_128 / flatten_round_constants:
// initial rf/2 x W, terminal rf/2 x W, internal rp; flat layout
// [initial][internal][terminal].
int zkg::p2_seed_constants(int field, int width, uint64_t seed, std::vector<uint64_t> &flat)
{
    if (p2_index(field, width) < 0) return set_err(ZKG_EINVAL, "unsupported Poseidon2 field %d width %d", field, width);
    SmallRng r(seed);
    const int half = 4, rp = rounds_p_of(field);
    std::vector<uint64_t> init(half * width), term(half * width), internal(rp);
    for (auto &v : init) v = sample_canon(r, field);
    for (auto &v : term) v = sample_canon(r, field);
    for (auto &v : internal) v = sample_canon(r, field);
    flat.clear();
    flat.insert(flat.end(), init.begin(), init.end());
    flat.insert(flat.end(), internal.begin(), internal.end());
    flat.insert(flat.end(), term.begin(), term.end());
    return ZKG_OK;
}

int zkg::p2_internal_rounds(int field) { return rounds_p_of(field); }

extern "C" int zkg_constants_load(zkg_ctx *ctx, int field, int width, uint64_t seed)
{
    if (!ctx) return set_err(ZKG_EINVAL, "null ctx");
    std::vector<uint64_t> flat;
    if (int rc = p2_seed_constants(field, width, seed, flat)) return rc;
    int rc = p2_replace(ctx, field, width, flat);
    if (!rc) ctx->p2[p2_index(field, width)].seed = seed;
    return rc;
}

// An explicit flat table [initial][internal][terminal] checked against the
// compiled schedule (8 external rounds, the field's internal rounds) and for
// canonical values.
static int p2_flat_from_caller(int field, int width, const void *flat, uint32_t n_ext, uint32_t n_int,
                               std::vector<uint64_t> &v)
{
    if (p2_index(field, width) < 0) return set_err(ZKG_EINVAL, "unsupported Poseidon2 field %d width %d", field, width);
    if (n_ext != 8 || (int)n_int != rounds_p_of(field))
        return set_err(ZKG_EINVAL, "rounds (%u, %u) differ from the compiled schedule (8, %d)", n_ext, n_int,
                       rounds_p_of(field));
    size_t n = (size_t)n_ext * width + n_int;
    v.resize(n);
    for (size_t i = 0; i < n; i++) {
        uint64_t x = field == FIELD_GL ? ((const uint64_t *)flat)[i] : ((const uint32_t *)flat)[i];
        if (x >= (field == FIELD_GL ? Gl::P : (uint64_t)Bb::P)) return set_err(ZKG_EINVAL, "constant %zu not canonical", i);
        v[i] = x;
    }
    return ZKG_OK;
}

extern "C" int zkg_constants_set(zkg_ctx *ctx, int field, int width, const void *flat, uint32_t n_ext, uint32_t n_int)
{
    if (!ctx || !flat) return set_err(ZKG_EINVAL, "null argument");
    std::vector<uint64_t> v;
    if (int rc = p2_flat_from_caller(field, width, flat, n_ext, n_int, v)) return rc;
    return p2_replace(ctx, field, width, v);
}

// ------------------------------------------------------------------ transforms
extern "C" int zkg_ntt_batch(zkg_ctx *ctx, int field, zkg_buf *buf, uint64_t height, uint32_t width, int inverse)
{
    if (!ctx || !buf) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_field(field)) return rc;
    if (!is_pow2(height) || ilog2u(height) > 22) return set_err(ZKG_EINVAL, "height %llu not a power of two <= 2^22", (unsigned long long)height);
    if (span_bytes(height, width, esize(field)) > buf->bytes) return set_err(ZKG_EINVAL, "buffer too small");
    if (height == 1 || width == 0) return ZKG_OK;
    uint64_t tb = ntt_tmp_bytes(field, height, width);
    void *tmp = nullptr;
    if (tb) {
        int rc = dev_alloc(ctx, tb, &tmp);
        if (rc) return rc;
    }
    hipError_t e = launch_ntt(field, ctx->tw[field], buf->ptr, tmp, height, width, inverse, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, tmp);
    if (e != hipSuccess) return hip_err(e, "ntt");
    return ZKG_OK;
}

int zkg::run_lde(zkg_ctx *ctx, hipStream_t st, int field, const void *in, uint64_t height, uint32_t width,
                 uint32_t log_blowup, uint64_t shift, void *out, void *tmp)
{
    const CosetSet *cs;
    int rc = get_coset(ctx, field, ilog2u(height), log_blowup, shift, &cs);
    if (rc) return rc;
    HIP_TRY(launch_coset_lde(field, ctx->tw[field], *cs, in, out, tmp, height, width, log_blowup, st), "coset lde");
    return ZKG_OK;
}

static int check_lde_shape(int field, uint64_t height, uint32_t width, uint32_t log_blowup)
{
    if (int rc = check_field(field)) return rc;
    if (!is_pow2(height) || ilog2u(height) > 22)
        return set_err(ZKG_EINVAL, "height %llu not a power of two <= 2^22", (unsigned long long)height);
    if (log_blowup > 2 || ilog2u(height) + (int)log_blowup > LOG_TW_TOP)
        return set_err(ZKG_EINVAL, "log_blowup %u unsupported (0..2)", log_blowup);
    if (width == 0) return set_err(ZKG_EINVAL, "width 0");
    return ZKG_OK;
}

extern "C" int zkg_coset_lde_batch(zkg_ctx *ctx, int field, const zkg_buf *in, uint64_t height, uint32_t width,
                                   uint32_t log_blowup, uint64_t shift, zkg_buf *out)
{
    if (!ctx || !in || !out) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_lde_shape(field, height, width, log_blowup)) return rc;
    uint64_t es = esize(field);
    if (span_bytes(height, width, es) > in->bytes || span_bytes(height << log_blowup, width, es) > out->bytes)
        return set_err(ZKG_EINVAL, "buffer too small");
    if (shift >= (field == FIELD_GL ? Gl::P : (uint64_t)Bb::P) || shift == 0)
        return set_err(ZKG_EINVAL, "shift must be a nonzero canonical element");
    uint64_t tb = coset_lde_tmp_bytes(field, height, width, log_blowup);
    void *tmp = nullptr;
    if (tb) {
        int rc = dev_alloc(ctx, tb, &tmp);
        if (rc) return rc;
    }
    int rc = run_lde(ctx, ctx->stream, field, in->ptr, height, width, log_blowup, shift, out->ptr, tmp);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, tmp);
    if (rc) return rc;
    if (e != hipSuccess) return hip_err(e, "lde sync");
    return ZKG_OK;
}

// ------------------------------------------------------------------ Poseidon2
extern "C" int zkg_poseidon2_permute(zkg_ctx *ctx, int field, int width, zkg_buf *states, uint64_t n)
{
    if (!ctx || !states) return set_err(ZKG_EINVAL, "null argument");
    const P2Set *p2 = get_p2(ctx, field, width);
    if (!p2) return set_err(ZKG_ESTATE, "Poseidon2 constants for field %d width %d not loaded", field, width);
    if (span_bytes(n, width, esize(field)) > states->bytes) return set_err(ZKG_EINVAL, "buffer too small");
    HIP_TRY(launch_p2_permute(field, width, *p2, states->ptr, n, ctx->stream), "permute");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return ZKG_OK;
}

// The trait's explicit-constant form: the caller's table goes into a set of
// its own for this call only, so the context's resident constants (which
// commits, MMCS and FRI use) are left as they were.
template <class T>
static int permute_host(zkg_ctx *ctx, int field, T *states, uint64_t n, const T *rc, uint32_t n_ext, uint32_t n_int)
{
    if (!ctx || (!states && n) || !rc) return set_err(ZKG_EINVAL, "null argument");
    std::vector<uint64_t> flat;
    int r = p2_flat_from_caller(field, 16, rc, n_ext, n_int, flat);
    if (r) return r;
    if (!n) return ZKG_OK;
    const uint64_t bytes = span_bytes(n, 16, sizeof(T));
    if (bytes > ctx->hbm_budget) return set_err(ZKG_EINVAL, "%llu states exceed the HBM budget", (unsigned long long)n);
    P2Set once;
    if ((r = p2_build_flat(ctx, field, 16, flat, once))) return r;
    void *d;
    if ((r = dev_alloc(ctx, bytes, &d))) {
        p2_release(ctx, once);
        return r;
    }
    hipError_t e = hipMemcpy(d, states, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_p2_permute(field, 16, once, d, n, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(states, d, bytes, hipMemcpyDeviceToHost);
    dev_free(ctx, d);
    p2_release(ctx, once);
    if (e != hipSuccess) return hip_err(e, "batch permute");
    return ZKG_OK;
}

extern "C" int zkg_poseidon2_batch_permute_goldilocks(zkg_ctx *ctx, uint64_t *states, uint64_t n,
                                                      const uint64_t *rc, uint32_t n_ext, uint32_t n_int)
{
    return permute_host<uint64_t>(ctx, FIELD_GL, states, n, rc, n_ext, n_int);
}

extern "C" int zkg_poseidon2_batch_permute_babybear(zkg_ctx *ctx, uint32_t *states, uint64_t n, const uint32_t *rc,
                                                    uint32_t n_ext, uint32_t n_int)
{
    return permute_host<uint32_t>(ctx, FIELD_BB, states, n, rc, n_ext, n_int);
}

extern "C" int zkg_hash_rows(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *mat, uint64_t height,
                             uint32_t mat_width, zkg_buf *digests)
{
    if (!ctx || !mat || !digests) return set_err(ZKG_EINVAL, "null argument");
    const P2Set *p2 = get_p2(ctx, field, p2_width);
    if (!p2) return set_err(ZKG_ESTATE, "Poseidon2 constants for field %d width %d not loaded", field, p2_width);
    uint64_t es = esize(field);
    if (span_bytes(height, mat_width, es) > mat->bytes || span_bytes(height, p2_width / 2, es) > digests->bytes)
        return set_err(ZKG_EINVAL, "buffer too small");
    HIP_TRY(launch_hash_rows(field, p2_width, *p2, mat->ptr, height, mat_width, digests->ptr, ctx->stream), "hash rows");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return ZKG_OK;
}

// ------------------------------------------------------------------ trees
int zkg::tree_alloc(zkg_ctx *ctx, int field, int p2_width, uint64_t n_leaves, const void *mat, uint32_t mat_width,
                    zkg_tree **out)
{
    zkg_tree *t = new zkg_tree();
    t->ctx = ctx;
    t->field = field;
    t->p2_width = p2_width;
    t->digest = p2_width / 2;
    t->n_leaves = n_leaves;
    t->nlayers = ilog2u(n_leaves) + 1;
    t->mat = mat;
    t->mat_width = mat_width;
    uint64_t off = 0, es = esize(field);
    for (int i = 0; i < t->nlayers; i++) {
        t->layer_off.push_back(off);
        t->layer_len.push_back(n_leaves >> i);
        off += (n_leaves >> i) * t->digest * es;
        off = (off + 255) & ~255ull;
    }
    int rc = dev_alloc(ctx, off, &t->base);
    if (rc) {
        delete t;
        return rc;
    }
    for (int i = 0; i < t->nlayers; i++) t->layers.push_back((char *)t->base + t->layer_off[i]);
    *out = t;
    return ZKG_OK;
}

int zkg::tree_build(zkg_ctx *ctx, hipStream_t st, zkg_tree *t, bool plain)
{
    const P2Set *p2 = get_p2(ctx, t->field, t->p2_width);
    HIP_TRY(launch_hash_rows(t->field, t->p2_width, *p2, t->mat, t->n_leaves, t->mat_width, t->layers[0], st),
            "hash rows");
    if (plain) HIP_TRY(hipStreamSynchronize(st), "sync");
    HIP_TRY(launch_merkle_tree(t->field, t->p2_width, *p2, t->layers.data(), t->nlayers, t->n_leaves, plain ? 1 : 9, st,
                               plain),
            "merkle");
    return ZKG_OK;
}

extern "C" int zkg_merkle_commit(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *mat, uint64_t height,
                                 uint32_t width, zkg_tree **out)
{
    if (!ctx || !mat || !out) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_field(field)) return rc;
    if (!get_p2(ctx, field, p2_width))
        return set_err(ZKG_ESTATE, "Poseidon2 constants for field %d width %d not loaded", field, p2_width);
    if (!is_pow2(height) || ilog2u(height) > 31) return set_err(ZKG_EINVAL, "height must be a power of two");
    if (span_bytes(height, width, esize(field)) > mat->bytes) return set_err(ZKG_EINVAL, "buffer too small");
    zkg_tree *t;
    int rc = tree_alloc(ctx, field, p2_width, height, mat->ptr, width, &t);
    if (rc) return rc;
    rc = tree_build(ctx, ctx->stream, t, false);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (!rc && e != hipSuccess) rc = hip_err(e, "merkle sync");
    if (rc) {
        zkg_tree_free(t);
        return rc;
    }
    *out = t;
    return ZKG_OK;
}

extern "C" int zkg_tree_root(const zkg_tree *t, void *root)
{
    if (!t || !root) return set_err(ZKG_EINVAL, "null argument");
    HIP_TRY(hipStreamSynchronize(t->ctx->stream), "sync");
    HIP_TRY(hipMemcpy(root, t->layers.back(), t->digest * esize(t->field), hipMemcpyDeviceToHost), "root");
    return ZKG_OK;
}

extern "C" int zkg_tree_num_layers(const zkg_tree *t, uint32_t *n)
{
    if (!t || !n) return set_err(ZKG_EINVAL, "null argument");
    *n = (uint32_t)t->nlayers;
    return ZKG_OK;
}

extern "C" int zkg_tree_digest_elems(const zkg_tree *t, uint32_t *n)
{
    if (!t || !n) return set_err(ZKG_EINVAL, "null argument");
    *n = (uint32_t)t->digest;
    return ZKG_OK;
}

extern "C" int zkg_tree_layer_len(const zkg_tree *t, uint32_t layer, uint64_t *n)
{
    if (!t || !n) return set_err(ZKG_EINVAL, "null argument");
    if ((int)layer >= t->nlayers) return set_err(ZKG_EINVAL, "layer %u of %d", layer, t->nlayers);
    *n = t->layer_len[layer];
    return ZKG_OK;
}

extern "C" int zkg_tree_layer(const zkg_tree *t, uint32_t layer, void *out, uint64_t cap)
{
    if (!t || !out) return set_err(ZKG_EINVAL, "null argument");
    if ((int)layer >= t->nlayers) return set_err(ZKG_EINVAL, "layer %u of %d", layer, t->nlayers);
    uint64_t bytes = t->layer_len[layer] * t->digest * esize(t->field);
    if (bytes > cap) return set_err(ZKG_EINVAL, "output too small");
    HIP_TRY(hipStreamSynchronize(t->ctx->stream), "sync");
    HIP_TRY(hipMemcpy(out, t->layers[layer], bytes, hipMemcpyDeviceToHost), "layer");
    return ZKG_OK;
}

int zkg::tree_open(const zkg_tree *t, const uint64_t *idx, uint32_t n, void *rows_out, void *paths_out)
{
    if (t->mmcs) return set_err(ZKG_EINVAL, "tree has several matrices: use zkg_mmcs_open");
    if (!n) return ZKG_OK;
    for (uint32_t i = 0; i < n; i++)
        if (idx[i] >= t->n_leaves) return set_err(ZKG_EINVAL, "index %llu out of range %llu", (unsigned long long)idx[i],
                                                  (unsigned long long)t->n_leaves);
    zkg_ctx *ctx = t->ctx;
    uint64_t es = esize(t->field);
    uint64_t rb = (uint64_t)n * t->mat_width * es, pb = (uint64_t)n * (t->nlayers - 1) * t->digest * es;
    void *d_idx, *d_rows, *d_paths;
    int rc = dev_alloc(ctx, n * 8ull, &d_idx);
    if (rc) return rc;
    if ((rc = dev_alloc(ctx, rb, &d_rows))) {
        dev_free(ctx, d_idx);
        return rc;
    }
    if ((rc = dev_alloc(ctx, pb, &d_paths))) {
        dev_free(ctx, d_idx);
        dev_free(ctx, d_rows);
        return rc;
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(d_idx, idx, n * 8ull, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_gather_open(t->field, t->digest, t->mat, t->mat_width, t->layers.data(), t->nlayers,
                               (const uint64_t *)d_idx, n, d_rows, d_paths, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && rows_out && rb) e = hipMemcpy(rows_out, d_rows, rb, hipMemcpyDeviceToHost);
    if (e == hipSuccess && paths_out && pb) e = hipMemcpy(paths_out, d_paths, pb, hipMemcpyDeviceToHost);
    dev_free(ctx, d_idx);
    dev_free(ctx, d_rows);
    dev_free(ctx, d_paths);
    if (e != hipSuccess) return hip_err(e, "open");
    return ZKG_OK;
}

extern "C" int zkg_tree_open(const zkg_tree *t, const uint64_t *idx, uint32_t n, void *rows_out, void *paths_out)
{
    if (!t || (!idx && n)) return set_err(ZKG_EINVAL, "null argument");
    return tree_open(t, idx, n, rows_out, paths_out);
}

extern "C" int zkg_tree_free(zkg_tree *t)
{
    if (!t) return ZKG_OK;
    (void)hipStreamSynchronize(t->ctx->stream);
    dev_free(t->ctx, t->base);
    if (t->desc) dev_free(t->ctx, t->desc);
    delete t;
    return ZKG_OK;
}

// ------------------------------------------------------------------ mixed-height MMCS
static uint64_t next_pow2_u64(uint64_t x)
{
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// p3 MerkleTree::new over several matrices (vendor/p3-merkle-tree-gpu/src/
// merkle_tree.rs:84-151, first_digest_layer :153-213, compress_and_inject
// :272-350): matrices sorted tallest first (stable); the tallest group's
// concatenated rows are hashed into layer 0, padded with zero digests to even
// length unless it has one row; each next layer has plen/2 compressions,
// padded to even (a single root when plen = 2); a group whose heights round up
// to next_pow2(plen/2) is injected there.
extern "C" int zkg_mmcs_commit(zkg_ctx *ctx, int field, int p2_width, uint32_t n_mats, const zkg_buf *const *mats,
                               const uint64_t *heights, const uint32_t *widths, zkg_tree **out)
{
    if (!ctx || !mats || !heights || !widths || !out || !n_mats) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_field(field)) return rc;
    const P2Set *p2 = get_p2(ctx, field, p2_width);
    if (!p2) return set_err(ZKG_ESTATE, "Poseidon2 constants for field %d width %d not loaded", field, p2_width);
    const uint64_t es = esize(field);
    for (uint32_t i = 0; i < n_mats; i++) {
        if (!mats[i]) return set_err(ZKG_EINVAL, "matrix %u is null", i);
        if (!heights[i] || heights[i] > (1ull << 31)) return set_err(ZKG_EINVAL, "matrix %u: height %llu", i,
                                                                    (unsigned long long)heights[i]);
        if (span_bytes(heights[i], widths[i], es) > mats[i]->bytes) return set_err(ZKG_EINVAL, "matrix %u: buffer too small", i);
    }
    std::vector<uint32_t> ord(n_mats);
    for (uint32_t i = 0; i < n_mats; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return heights[a] > heights[b]; });
    for (uint32_t i = 1; i < n_mats; i++) {
        uint64_t a = heights[ord[i - 1]], b = heights[ord[i]];
        if (a != b && next_pow2_u64(a) == next_pow2_u64(b))
            return set_err(ZKG_EINVAL, "heights %llu and %llu round up to the same power of two",
                           (unsigned long long)a, (unsigned long long)b);
    }
    // layer lengths and the matrix group injected at each layer
    const uint64_t maxh = heights[ord[0]];
    std::vector<uint64_t> len{maxh == 1 ? 1 : maxh + (maxh & 1)};
    std::vector<std::pair<uint32_t, uint32_t>> group{{0, 0}};  // [first, count) into ord, per layer
    uint32_t cursor = 0;
    while (cursor < n_mats && heights[ord[cursor]] == maxh) cursor++;
    group[0] = {0, cursor};
    while (len.back() != 1) {
        const uint64_t plen = len.back(), np2 = next_pow2_u64(plen / 2);
        uint32_t k = 0;
        while (cursor + k < n_mats && next_pow2_u64(heights[ord[cursor + k]]) == np2) k++;
        group.push_back({cursor, k});
        cursor += k;
        len.push_back(plen == 2 ? 1 : ((plen / 2 + 1) & ~1ull));
    }
    if (cursor != n_mats) return set_err(ZKG_EINVAL, "a matrix height matches no tree layer");

    zkg_tree *t = new zkg_tree();
    t->ctx = ctx;
    t->field = field;
    t->p2_width = p2_width;
    t->digest = p2_width / 2;
    t->n_leaves = maxh;
    t->nlayers = (int)len.size();
    t->mmcs = true;
    t->max_height = maxh;
    for (uint32_t i = 0; i < n_mats; i++) {
        t->mats.push_back(MatDesc{mats[i]->ptr, widths[i]});
        t->mat_heights.push_back(heights[i]);
    }
    uint64_t off = 0;
    for (int i = 0; i < t->nlayers; i++) {
        t->layer_off.push_back(off);
        t->layer_len.push_back(len[i]);
        off += len[i] * t->digest * es;
        off = (off + 255) & ~255ull;
    }
    std::vector<MatDesc> desc;
    for (uint32_t i = 0; i < n_mats; i++) desc.push_back(MatDesc{mats[ord[i]]->ptr, widths[ord[i]]});
    int rc = dev_alloc(ctx, off, &t->base);
    if (!rc) rc = dev_alloc(ctx, desc.size() * sizeof(MatDesc), &t->desc);
    if (rc) {
        zkg_tree_free(t);
        return rc;
    }
    for (int i = 0; i < t->nlayers; i++) t->layers.push_back((char *)t->base + t->layer_off[i]);
    const MatDesc *d = (const MatDesc *)t->desc;
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(t->desc, desc.data(), desc.size() * sizeof(MatDesc), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(t->base, 0, off, st);  // zero-digest padding
    // A tallest group of one matrix is the single-matrix row hash; a run of
    // layers with nothing injected and power-of-two lengths is a plain tree
    // segment.  Both take the single-matrix kernels; other layers compress and
    // inject one node per lane.
    if (e == hipSuccess) {
        if (group[0].second == 1)
            e = launch_hash_rows(field, p2_width, *p2, mats[ord[0]]->ptr, maxh, widths[ord[0]], t->layers[0], st);
        else
            e = launch_mmcs_leaves(field, p2_width, *p2, d + group[0].first, (int)group[0].second, maxh,
                                   t->layers[0], st);
    }
    auto plain = [&](int i) { return group[i].second == 0 && is_pow2(len[i - 1]) && len[i] * 2 == len[i - 1]; };
    for (int i = 1; i < t->nlayers && e == hipSuccess;) {
        if (plain(i)) {
            int j = i + 1;
            while (j < t->nlayers && plain(j)) j++;
            e = launch_merkle_tree(field, p2_width, *p2, &t->layers[i - 1], j - i + 1, len[i - 1], 9, st, false);
            i = j;
            continue;
        }
        const uint32_t g0 = group[i].first, gn = group[i].second;
        e = launch_mmcs_layer(field, p2_width, *p2, t->layers[i - 1], len[i - 1] / 2, t->layers[i], d + g0, (int)gn,
                              gn ? heights[ord[g0]] : 0, st);
        i++;
    }
    // the host descriptor vector goes out of scope: finish the upload first
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        zkg_tree_free(t);
        return hip_err(e, "mmcs commit");
    }
    *out = t;
    return ZKG_OK;
}

static int log2_ceil_u64(uint64_t x)
{
    int l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

// The pieces of a batch of openings (matrix rows, sibling digests) gathered
// on the device in one launch into one block [0, na + nb), whose two parts are
// then copied to a and b.  Piece: `bytes` (a multiple of 4) from device
// address src to offset off of the block.
struct OpenPiece {
    const void *src;
    uint64_t off;
    uint32_t bytes;
};

static int gather_openings(zkg_ctx *ctx, const std::vector<OpenPiece> &pieces, uint64_t na, void *a, uint64_t nb,
                           void *b)
{
    if (pieces.empty() || na + nb == 0) return ZKG_OK;
    void *d_out = nullptr, *d_segs = nullptr;
    int rc = dev_alloc(ctx, na + nb, &d_out);
    if (!rc) rc = dev_alloc(ctx, pieces.size() * sizeof(GatherSeg), &d_segs);
    hipError_t e = hipSuccess;
    if (!rc) {
        std::vector<GatherSeg> segs(pieces.size());
        for (size_t i = 0; i < pieces.size(); i++)
            segs[i] = GatherSeg{(const uint32_t *)pieces[i].src, (uint32_t *)((uint8_t *)d_out + pieces[i].off),
                                pieces[i].bytes / 4, 0};
        e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(d_segs, segs.data(), segs.size() * sizeof(GatherSeg), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = launch_gather_segs((const GatherSeg *)d_segs, (uint32_t)segs.size(), ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && na) e = hipMemcpy(a, d_out, na, hipMemcpyDeviceToHost);
        if (e == hipSuccess && nb) e = hipMemcpy(b, (uint8_t *)d_out + na, nb, hipMemcpyDeviceToHost);
    }
    if (d_out) dev_free(ctx, d_out);
    if (d_segs) dev_free(ctx, d_segs);
    if (rc) return rc;
    if (e != hipSuccess) return hip_err(e, "gather openings");
    return ZKG_OK;
}

// Mmcs::open_batch (vendor/p3-merkle-tree-gpu/src/mmcs.rs:179-210): matrix j
// opens row index >> (log2_ceil(max_h) - log2_ceil(h_j)); the proof is the
// sibling digest of every layer below the root.
extern "C" int zkg_mmcs_open(const zkg_tree *t, const uint64_t *indices, uint32_t n, void *rows_out, void *paths_out)
{
    if (!t || (!indices && n)) return set_err(ZKG_EINVAL, "null argument");
    std::vector<MatDesc> mats = t->mats;
    std::vector<uint64_t> hs = t->mat_heights;
    if (!t->mmcs) {
        mats = {MatDesc{t->mat, t->mat_width}};
        hs = {t->n_leaves};
    }
    const uint64_t maxh = t->mmcs ? t->max_height : t->n_leaves, es = esize(t->field);
    const int logmax = log2_ceil_u64(maxh);
    uint64_t row_elems = 0;
    for (auto &m : mats) row_elems += m.width;
    for (uint32_t q = 0; q < n; q++) {
        const uint64_t index = indices[q];
        if (index >= maxh)
            return set_err(ZKG_EINVAL, "index %llu out of range %llu", (unsigned long long)index,
                           (unsigned long long)maxh);
        for (size_t j = 0; j < mats.size() && rows_out; j++)
            if ((index >> (logmax - log2_ceil_u64(hs[j]))) >= hs[j])  // a padding position of a shorter matrix
                return set_err(ZKG_EINVAL, "index %llu has no row in matrix %zu (height %llu)",
                               (unsigned long long)index, j, (unsigned long long)hs[j]);
    }
    if (!n || (!rows_out && !paths_out)) return ZKG_OK;
    // every piece of every opening in one gather: [rows | paths]
    const uint64_t row_bytes = row_elems * es, path_bytes = (uint64_t)logmax * t->digest * es;
    const uint64_t rows_total = rows_out ? n * row_bytes : 0, paths_total = paths_out ? n * path_bytes : 0;
    std::vector<OpenPiece> pieces;
    pieces.reserve((size_t)n * (mats.size() + logmax));
    for (uint32_t q = 0; q < n; q++) {
        const uint64_t index = indices[q];
        uint64_t off = q * row_bytes;
        for (size_t j = 0; j < mats.size() && rows_out; j++) {
            const uint64_t r = index >> (logmax - log2_ceil_u64(hs[j]));
            const uint64_t rb = mats[j].width * es;
            if (rb) pieces.push_back(OpenPiece{(const uint8_t *)mats[j].ptr + r * rb, off, (uint32_t)rb});
            off += rb;
        }
        for (int i = 0; i < logmax && paths_out; i++) {
            const uint64_t sib = (index >> i) ^ 1;
            pieces.push_back(OpenPiece{(const uint8_t *)t->layers[i] + sib * t->digest * es,
                                       rows_total + q * path_bytes + (uint64_t)i * t->digest * es,
                                       (uint32_t)(t->digest * es)});
        }
    }
    return gather_openings(t->ctx, pieces, rows_total, rows_out, paths_total, paths_out);
}

// Mmcs::verify_batch (vendor/p3-merkle-tree-gpu/src/mmcs.rs:210-330) for n
// openings on the device.  The host checks what p3 refuses with an error and
// regroups the rows per matrix ([opening][width], tallest matrix first, ties
// in the given order) for the verify kernel.
extern "C" int zkg_mmcs_verify_batch(zkg_ctx *ctx, int field, int p2_width, const void *root, uint32_t n_mats,
                                     const uint64_t *heights, const uint32_t *widths, uint32_t n,
                                     const uint64_t *indices, const void *rows, const void *paths, uint8_t *ok)
{
    if (!ctx || !root || !heights || !widths || (n && (!indices || !rows || !paths || !ok)))
        return set_err(ZKG_EINVAL, "null argument");
    if (!n_mats) return set_err(ZKG_EINVAL, "empty batch");
    if (int rc = check_field(field)) return rc;
    const P2Set *p2 = get_p2(ctx, field, p2_width);
    if (!p2) return set_err(ZKG_ESTATE, "Poseidon2 constants for field %d width %d not loaded", field, p2_width);
    if (!n) return ZKG_OK;
    const uint64_t es = esize(field), D = (uint64_t)p2_width / 2, P = field == FIELD_GL ? Gl::P : (uint64_t)Bb::P;
    for (uint32_t i = 0; i < n_mats; i++)
        if (!heights[i] || heights[i] > (1ull << 40)) return set_err(ZKG_EINVAL, "matrix %u: height %llu", i,
                                                                    (unsigned long long)heights[i]);
    std::vector<uint32_t> ord(n_mats);
    for (uint32_t i = 0; i < n_mats; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return heights[a] > heights[b]; });
    for (uint32_t i = 1; i < n_mats; i++) {
        const uint64_t a = heights[ord[i - 1]], b = heights[ord[i]];
        if (a != b && next_pow2_u64(a) == next_pow2_u64(b))
            return set_err(ZKG_EINVAL, "heights %llu and %llu round up to the same power of two",
                           (unsigned long long)a, (unsigned long long)b);
    }
    const uint64_t maxh = heights[ord[0]];
    const int logmax = log2_ceil_u64(maxh);
    // groups of equal padded height, tallest first, and the layer each enters at
    std::vector<int> gstart, glevel;
    for (uint32_t i = 0; i < n_mats; i++)
        if (!i || next_pow2_u64(heights[ord[i]]) != next_pow2_u64(heights[ord[i - 1]])) {
            gstart.push_back((int)i);
            glevel.push_back(logmax - log2_ceil_u64(next_pow2_u64(heights[ord[i]])));
        }
    const int n_groups = (int)gstart.size();
    gstart.push_back((int)n_mats);
    uint64_t row_elems = 0;
    std::vector<uint64_t> col_off(n_mats);  // offset of matrix i within one opening's row block
    for (uint32_t i = 0; i < n_mats; i++) {
        col_off[i] = row_elems;
        row_elems += widths[i];
    }
    auto value = [&](const void *base, uint64_t k) -> uint64_t {
        return es == 8 ? ((const uint64_t *)base)[k] : ((const uint32_t *)base)[k];
    };
    for (uint32_t q = 0; q < n; q++)
        if (indices[q] >= maxh)
            return set_err(ZKG_EINVAL, "opening %u: index %llu out of bounds for height %llu", q,
                           (unsigned long long)indices[q], (unsigned long long)maxh);
    for (uint64_t k = 0; k < (uint64_t)n * row_elems; k++)
        if (value(rows, k) >= P) return set_err(ZKG_EINVAL, "opened value %llu is not canonical", (unsigned long long)k);
    for (uint64_t k = 0; k < (uint64_t)n * logmax * D; k++)
        if (value(paths, k) >= P) return set_err(ZKG_EINVAL, "sibling word %llu is not canonical", (unsigned long long)k);
    for (uint64_t k = 0; k < D; k++)
        if (value(root, k) >= P) return set_err(ZKG_EINVAL, "root is not canonical");
    // device layout: per-matrix row blocks, proofs, indices, root, flags, descriptors, group tables
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    std::vector<uint64_t> mat_off(n_mats);
    uint64_t off = 0;
    for (uint32_t i = 0; i < n_mats; i++) {
        mat_off[i] = off;
        off = al(off + (uint64_t)n * widths[ord[i]] * es);
    }
    const uint64_t proof_off = off, idx_off = al(proof_off + (uint64_t)n * logmax * D * es);
    const uint64_t root_off = al(idx_off + (uint64_t)n * 8), ok_off = al(root_off + D * es);
    const uint64_t desc_off = al(ok_off + (uint64_t)n * 4), gs_off = al(desc_off + n_mats * sizeof(MatDesc));
    const uint64_t gl_off = al(gs_off + gstart.size() * sizeof(int)), total = al(gl_off + glevel.size() * sizeof(int));
    std::vector<char> host(total, 0);
    for (uint32_t i = 0; i < n_mats; i++) {
        const uint32_t m = ord[i];
        const uint64_t rb = (uint64_t)widths[m] * es;
        for (uint32_t q = 0; q < n; q++)
            std::memcpy(&host[mat_off[i] + q * rb], (const char *)rows + (q * row_elems + col_off[m]) * es, rb);
    }
    std::memcpy(&host[proof_off], paths, (uint64_t)n * logmax * D * es);
    std::memcpy(&host[idx_off], indices, (uint64_t)n * 8);
    std::memcpy(&host[root_off], root, D * es);
    std::memcpy(&host[gs_off], gstart.data(), gstart.size() * sizeof(int));
    std::memcpy(&host[gl_off], glevel.data(), glevel.size() * sizeof(int));
    void *d = nullptr;
    if (int rc = dev_alloc(ctx, total, &d)) return rc;
    char *db = (char *)d;
    std::vector<MatDesc> desc(n_mats);
    for (uint32_t i = 0; i < n_mats; i++) desc[i] = MatDesc{db + mat_off[i], widths[ord[i]]};
    std::memcpy(&host[desc_off], desc.data(), desc.size() * sizeof(MatDesc));
    hipStream_t st = ctx->stream;
    std::vector<uint32_t> flags(n);
    hipError_t e = hipMemcpyAsync(d, host.data(), total, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_mmcs_verify(field, p2_width, *p2, (const MatDesc *)(db + desc_off), (const int *)(db + gs_off),
                               (const int *)(db + gl_off), n_groups, n, (const uint64_t *)(db + idx_off),
                               db + proof_off, logmax, db + root_off, (uint32_t *)(db + ok_off), st);
    if (e == hipSuccess) e = hipMemcpyAsync(flags.data(), db + ok_off, (uint64_t)n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(ctx, d);
    if (e != hipSuccess) return hip_err(e, "mmcs verify");
    for (uint32_t q = 0; q < n; q++) ok[q] = flags[q] ? 1 : 0;
    return ZKG_OK;
}

// ------------------------------------------------------------------ DEEP quotient
namespace {

struct Ef4H {
    uint32_t c[4];
};
Ef4H ef4h_mont(const uint32_t *canon)
{
    return Ef4H{{Bb::from_canon(canon[0]), Bb::from_canon(canon[1]), Bb::from_canon(canon[2]),
                 Bb::from_canon(canon[3])}};
}
Ef4H ef4h_mul(const Ef4H &a, const Ef4H &b)
{
    Ef4 r = ef4_mul(Ef4{{a.c[0], a.c[1], a.c[2], a.c[3]}}, Ef4{{b.c[0], b.c[1], b.c[2], b.c[3]}},
                    Bb::from_canon(11));
    return Ef4H{{r.c[0], r.c[1], r.c[2], r.c[3]}};
}
Ef4H ef4h_pow(Ef4H a, uint64_t e)
{
    Ef4H r{{Bb::one(), 0, 0, 0}};
    while (e) {
        if (e & 1) r = ef4h_mul(r, a);
        a = ef4h_mul(a, a);
        e >>= 1;
    }
    return r;
}
uint32_t bbh_pow(uint32_t a, uint64_t e)
{
    uint32_t r = Bb::one();
    while (e) {
        if (e & 1) r = Bb::mul(r, a);
        a = Bb::mul(a, a);
        e >>= 1;
    }
    return r;
}

uint64_t glh_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = Gl::mul(r, a);
        a = Gl::mul(a, a);
        e >>= 1;
    }
    return r;
}

// One DEEP request: BabyBear with EF4 challenges (4 x u32 each) or Goldilocks
// with base-field challenges (one u64 each).
struct DeepJob {
    int field = FIELD_BB;
    uint64_t height = 0;
    uint32_t width = 0;
    Ef4H z;              // BabyBear: Montgomery
    uint32_t g = 0;      // BabyBear shift, Montgomery
    uint64_t zg = 0, gg = 0;  // Goldilocks z and shift
    void *inv = nullptr;
    uint64_t es() const { return field == FIELD_GL ? 8 : 16; }  // bytes per challenge-field value
};

int deep_check(zkg_ctx *ctx, int field, const zkg_buf *lde, uint64_t height, uint32_t width, uint64_t shift,
               const void *z, DeepJob &j)
{
    if (!ctx || !lde || !z) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_field(field)) return rc;
    if (!is_pow2(height) || ilog2u(height) > LOG_TW_TOP) return set_err(ZKG_EINVAL, "LDE height must be 2^k <= 2^24");
    if (!width) return set_err(ZKG_EINVAL, "width 0");
    j.field = field;
    j.height = height;
    j.width = width;
    if (span_bytes(height, width, esize(field)) > lde->bytes) return set_err(ZKG_EINVAL, "buffer too small");
    if (field == FIELD_GL) {
        const uint64_t zz = *(const uint64_t *)z;
        if (zz >= Gl::P) return set_err(ZKG_EINVAL, "z is not canonical");
        if (shift % Gl::P == 0) return set_err(ZKG_EINVAL, "zero coset shift");
        j.zg = zz;
        j.gg = shift % Gl::P;
        if (glh_pow(j.zg, height) == glh_pow(j.gg, height)) return set_err(ZKG_EINVAL, "z lies on the LDE domain");
        return ZKG_OK;
    }
    const uint32_t *zw = (const uint32_t *)z;
    if (shift % Bb::P == 0) return set_err(ZKG_EINVAL, "zero coset shift");
    for (int i = 0; i < 4; i++)
        if (zw[i] >= Bb::P) return set_err(ZKG_EINVAL, "z is not canonical");
    j.z = ef4h_mont(zw);
    j.g = Bb::from_canon((uint32_t)(shift % Bb::P));
    // z must avoid the evaluation domain g H_height (x^height = g^height there)
    if (!zw[1] && !zw[2] && !zw[3] && bbh_pow(j.z.c[0], height) == bbh_pow(j.g, height))
        return set_err(ZKG_EINVAL, "z lies on the LDE domain");
    return ZKG_OK;
}

int deep_inv(zkg_ctx *ctx, DeepJob &j)
{
    int rc = dev_alloc(ctx, j.height * j.es(), &j.inv);
    if (rc) return rc;
    if (j.field == FIELD_GL)
        HIP_TRY(launch_deep_gl_inv(ctx->tw[FIELD_GL], j.inv, j.height, j.zg, j.gg, ctx->stream),
                "deep inverse denominators");
    else
        HIP_TRY(launch_deep_inv(ctx->tw[FIELD_BB], j.inv, j.height, j.z.c, j.g, Bb::from_canon(11), ctx->stream),
                "deep inverse denominators");
    return ZKG_OK;
}

int deep_values(zkg_ctx *ctx, const zkg_buf *lde, uint32_t log_blowup, DeepJob &j, void *ys_out)
{
    if (log_blowup > ilog2u(j.height)) return set_err(ZKG_EINVAL, "log_blowup exceeds log height");
    const uint64_t n = j.height >> log_blowup;
    void *partial = nullptr, *ys = nullptr;
    int rc = dev_alloc(ctx, deep_partial_blocks(n) * j.width * j.es(), &partial);
    if (!rc) rc = dev_alloc(ctx, (uint64_t)j.width * j.es(), &ys);
    hipError_t e = hipSuccess;
    if (!rc) {
        if (j.field == FIELD_GL) {
            // (z^n - g^n) / (n g^n)
            const uint64_t gn = glh_pow(j.gg, n);
            const uint64_t scale = Gl::mul(Gl::sub(glh_pow(j.zg, n), gn), glh_pow(Gl::mul(n % Gl::P, gn), Gl::P - 2));
            e = launch_deep_gl_open_values(ctx->tw[FIELD_GL], lde->ptr, j.height, j.width, n, j.inv, j.gg, scale,
                                           partial, ys, ctx->stream);
        } else {
            Ef4H zn = ef4h_pow(j.z, n);
            const uint32_t gn = bbh_pow(j.g, n);
            zn.c[0] = Bb::sub(zn.c[0], gn);
            const uint32_t sc = bbh_pow(Bb::mul(Bb::from_canon((uint32_t)(n % Bb::P)), gn), Bb::P - 2);
            Ef4H scale{{Bb::mul(zn.c[0], sc), Bb::mul(zn.c[1], sc), Bb::mul(zn.c[2], sc), Bb::mul(zn.c[3], sc)}};
            e = launch_deep_open_values(ctx->tw[FIELD_BB], lde->ptr, j.height, j.width, n, j.inv, j.z.c, j.g,
                                        Bb::from_canon(11), scale.c, partial, ys, ctx->stream);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(ys_out, ys, (uint64_t)j.width * j.es(), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    if (partial) dev_free(ctx, partial);
    if (ys) dev_free(ctx, ys);
    if (rc) return rc;
    if (e != hipSuccess) return hip_err(e, "deep open values");
    return ZKG_OK;
}

int deep_reduce(zkg_ctx *ctx, const zkg_buf *lde, const void *alpha, uint64_t alpha_offset, const void *ys,
                zkg_buf *ro, DeepJob &j)
{
    if (span_bytes(j.height, j.es()) > ro->bytes) return set_err(ZKG_EINVAL, "ro buffer too small");
    void *d_coef = nullptr;
    hipError_t e = hipSuccess;
    if (j.field == FIELD_GL) {
        const uint64_t a = *(const uint64_t *)alpha, *y = (const uint64_t *)ys;
        if (a >= Gl::P) return set_err(ZKG_EINVAL, "alpha is not canonical");
        for (uint32_t k = 0; k < j.width; k++)
            if (y[k] >= Gl::P) return set_err(ZKG_EINVAL, "opened values are not canonical");
        // c_k = alpha^off alpha^k, ry = sum_k c_k y_k
        std::vector<uint64_t> coef(j.width);
        uint64_t p = glh_pow(a, alpha_offset), ry = 0;
        for (uint32_t k = 0; k < j.width; k++) {
            coef[k] = p;
            ry = Gl::add(ry, Gl::mul(p, y[k]));
            p = Gl::mul(p, a);
        }
        int rc = dev_alloc(ctx, coef.size() * 8, &d_coef);
        if (rc) return rc;
        e = hipMemcpyAsync(d_coef, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = launch_deep_gl_reduce(lde->ptr, j.height, j.width, j.inv, d_coef, ry, ro->ptr, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // coef is host memory
    } else {
        const uint32_t *aw = (const uint32_t *)alpha, *yw = (const uint32_t *)ys;
        for (int i = 0; i < 4; i++)
            if (aw[i] >= Bb::P) return set_err(ZKG_EINVAL, "alpha is not canonical");
        for (uint64_t i = 0; i < (uint64_t)j.width * 4; i++)
            if (yw[i] >= Bb::P) return set_err(ZKG_EINVAL, "opened values are not canonical");
        const Ef4H a = ef4h_mont(aw);
        // c_k = alpha^off alpha^k (canonical, as 16-bit halves in doubles) and
        // ry = sum_k c_k y_k; Montgomery x canonical products come out canonical.
        Ef4H p = ef4h_pow(a, alpha_offset), ry{{0, 0, 0, 0}};
        std::vector<double> coef((size_t)j.width * 8);
        for (uint32_t k = 0; k < j.width; k++) {
            const Ef4H ck = ef4h_mul(p, Ef4H{{1, 0, 0, 0}});  // Montgomery -> canonical
            for (int c = 0; c < 4; c++) {
                coef[8 * k + c] = (double)(ck.c[c] & 0xFFFFu);
                coef[8 * k + 4 + c] = (double)(ck.c[c] >> 16);
            }
            Ef4H t = ef4h_mul(p, Ef4H{{yw[4 * k], yw[4 * k + 1], yw[4 * k + 2], yw[4 * k + 3]}});
            for (int c = 0; c < 4; c++) ry.c[c] = Bb::add(ry.c[c], t.c[c]);
            p = ef4h_mul(p, a);
        }
        int rc = dev_alloc(ctx, coef.size() * sizeof(double), &d_coef);
        if (rc) return rc;
        e = hipMemcpyAsync(d_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = launch_deep_reduce(lde->ptr, j.height, j.width, j.inv, d_coef, ry.c, Bb::from_canon(11), ro->ptr,
                                   ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // coef is host memory
    }
    dev_free(ctx, d_coef);
    if (e != hipSuccess) return hip_err(e, "deep reduce");
    return ZKG_OK;
}

void deep_release(zkg_ctx *ctx, DeepJob &j)
{
    if (j.inv) {
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx, j.inv);
        j.inv = nullptr;
    }
}

}  // namespace

extern "C" int zkg_deep_open_values(zkg_ctx *ctx, int field, const zkg_buf *lde, uint64_t height, uint32_t width,
                                    uint32_t log_blowup, uint64_t shift, const void *z, void *ys_out)
{
    DeepJob j;
    if (!ys_out) return set_err(ZKG_EINVAL, "null argument");
    int rc = deep_check(ctx, field, lde, height, width, shift, z, j);
    if (!rc) rc = deep_inv(ctx, j);
    if (!rc) rc = deep_values(ctx, lde, log_blowup, j, ys_out);
    if (ctx) deep_release(ctx, j);
    return rc;
}

extern "C" int zkg_deep_reduce(zkg_ctx *ctx, int field, const zkg_buf *lde, uint64_t height, uint32_t width,
                               uint64_t shift, const void *z, const void *alpha, uint64_t alpha_offset, const void *ys,
                               zkg_buf *ro)
{
    DeepJob j;
    if (!alpha || !ys || !ro) return set_err(ZKG_EINVAL, "null argument");
    int rc = deep_check(ctx, field, lde, height, width, shift, z, j);
    if (!rc) rc = deep_inv(ctx, j);
    if (!rc) rc = deep_reduce(ctx, lde, alpha, alpha_offset, ys, ro, j);
    if (ctx) deep_release(ctx, j);
    return rc;
}

extern "C" int zkg_deep_open(zkg_ctx *ctx, int field, const zkg_buf *lde, uint64_t height, uint32_t width,
                             uint32_t log_blowup, uint64_t shift, const void *z, const void *alpha,
                             uint64_t alpha_offset, void *ys_out, zkg_buf *ro)
{
    DeepJob j;
    if (!alpha || !ys_out || !ro) return set_err(ZKG_EINVAL, "null argument");
    int rc = deep_check(ctx, field, lde, height, width, shift, z, j);
    if (!rc) rc = deep_inv(ctx, j);
    if (!rc) rc = deep_values(ctx, lde, log_blowup, j, ys_out);
    if (!rc) rc = deep_reduce(ctx, lde, alpha, alpha_offset, ys_out, ro, j);
    if (ctx) deep_release(ctx, j);
    return rc;
}

// ------------------------------------------------------------------ commit
static uint64_t generator_of(int field) { return field == FIELD_GL ? 7 : 31; }

void zkg::stage_mark(zkg_ctx *ctx, hipStream_t st, int i)
{
    if (ctx->timing) (void)hipEventRecord(ctx->ev[i], st);
}

void zkg::stage_collect(zkg_ctx *ctx, int nmarks)
{
    if (!ctx->timing) return;
    (void)hipEventSynchronize(ctx->ev[nmarks - 1]);
    for (int i = 0; i < 4; i++) ctx->stage_ms[i] = 0;
    for (int i = 0; i + 1 < nmarks && i < 4; i++) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]);
        ctx->stage_ms[i] = ms;
    }
}

static int commit_common(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *trace, uint64_t height,
                         uint32_t width, uint32_t log_blowup, zkg_buf *lde_out, zkg_tree **out, bool plain)
{
    if (!ctx || !trace || !lde_out || !out) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_lde_shape(field, height, width, log_blowup)) return rc;
    if (!get_p2(ctx, field, p2_width))
        return set_err(ZKG_ESTATE, "Poseidon2 constants for field %d width %d not loaded", field, p2_width);
    uint64_t es = esize(field), big = height << log_blowup;
    if (span_bytes(height, width, es) > trace->bytes || span_bytes(big, width, es) > lde_out->bytes)
        return set_err(ZKG_EINVAL, "buffer too small");
    uint64_t tb = plain ? height * width * es : coset_lde_tmp_bytes(field, height, width, log_blowup);
    void *tmp = nullptr;
    int rc;
    if (tb && (rc = dev_alloc(ctx, tb, &tmp))) return rc;
    zkg_tree *t;
    if ((rc = tree_alloc(ctx, field, p2_width, big, lde_out->ptr, width, &t))) {
        dev_free(ctx, tmp);
        return rc;
    }
    hipStream_t st = ctx->stream;
    stage_mark(ctx, st, 0);
    if (plain) {
        hipError_t e = launch_lde_plain(field, ctx->tw[field], trace->ptr, lde_out->ptr, tmp, height, width, log_blowup,
                                        generator_of(field), st, true);
        if (e != hipSuccess) rc = hip_err(e, "plain lde");
    } else {
        rc = run_lde(ctx, st, field, trace->ptr, height, width, log_blowup, generator_of(field), lde_out->ptr, tmp);
    }
    stage_mark(ctx, st, 1);
    if (!rc) {
        const P2Set *p2 = get_p2(ctx, field, p2_width);
        hipError_t e = launch_hash_rows(field, p2_width, *p2, t->mat, big, width, t->layers[0], st);
        if (e == hipSuccess && plain) e = hipStreamSynchronize(st);
        stage_mark(ctx, st, 2);
        if (e == hipSuccess)
            e = launch_merkle_tree(field, p2_width, *p2, t->layers.data(), t->nlayers, big, plain ? 1 : 9, st, plain);
        stage_mark(ctx, st, 3);
        if (e != hipSuccess) rc = hip_err(e, "tree");
    }
    hipError_t e = hipStreamSynchronize(st);
    if (!rc && e != hipSuccess) rc = hip_err(e, "commit sync");
    stage_collect(ctx, 4);
    dev_free(ctx, tmp);
    if (rc) {
        zkg_tree_free(t);
        return rc;
    }
    *out = t;
    return ZKG_OK;
}

extern "C" int zkg_commit(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *trace, uint64_t height, uint32_t width,
                          uint32_t log_blowup, zkg_buf *lde_out, zkg_tree **out)
{
    return commit_common(ctx, field, p2_width, trace, height, width, log_blowup, lde_out, out, false);
}

extern "C" int zkg_commit_plain(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *trace, uint64_t height,
                                uint32_t width, uint32_t log_blowup, zkg_buf *lde_out, zkg_tree **out)
{
    return commit_common(ctx, field, p2_width, trace, height, width, log_blowup, lde_out, out, true);
}

// ------------------------------------------------------------------ FRI
extern "C" int zkg_fri_fold(zkg_ctx *ctx, int field, const zkg_buf *evals, uint64_t n_out, uint64_t alpha,
                            const zkg_buf *itw, zkg_buf *out)
{
    if (!ctx || !evals || !itw || !out) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_field(field)) return rc;
    uint64_t es = esize(field);
    if (span_bytes(n_out, 2, es) > evals->bytes || span_bytes(n_out, es) > itw->bytes || span_bytes(n_out, es) > out->bytes)
        return set_err(ZKG_EINVAL, "buffer too small");
    if (alpha >= (field == FIELD_GL ? Gl::P : (uint64_t)Bb::P)) return set_err(ZKG_EINVAL, "alpha not canonical");
    HIP_TRY(launch_fri_fold_explicit(field, evals->ptr, n_out, alpha, itw->ptr, out->ptr, ctx->stream), "fri fold");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return ZKG_OK;
}

// Canonical challenge words to the fold kernels' register form: EF4, four
// uint32_t to Montgomery form; base field, one uint64_t (Goldilocks: its two
// halves; BabyBear: Montgomery form).
static int fri_beta_words(int ext, int field, const void *beta, uint32_t br[4])
{
    if (ext) {
        const uint32_t *b = (const uint32_t *)beta;
        for (int i = 0; i < 4; i++) {
            if (b[i] >= Bb::P) return set_err(ZKG_EINVAL, "beta is not canonical");
            br[i] = Bb::from_canon(b[i]);
        }
        return ZKG_OK;
    }
    const uint64_t v = *(const uint64_t *)beta;
    if (v >= (field == FIELD_GL ? Gl::P : (uint64_t)Bb::P)) return set_err(ZKG_EINVAL, "beta is not canonical");
    br[0] = field == FIELD_GL ? (uint32_t)v : Bb::from_canon((uint32_t)v);
    br[1] = field == FIELD_GL ? (uint32_t)(v >> 32) : 0;
    br[2] = br[3] = 0;
    return ZKG_OK;
}

extern "C" int zkg_fri_fold_matrix(zkg_ctx *ctx, int ext, int field, const zkg_buf *evals, uint64_t n_out,
                                   const void *beta, zkg_buf *out)
{
    if (!ctx || !evals || !beta || !out) return set_err(ZKG_EINVAL, "null argument");
    if (ext) field = FIELD_BB;
    if (int rc = check_field(field)) return rc;
    if (!is_pow2(n_out) || ilog2u(n_out) + 1 > LOG_TW_TOP) return set_err(ZKG_EINVAL, "n_out must be a power of two <= 2^23");
    uint64_t es = ext ? 16 : esize(field);
    if (span_bytes(n_out, 2, es) > evals->bytes || span_bytes(n_out, es) > out->bytes) return set_err(ZKG_EINVAL, "buffer too small");
    uint32_t br[4];
    if (int rc = fri_beta_words(ext, field, beta, br)) return rc;
    P2Set none;
    HIP_TRY(launch_fri_fold(ext, field, ctx->tw[field], none, evals->ptr, n_out, br, out->ptr, nullptr, ctx->stream),
            "fold");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    return ZKG_OK;
}

// The commit phase of p3-fri 0.4.2 (external; the reference configures it
// with create_test_fri_params at zkf-backends/src/plonky3.rs:597): layer r of
// 2^(log_len - r) codeword elements is committed as a width-2 matrix (two
// elements per leaf, one sponge absorption: 8 BabyBear words for EF4, 2 field
// elements over a base field) and folded with beta_r (p3 fold_matrix, arity
// 2) into layer r + 1; the last fold leaves the final codeword.
//
// fri_setup checks the shape and allocates a phase's storage: the codewords of
// layers 0 .. n (layer 0 a copy of the caller's buffer, enqueued on the
// context stream, so openings stay valid after the caller reuses it) and the
// tree of every committed layer.
static int fri_setup(zkg_ctx *ctx, int ext, int field, const zkg_buf *cw, uint32_t log_len, uint32_t log_final,
                     zkg_fri **out)
{
    if (log_final >= log_len || log_len > 23 || log_final < 1)
        return set_err(ZKG_EINVAL, "need 1 <= log_final < log_len <= 23");
    const uint64_t es = ext ? 16 : esize(field);
    if ((es << log_len) > cw->bytes) return set_err(ZKG_EINVAL, "codeword buffer too small");
    const int nl = (int)(log_len - log_final);
    zkg_fri *f = new zkg_fri();
    f->ctx = ctx;
    f->field = field;
    f->ext = ext;
    f->elem_bytes = (uint32_t)es;
    f->log_len = log_len;
    f->log_final = log_final;
    int rc = ZKG_OK;
    for (int r = 0; r <= nl && !rc; r++) {
        void *ptr;
        rc = dev_alloc(ctx, es << (log_len - r), &ptr);
        if (!rc) f->codewords.push_back(ptr);
    }
    for (int r = 0; r < nl && !rc; r++) {
        zkg_tree *t;
        rc = tree_alloc(ctx, field, 16, 1ull << (log_len - 1 - r), f->codewords[r], ext ? 8 : 2, &t);
        if (!rc) f->trees.push_back(t);
    }
    if (!rc) {
        hipError_t e = hipMemcpyAsync(f->codewords[0], cw->ptr, es << log_len, hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) rc = hip_err(e, "copy codeword");
    }
    if (rc) {
        zkg_fri_free(f);
        return rc;
    }
    *out = f;
    return ZKG_OK;
}

// The whole phase in one call, every beta given (br[r]: fold r's challenge in
// the fold kernels' register form).  The fold chain runs on the context
// stream: layer 0's leaf hash, then each fold fused with the leaf hash of the
// layer it produces (event r: layer r's leaves ready).  Layer r's tree needs
// only those leaves, so trees of more than fri_tops_max_leaves() leaves run on
// the side streams beside the later folds.  The whole chain is enqueued before
// any tree: with the trees' launches between its folds, the host's enqueue
// rate paced the short folds of the small layers.  The trees of the small
// layers are latency-bound launch chains; they are built together on the
// context stream once the chain has produced all their leaves, one launch per
// few levels of all of them.
static int fri_one_call(zkg_ctx *ctx, int ext, int field, const zkg_buf *cw, uint32_t log_len, uint32_t log_final,
                        const std::vector<std::array<uint32_t, 4>> &br, zkg_fri **out)
{
    const P2Set *p2 = get_p2(ctx, field, 16);
    if (!p2) return set_err(ZKG_ESTATE, "Poseidon2-16 constants for field %d not loaded", field);
    zkg_fri *f;
    if (int rc = fri_setup(ctx, ext, field, cw, log_len, log_final, &f)) return rc;
    const int nl = (int)f->trees.size();
    hipStream_t st = ctx->stream;
    stage_mark(ctx, st, 0);
    zkg_tree *t0 = f->trees[0];
    hipError_t e = launch_hash_rows(field, 16, *p2, f->codewords[0], t0->n_leaves, ext ? 8 : 2, t0->layers[0], st);
    if (e == hipSuccess) e = hipEventRecord(ctx->fri_ev[0], st);
    for (int r = 0; r < nl && e == hipSuccess; r++) {
        void *next_digests = (r + 1 < nl) ? f->trees[r + 1]->layers[0] : nullptr;
        e = launch_fri_fold(ext, field, ctx->tw[field], *p2, f->codewords[r], 1ull << (log_len - 1 - r), br[r].data(),
                            f->codewords[r + 1], next_digests, st);
        if (e == hipSuccess) e = hipEventRecord(ctx->fri_ev[r + 1], st);
    }
    std::vector<MerkleTop> tops;
    for (int r = 0; r < nl && e == hipSuccess; r++) {
        zkg_tree *t = f->trees[r];
        if (t->n_leaves <= fri_tops_max_leaves()) {
            tops.push_back(MerkleTop{t->layers.data(), t->nlayers, t->n_leaves});
            continue;
        }
        hipStream_t side = ctx->side[r & 1];
        e = hipStreamWaitEvent(side, ctx->fri_ev[r], 0);
        if (e == hipSuccess)
            e = launch_merkle_tree(field, 16, *p2, t->layers.data(), t->nlayers, t->n_leaves, 9, side, false);
    }
    if (e == hipSuccess && !tops.empty()) e = launch_merkle_tops(field, 16, *p2, tops.data(), (int)tops.size(), st);
    for (int i = 0; i < 2; i++) {  // join the side streams
        hipError_t e2 = hipEventRecord(ctx->side_join[i], ctx->side[i]);
        if (e2 == hipSuccess) e2 = hipStreamWaitEvent(st, ctx->side_join[i], 0);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    stage_mark(ctx, st, 1);
    if (ctx->timing) {
        (void)hipEventSynchronize(ctx->ev[1]);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
        for (int i = 0; i < 4; i++) ctx->stage_ms[i] = 0;
        ctx->stage_ms[3] = ms;
    }
    if (e != hipSuccess) {
        int rc = hip_err(e, "fri commit phase");
        zkg_fri_free(f);
        return rc;
    }
    f->folds = nl;
    *out = f;
    return ZKG_OK;
}

// EF4 over BabyBear: betas[4 r .. 4 r + 3] (reduced mod p).
extern "C" int zkg_fri_commit_phase(zkg_ctx *ctx, const zkg_buf *cw, uint32_t log_len, uint32_t log_final,
                                    const uint32_t *betas, zkg_fri **out)
{
    if (!ctx || !cw || !betas || !out) return set_err(ZKG_EINVAL, "null argument");
    if (log_final >= log_len || log_len > 23 || log_final < 1)
        return set_err(ZKG_EINVAL, "need 1 <= log_final < log_len <= 23");
    std::vector<std::array<uint32_t, 4>> br(log_len - log_final);
    for (size_t r = 0; r < br.size(); r++)
        for (int i = 0; i < 4; i++) {
            const uint32_t v = betas[4 * r + i];
            br[r][i] = Bb::from_canon(v >= Bb::P ? v - Bb::P : v);
        }
    return fri_one_call(ctx, 1, FIELD_BB, cw, log_len, log_final, br, out);
}

// A base field, Challenge = Val: the reference's own configuration
// (`type Challenge<F> = F`, zkf-backends/src/plonky3.rs:55), whose
// ExtensionMmcs is then the field's Poseidon2-16 MMCS; betas[r] canonical.
extern "C" int zkg_fri_commit_phase_base(zkg_ctx *ctx, int field, const zkg_buf *cw, uint32_t log_len,
                                         uint32_t log_final, const uint64_t *betas, zkg_fri **out)
{
    if (!ctx || !cw || !betas || !out) return set_err(ZKG_EINVAL, "null argument");
    if (int rc = check_field(field)) return rc;
    if (log_final >= log_len || log_len > 23 || log_final < 1)
        return set_err(ZKG_EINVAL, "need 1 <= log_final < log_len <= 23");
    std::vector<std::array<uint32_t, 4>> br(log_len - log_final);
    for (size_t r = 0; r < br.size(); r++)
        if (int rc = fri_beta_words(0, field, &betas[r], br[r].data()))
            return set_err(rc, "beta %zu is not a canonical field element", r);
    return fri_one_call(ctx, 0, field, cw, log_len, log_final, br, out);
}

// The commit phase one layer at a time, in p3-fri's transcript order (the
// reference's prover runs it through TwoAdicFriPcs, plonky3.rs:68,597):
// zkg_fri_begin commits layer 0; each zkg_fri_fold_next(beta_r) folds layer r
// into layer r + 1 and commits that layer, the fold fused with its leaf hash,
// until the last call leaves the final codeword.  Every step runs on the
// context stream; the caller reads layer r's root (zkg_fri_layer_root) before
// sampling beta_r.  All layer storage is allocated by zkg_fri_begin.
extern "C" int zkg_fri_begin(zkg_ctx *ctx, int ext, int field, const zkg_buf *cw, uint32_t log_len,
                             uint32_t log_final, zkg_fri **out)
{
    if (!ctx || !cw || !out) return set_err(ZKG_EINVAL, "null argument");
    if (ext && field != FIELD_BB) return set_err(ZKG_EINVAL, "EF4 challenges need the BabyBear field");
    if (int rc = check_field(field)) return rc;
    const P2Set *p2 = get_p2(ctx, field, 16);
    if (!p2) return set_err(ZKG_ESTATE, "Poseidon2-16 constants for field %d not loaded", field);
    zkg_fri *f;
    if (int rc = fri_setup(ctx, ext, field, cw, log_len, log_final, &f)) return rc;
    f->stepwise = true;
    f->p2_epoch = ctx->p2_epoch;
    hipStream_t st = ctx->stream;
    zkg_tree *t = f->trees[0];
    hipError_t e = launch_hash_rows(field, 16, *p2, f->codewords[0], t->n_leaves, ext ? 8 : 2, t->layers[0], st);
    if (e == hipSuccess) e = launch_merkle_tree(field, 16, *p2, t->layers.data(), t->nlayers, t->n_leaves, 9, st, false);
    if (e != hipSuccess) {
        int rc = hip_err(e, "fri begin");
        zkg_fri_free(f);
        return rc;
    }
    *out = f;
    return ZKG_OK;
}

extern "C" int zkg_fri_fold_next(zkg_fri *f, const void *beta)
{
    if (!f || !beta) return set_err(ZKG_EINVAL, "null argument");
    if (!f->stepwise) return set_err(ZKG_ESTATE, "this commit phase was built in one call");
    const int nl = (int)f->trees.size();
    if (f->folds >= nl) return set_err(ZKG_ESTATE, "the commit phase is complete");
    zkg_ctx *ctx = f->ctx;
    if (ctx->p2_epoch != f->p2_epoch)
        return set_err(ZKG_ESTATE, "Poseidon2 constants were reloaded during the commit phase");
    const P2Set *p2 = get_p2(ctx, f->field, 16);
    if (!p2) return set_err(ZKG_ESTATE, "Poseidon2-16 constants not loaded");
    uint32_t br[4];
    if (int rc = fri_beta_words(f->ext, f->field, beta, br)) return rc;
    const int r = f->folds;
    hipStream_t st = ctx->stream;
    zkg_tree *next = r + 1 < nl ? f->trees[r + 1] : nullptr;
    hipError_t e = launch_fri_fold(f->ext, f->field, ctx->tw[f->field], *p2, f->codewords[r],
                                   1ull << (f->log_len - 1 - r), br, f->codewords[r + 1],
                                   next ? next->layers[0] : nullptr, st);
    if (e == hipSuccess && next)
        e = launch_merkle_tree(f->field, 16, *p2, next->layers.data(), next->nlayers, next->n_leaves, 9, st, false);
    if (e != hipSuccess) return hip_err(e, "fri fold");
    f->folds++;
    return ZKG_OK;
}

extern "C" int zkg_fri_num_layers(const zkg_fri *f, uint32_t *n)
{
    if (!f || !n) return set_err(ZKG_EINVAL, "null argument");
    *n = (uint32_t)f->trees.size();
    return ZKG_OK;
}

extern "C" int zkg_fri_layer_root(const zkg_fri *f, uint32_t layer, uint32_t *root8)
{
    if (!f || !root8) return set_err(ZKG_EINVAL, "null argument");
    if (layer >= f->trees.size()) return set_err(ZKG_EINVAL, "layer out of range");
    if ((int)layer > f->folds) return set_err(ZKG_ESTATE, "layer %u is not committed yet", layer);
    return zkg_tree_root(f->trees[layer], root8);
}

extern "C" int zkg_fri_final(const zkg_fri *f, uint32_t *out)
{
    if (!f || !out) return set_err(ZKG_EINVAL, "null argument");
    if (f->folds < (int)f->trees.size()) return set_err(ZKG_ESTATE, "the commit phase is not complete");
    HIP_TRY(hipStreamSynchronize(f->ctx->stream), "sync");
    HIP_TRY(hipMemcpy(out, f->codewords.back(), (uint64_t)f->elem_bytes << f->log_final, hipMemcpyDeviceToHost),
            "final");
    return ZKG_OK;
}

extern "C" int zkg_fri_open(const zkg_fri *f, const uint64_t *idx, uint32_t n, uint32_t *rows_out, uint32_t *paths_out)
{
    if (!f || (!idx && n)) return set_err(ZKG_EINVAL, "null argument");
    const int nl = (int)f->trees.size();
    if (f->folds < nl) return set_err(ZKG_ESTATE, "the commit phase is not complete");
    for (uint32_t q = 0; q < n; q++)
        if (idx[q] >= (1ull << f->log_len)) return set_err(ZKG_EINVAL, "query index out of range");
    // per (query, layer): one width-2 row (2 elements) and the sibling digests
    // (8 field elements each); every piece of every query in one gather
    const uint64_t row_bytes = 2ull * f->elem_bytes;
    const uint64_t digest_bytes = 8ull * (f->field == FIELD_GL ? 8 : 4);
    uint64_t per_query_path = 0;
    for (int r = 0; r < nl; r++) per_query_path += (uint64_t)(f->trees[r]->nlayers - 1) * digest_bytes;
    const uint64_t rows_total = rows_out ? n * nl * row_bytes : 0, paths_total = paths_out ? n * per_query_path : 0;
    std::vector<OpenPiece> pieces;
    for (uint32_t q = 0; q < n; q++) {
        uint64_t path_off = rows_total + q * per_query_path;
        for (int r = 0; r < nl; r++) {
            const zkg_tree *t = f->trees[r];
            const uint64_t leaf = idx[q] >> (r + 1);
            if (rows_out)
                pieces.push_back(OpenPiece{(const uint8_t *)f->codewords[r] + leaf * row_bytes,
                                           ((uint64_t)q * nl + r) * row_bytes, (uint32_t)row_bytes});
            for (int i = 0; i + 1 < t->nlayers && paths_out; i++) {
                const uint64_t sib = (leaf >> i) ^ 1;
                pieces.push_back(OpenPiece{(const uint8_t *)t->layers[i] + sib * digest_bytes, path_off,
                                           (uint32_t)digest_bytes});
                path_off += digest_bytes;
            }
        }
    }
    return gather_openings(f->ctx, pieces, rows_total, rows_out, paths_total, paths_out);
}

extern "C" int zkg_fri_free(zkg_fri *f)
{
    if (!f) return ZKG_OK;
    (void)hipStreamSynchronize(f->ctx->stream);
    for (zkg_tree *t : f->trees) {
        t->mat = nullptr;
        zkg_tree_free(t);
    }
    for (void *p : f->codewords) dev_free(f->ctx, p);
    delete f;
    return ZKG_OK;
}
