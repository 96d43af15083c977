// zkg_driver.hip -- the runtime's GPU dispatch driver over libzkg.
//
// Host-side mirror of zkf-runtime's GpuDispatchDriver (metal_driver.rs:55-93)
// and GpuBufferAllocator (buffer_bridge.rs:77-84) for the commit-path
// ProverOps (graph.rs:69-108), with the node payloads of execution.rs:51-106.
// The Metal driver (metal_dispatch_macos.rs:305-779) reads each input slot
// into host memory, runs one accelerator call and writes the node's primary
// output slot; here the BufferBridge slots are device buffers, so NTT, LDE
// and FRI fold run on HBM-resident data.  The Poseidon batch and Merkle layer
// arms keep the Metal driver's state packing (16-lane states, zero padded),
// which is defined on the host.  A node this lane does not take returns
// ZKG_EFALLBACK, the FallbackSignal that lets the runtime run it on the CPU.
#include "../../include/zkg.h"
#include "zkg_field.h"
#include "zkg_internal.h"
#include "zkg_runtime.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

using namespace zkg;

struct zkg_bridge {
    zkg_ctx *ctx = nullptr;
    std::map<uint32_t, zkg_buf *> slots;
};

namespace {

zkg_buf *slot_buf(const zkg_bridge *b, uint32_t slot)
{
    auto it = b->slots.find(slot);
    return it == b->slots.end() ? nullptr : it->second;
}

// GpuBufferAllocator::alloc_shared for a slot (an existing slot is replaced).
int slot_alloc(zkg_bridge *b, uint32_t slot, uint64_t bytes, zkg_buf **out)
{
    if (zkg_buf *old = slot_buf(b, slot)) {
        if (old->bytes == bytes) {
            *out = old;
            return ZKG_OK;
        }
        zkg_buf_free(old);
        b->slots.erase(slot);
    }
    zkg_buf *nb = nullptr;
    if (int rc = zkg_buf_alloc(b->ctx, std::max<uint64_t>(bytes, 1), &nb)) return rc;
    nb->bytes = bytes;  // the slot's logical size
    b->slots[slot] = nb;
    *out = nb;
    return ZKG_OK;
}

// The output buffer of a node whose output slot is also one of its inputs:
// a fresh buffer, put into the slot (freeing the input) only after the node's
// kernels are done.  Otherwise the slot's own buffer.
int out_buffer(zkg_bridge *b, uint32_t slot, uint64_t bytes, bool aliases_input, zkg_buf **out, zkg_buf **pending)
{
    *pending = nullptr;
    if (!aliases_input) return slot_alloc(b, slot, bytes, out);
    if (int rc = zkg_buf_alloc(b->ctx, std::max<uint64_t>(bytes, 1), out)) return rc;
    (*out)->bytes = bytes;
    *pending = *out;
    return ZKG_OK;
}

void install(zkg_bridge *b, uint32_t slot, zkg_buf *nb)
{
    if (zkg_buf *old = slot_buf(b, slot)) zkg_buf_free(old);
    b->slots[slot] = nb;
}

const char *field_name(int f) { return f == FIELD_GL ? "goldilocks" : "babybear"; }

template <class T>
int read_slot(const zkg_bridge *b, uint32_t slot, std::vector<T> &v)
{
    const zkg_buf *buf = slot_buf(b, slot);
    if (!buf) return set_err(ZKG_EINVAL, "slot %u is not resident", slot);
    v.assign(buf->bytes / sizeof(T), 0);
    return v.empty() ? ZKG_OK : zkg_buf_download(buf, 0, v.data(), v.size() * sizeof(T));
}

template <class T>
int write_slot(zkg_bridge *b, uint32_t slot, const std::vector<T> &v)
{
    zkg_buf *out = nullptr;
    if (int rc = slot_alloc(b, slot, v.size() * sizeof(T), &out)) return rc;
    return v.empty() ? ZKG_OK : zkg_buf_upload(out, 0, v.data(), v.size() * sizeof(T));
}

// PoseidonBatch (metal_dispatch_macos.rs:492-566): `count` states of `width`
// elements (chunks of the input, zero padded to 16 lanes), permuted with the
// constants of the aux slot or, without one, those of seed 42; the first
// `width` lanes of each state are the output.
template <class T>
int run_poseidon_batch(zkg_bridge *b, const zkg_node *nd, int field, uint64_t *in_bytes, uint64_t *out_bytes)
{
    if (nd->width > 16) return set_err(ZKG_EFALLBACK, "PoseidonBatch width %u exceeds 16", nd->width);
    std::vector<T> in;
    if (int rc = read_slot(b, nd->in_slot, in)) return rc;
    const uint64_t count = nd->size, w = std::max<uint32_t>(nd->width, 1);
    // the states live in host and device memory: a count whose states could
    // not fit the context's HBM budget is refused before anything is sized
    if (span_bytes(count, 16, sizeof(T)) > b->ctx->hbm_budget)
        return set_err(ZKG_EINVAL, "PoseidonBatch count %llu exceeds the HBM budget", (unsigned long long)count);
    std::vector<T> states(count * 16, 0);
    for (uint64_t p = 0; p < count; p++)
        for (uint64_t i = 0; i < nd->width && p * w + i < in.size(); i++) states[p * 16 + i] = in[p * w + i];
    std::vector<uint64_t> rc64;
    uint32_t n_ext = nd->n_ext, n_int = nd->n_int;
    if (nd->aux_slot && slot_buf(b, nd->aux_slot)) {
        std::vector<T> rc;
        if (int e = read_slot(b, nd->aux_slot, rc)) return e;
        // the flat table [initial][internal][terminal] of width-16 rounds
        if (rc.size() != (uint64_t)n_ext * 16 + n_int)
            return set_err(ZKG_EINVAL, "PoseidonBatch constants slot holds %zu elements, rounds (%u, %u) need %llu",
                           rc.size(), n_ext, n_int, (unsigned long long)((uint64_t)n_ext * 16 + n_int));
        rc64.assign(rc.begin(), rc.end());
    } else {
        if (int e = p2_seed_constants(field, 16, 42, rc64)) return e;
        n_ext = 8;
        n_int = (uint32_t)p2_internal_rounds(field);
    }
    if (rc64.empty()) return set_err(ZKG_EFALLBACK, "Poseidon round constants unavailable");
    int rc;
    if constexpr (sizeof(T) == 8) {
        rc = zkg_poseidon2_batch_permute_goldilocks(b->ctx, states.data(), count, rc64.data(), n_ext, n_int);
    } else {
        std::vector<uint32_t> rc32(rc64.begin(), rc64.end());
        rc = zkg_poseidon2_batch_permute_babybear(b->ctx, states.data(), count, rc32.data(), n_ext, n_int);
    }
    if (rc) return rc;
    std::vector<T> out(count * nd->width);
    for (uint64_t p = 0; p < count; p++)
        for (uint32_t i = 0; i < nd->width; i++) out[p * nd->width + i] = states[p * 16 + i];
    *in_bytes = in.size() * sizeof(T);
    *out_bytes = out.size() * sizeof(T);
    return write_slot(b, nd->out_slot, out);
}

// MerkleLayer (metal_dispatch_macos.rs:684-757): parent p compresses leaves 2p
// and min(2p + 1, n - 1); up to 8 elements of each leaf fill lanes 0..7 and
// 8..15 of a zero state, which is permuted with the constants of the runtime
// seed; the parent keeps the first out_bytes / (parents x element) lanes,
// clamped to 1..8.
template <class T>
int run_merkle_layer(zkg_bridge *b, const zkg_node *nd, int field, uint64_t *in_bytes, uint64_t *out_bytes)
{
    std::vector<T> in;
    if (int rc = read_slot(b, nd->in_slot, in)) return rc;
    const uint64_t leaves = nd->size;
    if (in.empty() || !leaves) return set_err(ZKG_EFALLBACK, "MerkleLayer missing %s leaves", field_name(field));
    const uint64_t lw = std::max<uint64_t>(in.size() / leaves, 1), parents = leaves / 2 + (leaves & 1);
    if (span_bytes(parents, 16, sizeof(T)) > b->ctx->hbm_budget)
        return set_err(ZKG_EINVAL, "MerkleLayer leaf_count %llu exceeds the HBM budget", (unsigned long long)leaves);
    const uint64_t ow = nd->out_bytes ? std::min<uint64_t>(std::max<uint64_t>(nd->out_bytes / sizeof(T) / parents, 1), 8)
                                      : 1;
    std::vector<uint64_t> rc64;
    if (int e = p2_seed_constants(field, 16, nd->seed, rc64)) return e;
    std::vector<T> states(parents * 16, 0);
    for (uint64_t p = 0; p < parents; p++) {
        const uint64_t l0 = 2 * p * lw, r0 = std::min<uint64_t>(2 * p + 1, leaves - 1) * lw;
        for (uint64_t i = 0; i < std::min<uint64_t>(lw, 8); i++) {
            if (l0 + i < in.size()) states[p * 16 + i] = in[l0 + i];
            if (r0 + i < in.size()) states[p * 16 + 8 + i] = in[r0 + i];
        }
    }
    const uint32_t n_int = (uint32_t)p2_internal_rounds(field);
    int rc;
    if constexpr (sizeof(T) == 8) {
        rc = zkg_poseidon2_batch_permute_goldilocks(b->ctx, states.data(), parents, rc64.data(), 8, n_int);
    } else {
        std::vector<uint32_t> rc32(rc64.begin(), rc64.end());
        rc = zkg_poseidon2_batch_permute_babybear(b->ctx, states.data(), parents, rc32.data(), 8, n_int);
    }
    if (rc) return rc;
    std::vector<T> out(parents * ow);
    for (uint64_t p = 0; p < parents; p++)
        for (uint64_t i = 0; i < ow; i++) out[p * ow + i] = states[p * 16 + i];
    *in_bytes = in.size() * sizeof(T);
    *out_bytes = out.size() * sizeof(T);
    return write_slot(b, nd->out_slot, out);
}

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

}  // namespace

// ------------------------------------------------------------------ BufferBridge
extern "C" int zkg_bridge_create(zkg_ctx *ctx, zkg_bridge **out)
{
    if (!ctx || !out) return set_err(ZKG_EINVAL, "null argument");
    *out = new zkg_bridge;
    (*out)->ctx = ctx;
    return ZKG_OK;
}

extern "C" void zkg_bridge_destroy(zkg_bridge *b)
{
    if (!b) return;
    for (auto &kv : b->slots) zkg_buf_free(kv.second);
    delete b;
}

extern "C" int zkg_bridge_alloc(zkg_bridge *b, uint32_t slot, uint64_t bytes)
{
    if (!b || !slot) return set_err(ZKG_EINVAL, !b ? "null bridge" : "slot 0 means no buffer");
    zkg_buf *out = nullptr;
    return slot_alloc(b, slot, bytes, &out);
}

extern "C" int zkg_bridge_free(zkg_bridge *b, uint32_t slot)
{
    if (!b) return set_err(ZKG_EINVAL, "null bridge");
    if (zkg_buf *buf = slot_buf(b, slot)) {
        zkg_buf_free(buf);
        b->slots.erase(slot);
    }
    return ZKG_OK;
}

extern "C" int zkg_bridge_write(zkg_bridge *b, uint32_t slot, const void *src, uint64_t bytes)
{
    if (!b || !slot || (!src && bytes)) return set_err(ZKG_EINVAL, "null argument or slot 0");
    zkg_buf *out = nullptr;
    if (int rc = slot_alloc(b, slot, bytes, &out)) return rc;
    return bytes ? zkg_buf_upload(out, 0, src, bytes) : ZKG_OK;
}

extern "C" int zkg_bridge_read(const zkg_bridge *b, uint32_t slot, void *dst, uint64_t cap, uint64_t *bytes)
{
    if (!b || !bytes || (!dst && cap)) return set_err(ZKG_EINVAL, "null argument");
    const zkg_buf *buf = slot_buf(b, slot);
    if (!buf) return set_err(ZKG_EINVAL, "slot %u is not resident", slot);
    *bytes = buf->bytes;
    if (!cap) return ZKG_OK;
    if (cap < buf->bytes) return set_err(ZKG_EINVAL, "slot %u holds %llu bytes", slot, (unsigned long long)buf->bytes);
    return buf->bytes ? zkg_buf_download(buf, 0, dst, buf->bytes) : ZKG_OK;
}

extern "C" zkg_buf *zkg_bridge_buffer(zkg_bridge *b, uint32_t slot) { return b ? slot_buf(b, slot) : nullptr; }

// ------------------------------------------------------------------ GpuDispatchDriver
extern "C" int zkg_driver_should_try(const zkg_node *nd)
{
    // size thresholds below which a launch does not pay (metal_dispatch_macos.rs:151-188)
    if (!nd) return 0;
    switch (nd->op) {
    case ZKG_OP_NTT: return nd->size >= (1u << 12);
    case ZKG_OP_LDE: return nd->size >= (1u << 10);
    case ZKG_OP_POSEIDON_BATCH: return nd->size >= 1024 && nd->width <= 16;
    case ZKG_OP_MERKLE_LAYER: return nd->size >= 1024;
    case ZKG_OP_FRI_FOLD: return nd->size >= (1u << 12);
    default: return 0;
    }
}

extern "C" int zkg_driver_execute(zkg_ctx *ctx, const zkg_node *nd, zkg_bridge *b, zkg_telemetry *tel)
{
    if (!ctx || !nd || !b) return set_err(ZKG_EINVAL, "null argument");
    if (b->ctx != ctx) return set_err(ZKG_EINVAL, "bridge of another context");
    if (nd->field != FIELD_GL && nd->field != FIELD_BB)
        return set_err(ZKG_EFALLBACK, "no payload for field %d on the HIP lane", nd->field);
    if (!nd->out_slot) return set_err(ZKG_EINVAL, "node has no output slot");
    const int f = nd->field;
    const uint64_t es = f == FIELD_GL ? 8 : 4;
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t in_bytes = 0, out_bytes = 0;
    const char *kernel = "";
    int rc = ZKG_OK;
    zkg_buf *pending = nullptr;  // output to install once the kernels are done
    switch (nd->op) {
    case ZKG_OP_NTT: {  // values transformed in natural order (NttAccelerator::forward_ntt / inverse_ntt)
        const zkg_buf *in = slot_buf(b, nd->in_slot);
        if (!in) return set_err(ZKG_EINVAL, "slot %u is not resident", nd->in_slot);
        const uint64_t w = std::max<uint32_t>(nd->width, 1), n = in->bytes / es;
        if (n % w || !is_pow2(n / w)) return set_err(ZKG_EINVAL, "NTT size must be a power of 2, got %llu", (unsigned long long)(n / w));
        zkg_buf *out = nullptr;
        if (nd->out_slot != nd->in_slot) {
            rc = slot_alloc(b, nd->out_slot, in->bytes, &out);
            if (!rc && in->bytes) {
                hipError_t e = hipMemcpyAsync(out->ptr, in->ptr, in->bytes, hipMemcpyDeviceToDevice, ctx->stream);
                if (e != hipSuccess) rc = hip_err(e, "NTT input copy");
            }
        } else {
            out = slot_buf(b, nd->out_slot);
        }
        if (!rc) rc = zkg_ntt_batch(ctx, f, out, n / w, (uint32_t)w, nd->inverse ? 1 : 0);
        in_bytes = out_bytes = n * es;
        kernel = f == FIELD_GL ? "hip-ntt-goldilocks" : "hip-ntt-babybear";
        break;
    }
    case ZKG_OP_LDE: {  // coset LDE at the blowup, rows bit-reversed (p3 coset_lde_batch + bit_reverse_rows)
        const zkg_buf *in = slot_buf(b, nd->in_slot);
        if (!in) return set_err(ZKG_EINVAL, "slot %u is not resident", nd->in_slot);
        const uint64_t w = std::max<uint32_t>(nd->width, 1), n = in->bytes / es / w;
        if (!is_pow2(nd->blowup) || nd->blowup > 256)
            return set_err(ZKG_EINVAL, "blowup %u is not a power of two <= 256", nd->blowup);
        const uint32_t lb = (uint32_t)__builtin_ctz(nd->blowup);
        zkg_buf *out = nullptr;
        rc = out_buffer(b, nd->out_slot, (n << lb) * w * es, nd->out_slot == nd->in_slot, &out, &pending);
        if (!rc) rc = zkg_coset_lde_batch(ctx, f, in, n, (uint32_t)w, lb, f == FIELD_GL ? 7 : 31, out);
        in_bytes = n * w * es;
        out_bytes = (n << lb) * w * es;
        kernel = f == FIELD_GL ? "hip-lde-goldilocks" : "hip-lde-babybear";
        break;
    }
    case ZKG_OP_POSEIDON_BATCH:
        rc = f == FIELD_GL ? run_poseidon_batch<uint64_t>(b, nd, f, &in_bytes, &out_bytes)
                           : run_poseidon_batch<uint32_t>(b, nd, f, &in_bytes, &out_bytes);
        kernel = f == FIELD_GL ? "hip-poseidon2-goldilocks" : "hip-poseidon2-babybear";
        break;
    case ZKG_OP_MERKLE_LAYER:
        rc = f == FIELD_GL ? run_merkle_layer<uint64_t>(b, nd, f, &in_bytes, &out_bytes)
                           : run_merkle_layer<uint32_t>(b, nd, f, &in_bytes, &out_bytes);
        kernel = f == FIELD_GL ? "hip-merkle-goldilocks-poseidon2" : "hip-merkle-babybear-poseidon2";
        break;
    case ZKG_OP_FRI_FOLD: {  // FriAccelerator::fold_*: alpha is the first element of its slot
        const zkg_buf *ev = slot_buf(b, nd->in_slot), *tw = slot_buf(b, nd->aux2_slot);
        if (!ev || !tw) return set_err(ZKG_EINVAL, "FriFold slots %u / %u are not resident", nd->in_slot, nd->aux2_slot);
        const uint64_t n_out = ev->bytes / es / 2;
        if (!n_out || tw->bytes / es < n_out)
            return set_err(ZKG_EFALLBACK, "FriFold %s: %llu evaluations, %llu inverse twiddles", field_name(f),
                           (unsigned long long)(ev->bytes / es), (unsigned long long)(tw->bytes / es));
        // every input slot must be resident; an empty alpha slot reads as 0
        const zkg_buf *a = slot_buf(b, nd->aux_slot);
        if (!a) return set_err(ZKG_EINVAL, "FriFold alpha slot %u is not resident", nd->aux_slot);
        uint64_t alpha = 0;
        if (a->bytes >= es) rc = zkg_buf_download(a, 0, &alpha, es);
        const bool alias = nd->out_slot == nd->in_slot || nd->out_slot == nd->aux_slot || nd->out_slot == nd->aux2_slot;
        zkg_buf *out = nullptr;
        if (!rc) rc = out_buffer(b, nd->out_slot, n_out * es, alias, &out, &pending);
        if (!rc) rc = zkg_fri_fold(ctx, f, ev, n_out, alpha, tw, out);
        in_bytes = ev->bytes;
        out_bytes = n_out * es;
        kernel = f == FIELD_GL ? "hip-fri-fold-goldilocks" : "hip-fri-fold-babybear";
        break;
    }
    default:
        return set_err(ZKG_EFALLBACK, "op %u is not on the HIP lane", nd->op);
    }
    if (!rc) rc = zkg_ctx_synchronize(ctx);
    if (pending) {
        if (rc)
            zkg_buf_free(pending);
        else
            install(b, nd->out_slot, pending);
    }
    if (rc) return rc;
    if (tel) {
        snprintf(tel->kernel, sizeof tel->kernel, "%s", kernel);
        tel->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        tel->input_bytes = in_bytes;
        tel->output_bytes = out_bytes;
    }
    return ZKG_OK;
}
