// zkg_runtime.h -- handle layouts and internal helpers of the libzkg runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/zkg.h"
#include "zkg_internal.h"

namespace zkg {

struct CosetKey {
    int field, log_n, log_b;
    uint64_t shift;
    bool operator<(const CosetKey &o) const
    {
        return std::tie(field, log_n, log_b, shift) < std::tie(o.field, o.log_n, o.log_b, o.shift);
    }
};

struct BatchState;  // zkg_batch.hip
struct ServeState;  // zkg_batch.hip
struct CeCache;     // zkg_constraints.hip

}  // namespace zkg

struct zkg_ctx {
    int device = 0;
    uint64_t hbm_budget = 0;
    uint64_t hbm_used = 0;
    hipStream_t stream = nullptr;
    mutable std::mutex mem_mu, table_mu;
    std::unordered_map<void *, uint64_t> alloc_sizes;  // live blocks
    std::multimap<uint64_t, void *> cache;              // freed blocks by size
    uint64_t cached_bytes = 0;
    zkg::TwiddleSet tw[2];
    zkg::P2Set p2[3];  // GL-16, GL-8, BB-16
    uint64_t p2_epoch = 0;  // bumped by every constant reload
    std::map<zkg::CosetKey, zkg::CosetSet> cosets;
    bool timing = false;
    hipEvent_t ev[8] = {};
    hipStream_t side[2] = {};      // FRI layer trees overlap the fold chain
    hipEvent_t side_join[2] = {};
    hipEvent_t fri_ev[32] = {};    // leaf digests of FRI layer r ready
    double stage_ms[4] = {0, 0, 0, 0};
    zkg::BatchState *batch = nullptr;
    zkg::ServeState *serve = nullptr;
    zkg::CeCache *ce_cache = nullptr;  // constraint programs seen by zkg_constraint_eval_gl
    bool ce_cache_set = false;          // the default cache is made on first use unless configured
};

struct zkg_buf {
    zkg_ctx *ctx;
    void *ptr;
    uint64_t bytes;
};

struct zkg_tree {
    zkg_ctx *ctx = nullptr;
    int field = 0, p2_width = 16, digest = 8;
    uint64_t n_leaves = 0;
    int nlayers = 0;
    const void *mat = nullptr;  // leaf matrix (not owned)
    uint32_t mat_width = 0;
    void *base = nullptr;  // one allocation for every digest layer
    std::vector<uint64_t> layer_off;
    std::vector<uint64_t> layer_len;  // digests per layer
    std::vector<void *> layers;
    // Mixed-height MMCS trees: the committed matrices in commit order (not
    // owned) and the device descriptor array the layers were built from.
    bool mmcs = false;
    std::vector<zkg::MatDesc> mats;
    std::vector<uint64_t> mat_heights;
    uint64_t max_height = 0;
    void *desc = nullptr;
};

struct zkg_fri {
    zkg_ctx *ctx = nullptr;
    int field = 1;                  // FIELD_BB: EF4 or base-field BabyBear; FIELD_GL: base field
    uint32_t elem_bytes = 16;       // one codeword element: 16 (EF4), 8 (Goldilocks), 4 (BabyBear)
    uint32_t log_len = 0, log_final = 0;
    std::vector<void *> codewords;  // layer r codeword (2^(log_len - r) elements), last = final
    std::vector<zkg_tree *> trees;  // layer r tree over width-2 rows of codeword r
    int ext = 1;                    // EF4 challenges (1) or base-field challenges (0)
    int folds = 0;                  // folds done: layers 0..folds are committed (all once folds = n layers)
    bool stepwise = false;          // built by zkg_fri_begin / zkg_fri_fold_next
    uint64_t p2_epoch = 0;          // the context's constant epoch at zkg_fri_begin
};

namespace zkg {

int set_err(int code, const char *fmt, ...);
int hip_err(hipError_t e, const char *what);
int dev_alloc(zkg_ctx *ctx, uint64_t bytes, void **out);
void dev_free(zkg_ctx *ctx, void *p);
void cache_release(zkg_ctx *ctx);
// Bytes of live device blocks (read under the allocator's lock).
uint64_t hbm_in_use(const zkg_ctx *ctx);
int get_coset(zkg_ctx *ctx, int field, int log_n, int log_b, uint64_t shift, const CosetSet **out);
const P2Set *get_p2(zkg_ctx *ctx, int field, int width);
int run_lde(zkg_ctx *ctx, hipStream_t st, int field, const void *in, uint64_t height, uint32_t width,
            uint32_t log_blowup, uint64_t shift, void *out, void *tmp);
int tree_alloc(zkg_ctx *ctx, int field, int p2_width, uint64_t n_leaves, const void *mat, uint32_t mat_width,
               zkg_tree **out);
int tree_build(zkg_ctx *ctx, hipStream_t st, zkg_tree *t, bool plain);
int tree_open(const zkg_tree *t, const uint64_t *idx, uint32_t n, void *rows_out, void *paths_out);
void stage_mark(zkg_ctx *ctx, hipStream_t st, int i);
void stage_collect(zkg_ctx *ctx, int nmarks);
void batch_release(zkg_ctx *ctx);
void serve_shutdown(zkg_ctx *ctx);
void ce_cache_release(zkg_ctx *ctx);
// Looks up this device's address of the no-allocation scratch array
// (zkg_noalloc.hip); called once per context with its device current.
int noalloc_setup(int device);
// Poseidon2 constants drawn from SmallRng(seed), flat [initial][internal][terminal].
int p2_seed_constants(int field, int width, uint64_t seed, std::vector<uint64_t> &flat);
int p2_internal_rounds(int field);

// a * b * c, or UINT64_MAX when the product does not fit in 64 bits.  Sizes
// checked against a buffer go through this, so that a caller's huge count
// fails the check instead of wrapping below the buffer size.
inline uint64_t span_bytes(uint64_t a, uint64_t b, uint64_t c = 1)
{
    uint64_t r;
    if (__builtin_mul_overflow(a, b, &r) || __builtin_mul_overflow(r, c, &r)) return UINT64_MAX;
    return r;
}
// [off, off + bytes) lies inside a buffer of `size` bytes.
inline bool range_fits(uint64_t off, uint64_t bytes, uint64_t size) { return off <= size && bytes <= size - off; }

}  // namespace zkg
