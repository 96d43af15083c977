// zkg_internal.h -- declarations shared by the kernel and runtime translation
// units of libzkg.  Not part of the public C ABI (include/zkg.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "zkg_plans.h"

namespace zkg {

enum FieldId : int { FIELD_GL = 0, FIELD_BB = 1 };

constexpr int LOG_TW_TILE = 11;   // largest in-LDS transform: 2^11 points
constexpr int LOG_TW_TOP = 24;    // four-step / coset tables cover orders <= 2^24
constexpr int TW_SPLIT = 12;      // omega^e = lo[e & 4095] * hi[e >> 12]

// Inter-round twiddles of the 2^L plan, round I (not the last; at most two
// such rounds): omega_Mi^(a q) at rtw_offset(L, I) + q (Mi / R) + a, where Mi
// is the round's sub-problem size and R its radix.
__host__ __device__ constexpr int rtw_offset(int L, int I) { return (1 << (L + 1)) + (I ? (1 << L) : 0); }
constexpr int RTW_ENTRIES = 1 << (LOG_TW_TILE + 2);

// Resident twiddle tables of one field, register form (BabyBear Montgomery).
struct TwiddleSet {
    void *tile_fwd = nullptr;  // omega_2048^j, j < 1024
    void *tile_inv = nullptr;  // omega_2048^-j
    void *round_fwd[LDE_PLANS] = {}, *round_inv[LDE_PLANS] = {};  // RTW_ENTRIES inter-round twiddles per plan
    void *top_lo_fwd = nullptr, *top_hi_fwd = nullptr;  // omega_2^24^i, omega_2^24^(4096 i)
    void *top_lo_inv = nullptr, *top_hi_inv = nullptr;
};

// Coset weights for one (N, blowup, shift): chunk b of the LDE evaluates the
// coefficients times (s_b^k / N), s_b = shift * omega_{BN}^{rev_B(b)}, stored as
// lo[b][k & 4095] (with 1/N folded in) and hi[b][k >> 12].
struct CosetSet {
    void *lo = nullptr;  // B * 4096
    void *hi = nullptr;  // B * 1024
    uint32_t log_n = 0, log_blowup = 0;
};

// Poseidon2 constants in register form, flat [initial][internal][terminal].
struct P2Set {
    void *rc = nullptr;
    void *diag = nullptr;
    void *rc_f64 = nullptr;  // BabyBear only: canonical constants as doubles
    uint64_t seed = 0;
    bool loaded = false;
};

// One matrix of a mixed-height MMCS commit (device pointer, row-major).
struct MatDesc {
    const void *ptr;
    uint64_t width;
};

struct PassTimes;

// --- NTT / LDE (zkg_ntt.hip) -------------------------------------------------
// All matrices are row-major (height x width) in HBM, canonical values.
hipError_t launch_coset_lde(int field, const TwiddleSet &tw, const CosetSet &cs, const void *in,
                            void *out, void *tmp, uint64_t height, uint32_t width,
                            uint32_t log_blowup, hipStream_t st);
// Bytes of scratch that launch_coset_lde needs (0 for single-pass sizes).
uint64_t coset_lde_tmp_bytes(int field, uint64_t height, uint32_t width, uint32_t log_blowup);
hipError_t launch_ntt(int field, const TwiddleSet &tw, void *data, void *tmp, uint64_t height,
                      uint32_t width, int inverse, hipStream_t st);
uint64_t ntt_tmp_bytes(int field, uint64_t height, uint32_t width);
// Fills the RTW_ENTRIES inter-round table (register form) of one direction
// and plan from omega_2048 of that direction (register form).
void fill_round_twiddles(int plan, int field, uint64_t w2048, void *table);
// The plan a transform of this shape runs (zkg_tuning.hip: override > learned
// > static table > plan 0).
int lde_plan_for(int field, int log_n, uint32_t width, uint32_t log_blowup);
// The coset LDE with an explicit plan (the tuner's timing runs).
hipError_t launch_coset_lde_plan(int plan, int field, const TwiddleSet &tw, const CosetSet &cs, const void *in,
                                 void *out, void *tmp, uint64_t height, uint32_t width, uint32_t log_blowup,
                                 hipStream_t st);
// Work-form splits of the row hash and Merkle levels for (field, Poseidon2
// width), through the same layers (zkg_tuning.hip): log2 sizes, or levels
// per fused launch.
enum SplitKind { SPLIT_LANES = 0, SPLIT_CHUNKS = 1, SPLIT_HASH_CHUNKS = 2, SPLIT_FUSED = 3 };
int split_value(int field, int width, int which);
// One compiled instance of the transforms per plan (zkg_ntt.hip built with
// -DZKG_LDE_VARIANT=k into namespace lde_plan<k>).
#define ZKG_DECLARE_PLAN(NS)                                                                                       \
    namespace NS {                                                                                                 \
    hipError_t coset_lde(int field, const TwiddleSet &tw, const CosetSet &cs, const void *in, void *out, void *tmp,  \
                         uint64_t height, uint32_t width, uint32_t log_blowup, hipStream_t st);                       \
    hipError_t ntt(int field, const TwiddleSet &tw, void *data, void *tmp, uint64_t height, uint32_t width,          \
                   int inverse, hipStream_t st);                                                                   \
    void fill_round_twiddles(int field, uint64_t w2048, void *table);                                             \
    }
ZKG_DECLARE_PLAN(lde_plan0)
ZKG_DECLARE_PLAN(lde_plan1)
ZKG_DECLARE_PLAN(lde_plan2)
ZKG_DECLARE_PLAN(lde_plan3)
static_assert(LDE_PLANS == 4, "one compiled instance per plan");
// Unfused per-stage radix-2 path (one HBM pass per stage): the plain GPU
// denominator used for comparison with the fused path.
// tmp: height * width elements.
hipError_t launch_lde_plain(int field, const TwiddleSet &tw, const void *in, void *out, void *tmp,
                            uint64_t height, uint32_t width, uint32_t log_blowup, uint64_t shift,
                            hipStream_t st, bool sync_each);

// --- Poseidon2 / Merkle / FRI (zkg_hash.hip) -------------------------------
hipError_t launch_p2_permute(int field, int width, const P2Set &p2, void *states, uint64_t n,
                             hipStream_t st);
hipError_t launch_hash_rows(int field, int width, const P2Set &p2, const void *mat,
                            uint64_t height, uint32_t mat_width, void *digests, hipStream_t st);
// Builds levels 1.. of a power-of-two tree whose level 0 (n digests) is
// layers[0]; layers[i] receives n >> i digests.  fuse_levels = 1 gives one
// launch per level (the plain path).
hipError_t launch_merkle_tree(int field, int width, const P2Set &p2, void *const *layers,
                              int nlayers, uint64_t n_leaves, int fuse_levels, hipStream_t st,
                              bool sync_each);
// Several small trees built together (width 16, the FRI layer trees), W
// lanes per compression: one launch per 5 levels of all of them.
// Tree t: levels 1.. over layers[0] (n_leaves digests), nlayers =
// log2(n_leaves) + 1.
struct MerkleTop {
    void *const *layers;
    int nlayers;
    uint64_t n_leaves;
};
hipError_t launch_merkle_tops(int field, int width, const P2Set &p2, const MerkleTop *trees, int ntrees,
                              hipStream_t st);
// FRI layer trees of at most this many leaves are built together after the
// fold chain (ZKG_FRI_TOPS_LOG2 overrides the default 2^16).
uint64_t fri_tops_max_leaves();
hipError_t launch_gen(int field, void *out, uint64_t seed, uint64_t start, uint64_t n,
                      hipStream_t st);

// --- Shared launches over job tables (zkg_hash.hip) -------------------------
// Row hashes of many matrices in one persistent launch: job j hashes `rows`
// rows of `mat` (row-major, `width` elements per row) into `digests`; its
// tiles of HASH_TILE_ROWS rows are numbered tile0 .. tile0 + ceil(rows /
// HASH_TILE_ROWS) - 1, tile0 ascending in table order.  A grid sized to the
// chip takes tiles in that order from a device counter pair (next tile, blocks
// done), which the launch's last block resets to zero, so the same launch can
// run again (a graph replay) without clearing it.
constexpr int HASH_TILE_ROWS = 256;
struct HashJob {
    const void *mat;
    void *digests;
    uint64_t rows;
    uint32_t width;
    uint32_t tile0;
};
// One launch's worth of a tree's levels, as the single-tree commit plans them:
// `form` SEG_FUSED (one compression per lane, nb1 first-level compressions per
// block), SEG_CHUNKS (Goldilocks, one 4-chunk of the state per lane) or
// SEG_LANES (W lanes per compression); k levels from `in` into out[0..k-1];
// `blocks` blocks, the first of them `first` in a shared launch.
constexpr int SEG_MAX_LEVELS = 9;
enum SegForm : int32_t { SEG_FUSED = 0, SEG_CHUNKS = 1, SEG_LANES = 2 };
struct MerkleSeg {
    const void *in;
    void *out[SEG_MAX_LEVELS];
    uint32_t first, blocks;
    int32_t form, k, nb1, unused;
};
// The launches of one power-of-two tree (levels 1.. over layers[0] of n_leaves
// digests), in order: segment r needs segment r - 1 complete.
int merkle_plan(int field, int width, void *const *layers, int nlayers, uint64_t n_leaves,
                std::vector<MerkleSeg> &segs);
// jobs, ctr: device memory (ctr: two uint32_t, zero before the first launch).
hipError_t launch_hash_jobs(int field, int width, const P2Set &p2, const HashJob *jobs, int njobs, uint32_t ntiles,
                            uint32_t *ctr, hipStream_t st);
// segs: device memory, `first` ascending, nblocks = sum of blocks.
hipError_t launch_merkle_segs(int field, int width, const P2Set &p2, const MerkleSeg *segs, int nsegs,
                              uint32_t nblocks, hipStream_t st);
// Mixed-height MMCS.  Leaves: digest i = sponge over row i of every matrix of
// the group (device array of n MatDesc), rows concatenated.  Layer: node i of
// the next layer (i < half) = compress(prev[2i], prev[2i+1]), then, if n_inj,
// compress(node, group hash of row i) for i < inj_len, compress(node, 0) above.
hipError_t launch_mmcs_verify(int field, int width, const P2Set &p2, const MatDesc *desc, const int *gstart,
                              const int *glevel, int n_groups, uint32_t n_q, const uint64_t *idx, const void *proof,
                              int log_h, const void *root, uint32_t *ok, hipStream_t st);
hipError_t launch_mmcs_leaves(int field, int width, const P2Set &p2, const MatDesc *group, int n,
                              uint64_t rows, void *digests, hipStream_t st);
hipError_t launch_mmcs_layer(int field, int width, const P2Set &p2, const void *prev, uint64_t half,
                             void *next, const MatDesc *inj, int n_inj, uint64_t inj_len, hipStream_t st);
// FRI fold over EF4 (ext = 1) or a base field; twiddles from the resident
// inverse top tables.  If leaf_digests is not null, also hashes rows of the
// folded codeword viewed as a width-2 (ext) matrix into leaf digests (fused).
hipError_t launch_fri_fold(int ext, int field, const TwiddleSet &tw, const P2Set &p2,
                           const void *in, uint64_t n_out, const uint32_t *beta_mont,
                           void *out, void *leaf_digests, hipStream_t st);
// FriAccelerator semantics: explicit inverse twiddles.
hipError_t launch_fri_fold_explicit(int field, const void *in, uint64_t n_out, uint64_t alpha,
                                    const void *inv_tw, void *out, hipStream_t st);
hipError_t launch_gather_open(int field, int digest, const void *mat, uint32_t mat_width,
                              void *const *layers, int nlayers, const uint64_t *d_indices,
                              uint32_t n, void *rows_out, void *paths_out, hipStream_t st);
// One piece of an opening (a matrix row, a sibling digest): `words` 32-bit
// words from src to dst, both device memory.
struct GatherSeg {
    const uint32_t *src;
    uint32_t *dst;
    uint32_t words, pad;
};
// segs: device memory, n pieces, one block each.
hipError_t launch_gather_segs(const GatherSeg *segs, uint32_t n, hipStream_t st);

// --- DEEP quotient (zkg_deep.hip), BabyBear LDE, EF4 challenges -------------
// EF4 arguments are 4 words; "mont" = Montgomery form, otherwise canonical.
// inv: height EF4 = 1 / (z - x_i) (Montgomery), x_i = g w_H^rev(i).
hipError_t launch_deep_inv(const TwiddleSet &tw, void *inv, uint64_t height, const uint32_t *z_mont,
                           uint32_t g_mont, uint32_t w11, hipStream_t st);
// Blocks of the column-sum pass over n rows (partial: blocks * width EF4).
uint64_t deep_partial_blocks(uint64_t n);
// ys (width EF4, canonical) = scale * sum_{i<n} e_ik x_i inv_i.
hipError_t launch_deep_open_values(const TwiddleSet &tw, const void *lde, uint64_t height, uint32_t width,
                                   uint64_t n, const void *inv, const uint32_t *z_mont, uint32_t g_mont,
                                   uint32_t w11, const uint32_t *scale_mont, void *partial, void *ys,
                                   hipStream_t st);
// ro_i += (ry - sum_k c_k e_ik) * inv_i; coef: width x 8 doubles, the 16-bit
// halves [lo0..lo3, hi0..hi3] of the canonical c_k = alpha^off alpha^k; ry and
// ro canonical.
hipError_t launch_deep_reduce(const void *lde, uint64_t height, uint32_t width, const void *inv,
                              const void *coef, const uint32_t *ry, uint32_t w11, void *ro, hipStream_t st);

// Goldilocks with base-field z (canonical u64 throughout): inv (height u64),
// opened values ys (width u64) = scale * sum_{i<n} e_ik x_i inv_i, and
// ro_i += (ry - sum_k coef_k e_ik) * inv_i.
hipError_t launch_deep_gl_inv(const TwiddleSet &tw, void *inv, uint64_t height, uint64_t z, uint64_t g,
                              hipStream_t st);
hipError_t launch_deep_gl_open_values(const TwiddleSet &tw, const void *lde, uint64_t height, uint32_t width,
                                      uint64_t n, const void *inv, uint64_t g, uint64_t scale, void *partial,
                                      void *ys, hipStream_t st);
hipError_t launch_deep_gl_reduce(const void *lde, uint64_t height, uint32_t width, const void *inv,
                                 const void *coef, uint64_t ry, void *ro, hipStream_t st);

// Register-form conversion helpers on the host.
uint32_t bb_to_mont_host(uint32_t x);
uint32_t bb_from_mont_host(uint32_t x);

}  // namespace zkg
