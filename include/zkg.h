/*
 * zkg.h -- C ABI of libzkg, the MI355X (gfx950) STARK trace-commitment
 * accelerator.  It stands where ZirOS names a `zkf-gpu` crate (README.md:206,
 * absent from the tree) and replaces the Metal entry points of zkf-metal for
 * the coset LDE, Poseidon2 row hashing, Merkle commitment and FRI folding.
 *
 * Conventions (after zkf-ffi/src/lib.rs:19-74,164-180): opaque handles; every
 * function returns int, 0 on success and a negative ZKG_E* code on failure,
 * with a thread-local message readable through zkg_last_error().  The caller
 * owns host memory; device buffers and trees belong to the library until the
 * caller frees them.  A context may be used by one submitting thread at a time;
 * the serving loop owns an internal thread of its own.
 *
 * Field elements cross the boundary in canonical form, little-endian:
 * Goldilocks as uint64_t, BabyBear as uint32_t, a BabyBear degree-4 extension
 * element (x^4 - 11) as 4 uint32_t coefficients, lowest first.  Matrices are
 * row-major (height x width).  Heights are powers of two.
 */
#ifndef ZKG_H
#define ZKG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZKG_ABI_VERSION 1

enum zkg_field { ZKG_GOLDILOCKS = 0, ZKG_BABYBEAR = 1 };

enum zkg_status {
    ZKG_OK = 0,
    ZKG_EINVAL = -1,   /* bad argument (shape, field, width, null) */
    ZKG_EDEVICE = -2,  /* HIP runtime error */
    ZKG_ENOMEM = -3,   /* device allocation failed or HBM budget exceeded */
    ZKG_ESTATE = -4,   /* constants not loaded, serving loop not running, ... */
    ZKG_EBUSY = -5,    /* serve ticket not finished yet */
    ZKG_EFALLBACK = -6, /* the dispatch driver does not take this node: the CPU may */
};

typedef struct zkg_ctx zkg_ctx;
typedef struct zkg_buf zkg_buf;
typedef struct zkg_tree zkg_tree;
typedef struct zkg_fri zkg_fri;

/* ---- errors and versions ------------------------------------------------ */
/* Replaces zkf-ffi's zkf_last_error_message (zkf-ffi/src/lib.rs:164-180). */
const char *zkg_last_error(void);
int zkg_abi_version(void);

/* ---- device context ------------------------------------------------------ */
/* Replaces the process-global MetalContext (zkf-metal/src/device.rs:1469-1480).
 * hbm_budget_bytes = 0 means 288 GiB.  Twiddle tables for both fields become
 * resident at creation. */
int zkg_ctx_create(int device_ordinal, uint64_t hbm_budget_bytes, zkg_ctx **out);
int zkg_ctx_destroy(zkg_ctx *ctx);
int zkg_ctx_synchronize(zkg_ctx *ctx);
/* Bytes of HBM this context currently holds (buffers, trees, tables). */
int zkg_ctx_hbm_used(const zkg_ctx *ctx, uint64_t *bytes);

/* ---- device buffers ------------------------------------------------------ */
/* Replaces GpuBufferAllocator::{alloc_shared, free_shared}
 * (zkf-runtime/src/buffer_bridge.rs:77-84); HBM is discrete memory, so a
 * handle comes back rather than a host pointer. */
int zkg_buf_alloc(zkg_ctx *ctx, uint64_t bytes, zkg_buf **out);
int zkg_buf_free(zkg_buf *buf);
int zkg_buf_upload(zkg_buf *buf, uint64_t offset, const void *src, uint64_t bytes);
int zkg_buf_download(const zkg_buf *buf, uint64_t offset, void *dst, uint64_t bytes);
int zkg_buf_size(const zkg_buf *buf, uint64_t *bytes);
/* Fills n elements from element `start` of the synthetic stream `seed`:
 * element i is the first v_a = mix64(seed + (i+1) * 0x9e3779b97f4a7c15 +
 * a * 0x632be59bd9b4e019), a = 0, 1, ..., (>> 33 for BabyBear) with v_a < p,
 * mix64 being the SplitMix64 finaliser. */
int zkg_fill_random(zkg_ctx *ctx, int field, zkg_buf *buf, uint64_t seed, uint64_t start, uint64_t n);

/* ---- Poseidon2 constants ------------------------------------------------ */
/* Keeps the round constants of Poseidon2<field, width> derived from
 * SmallRng::seed_from_u64(seed) resident (the reference rebuilds them per call:
 * zkf-metal/src/poseidon2/goldilocks.rs:42-83, babybear.rs:76-111).
 * Supported: (GOLDILOCKS, 16), (GOLDILOCKS, 8), (BABYBEAR, 16). */
int zkg_constants_load(zkg_ctx *ctx, int field, int width, uint64_t seed);
/* Replaces the constants with an explicit flat [initial][internal][terminal]
 * table (rounds_f * width + rounds_p canonical values), as the reference's
 * Poseidon2Accelerator passes them. */
int zkg_constants_set(zkg_ctx *ctx, int field, int width, const void *flat, uint32_t n_external_rounds,
                      uint32_t n_internal_rounds);

/* ---- transforms ---------------------------------------------------------- */
/* p3 TwoAdicSubgroupDft::dft_batch / idft_batch on every column, in place,
 * natural order in and out, inverse scaled by 1/height.  Replaces MetalDft
 * (zkf-metal/src/ntt/p3_adapter.rs:69-75,78-106). */
int zkg_ntt_batch(zkg_ctx *ctx, int field, zkg_buf *buf, uint64_t height, uint32_t width, int inverse);
/* No counterpart in the reference: zkg_ntt_batch for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The four-step transposition buffer
 * (height x width elements for heights above 2^11) lives in a fixed device
 * array of ZKG_NOALLOC_SCRATCH_BYTES per GPU instead of an allocation; larger
 * transforms are ZKG_EINVAL.  Stream-ordered: returns once enqueued on the
 * context's stream (zkg_buf_download and zkg_ctx_synchronize wait for it).
 * All calls on one GPU share the array: calls on different contexts of one
 * GPU that overlap in time race on it. */
#define ZKG_NOALLOC_SCRATCH_BYTES (64ull << 20)
int zkg_ntt_batch_noalloc(zkg_ctx *ctx, int field, zkg_buf *buf, uint64_t height, uint32_t width, int inverse);
/* Coset LDE as TwoAdicFriPcs::commit uses RuntimeDft
 * (zkf-backends/src/plonky3.rs:101-257): out (height << log_blowup rows) =
 * bit_reverse_rows(coset_lde_batch(in, log_blowup, shift)).  in is unchanged. */
int zkg_coset_lde_batch(zkg_ctx *ctx, int field, const zkg_buf *in, uint64_t height, uint32_t width,
                        uint32_t log_blowup, uint64_t shift, zkg_buf *out);

/* ---- Poseidon2 ----------------------------------------------------------- */
/* Replaces Poseidon2Accelerator::batch_permute_* (zkf-core/src/acceleration.rs:
 * 110-140): n states of `width` elements, in place, with the loaded constants. */
int zkg_poseidon2_permute(zkg_ctx *ctx, int field, int width, zkg_buf *states, uint64_t n);
/* Host-memory form with an explicit constant table, exactly the trait's
 * batch_permute_goldilocks / batch_permute_babybear (width 16). */
int zkg_poseidon2_batch_permute_goldilocks(zkg_ctx *ctx, uint64_t *states, uint64_t n_perms,
                                           const uint64_t *round_constants, uint32_t n_external_rounds,
                                           uint32_t n_internal_rounds);
int zkg_poseidon2_batch_permute_babybear(zkg_ctx *ctx, uint32_t *states, uint64_t n_perms,
                                         const uint32_t *round_constants, uint32_t n_external_rounds,
                                         uint32_t n_internal_rounds);
/* PaddingFreeSponge<P, width, width/2, width/2> of every row of a matrix:
 * digests receives height * width/2 elements. */
int zkg_hash_rows(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *mat, uint64_t height,
                  uint32_t mat_width, zkg_buf *digests);

/* ---- Merkle commitment (Mmcs) -------------------------------------------- */
/* Replaces MetalMerkleTreeMmcs::commit (zkf-metal/src/merkle.rs:76-105) for a
 * single matrix of power-of-two height: digest layer 0 = row hashes, layer i+1
 * = TruncatedPermutation of pairs (vendor/p3-merkle-tree-gpu/src/
 * merkle_tree.rs:84-151).  The tree keeps a reference to `mat` for openings:
 * the caller keeps it alive until zkg_tree_free. */
int zkg_merkle_commit(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *mat, uint64_t height,
                      uint32_t width, zkg_tree **out);
int zkg_tree_root(const zkg_tree *tree, void *root_out);
int zkg_tree_num_layers(const zkg_tree *tree, uint32_t *n);
int zkg_tree_digest_elems(const zkg_tree *tree, uint32_t *n);
/* Digests in layer `layer` (height >> layer for single-matrix trees; the
 * padded length for zkg_mmcs_commit trees). */
int zkg_tree_layer_len(const zkg_tree *tree, uint32_t layer, uint64_t *n);
/* Copies digest layer `layer` (zkg_tree_layer_len digests) to host memory. */
int zkg_tree_layer(const zkg_tree *tree, uint32_t layer, void *out, uint64_t out_bytes);
/* Mmcs::open_batch (vendor/p3-merkle-tree-gpu/src/mmcs.rs:179-210) for n
 * indices: rows_out gets n * width elements, paths_out n * (layers-1) digests. */
int zkg_tree_open(const zkg_tree *tree, const uint64_t *indices, uint32_t n, void *rows_out, void *paths_out);
int zkg_tree_free(zkg_tree *tree);

/* Mmcs::commit over several matrices of possibly different, not necessarily
 * power-of-two heights (vendor/p3-merkle-tree-gpu/src/merkle_tree.rs:84-151,
 * compress_and_inject :272-350).  Matrices are taken tallest first (stable);
 * the rows of the tallest ones, concatenated in commit order, are hashed into
 * layer 0, padded with zero digests to even length; each next layer
 * compresses pairs (padded to even, a single root last) and, where a shorter
 * group's heights round up to that layer's size, compresses the result with
 * the hash of the group's rows.  Heights that round up to the same power of
 * two must be equal (as p3 asserts).  The tree references the matrices for
 * openings: the caller keeps them alive until zkg_tree_free.
 * zkg_tree_root/num_layers/layer/free apply; zkg_tree_layer copies the
 * layer's padded length. */
int zkg_mmcs_commit(zkg_ctx *ctx, int field, int p2_width, uint32_t n_mats, const zkg_buf *const *mats,
                    const uint64_t *heights, const uint32_t *widths, zkg_tree **out);
/* Mmcs::open_batch (vendor/p3-merkle-tree-gpu/src/mmcs.rs:179-210) for n
 * indices below the tallest height: per index, matrix j's row
 * index >> (log2_ceil(max_h) - log2_ceil(h_j)) for every matrix in commit
 * order (sum of widths elements), and log2_ceil(max_h) sibling digests.
 * Also accepts single-matrix trees. */
int zkg_mmcs_open(const zkg_tree *tree, const uint64_t *indices, uint32_t n, void *rows_out, void *paths_out);
/* Mmcs::verify_batch (vendor/p3-merkle-tree-gpu/src/mmcs.rs:210-330) for n
 * openings at once, on the device: ok[q] = 1 if opening q (its rows and
 * path, laid out as zkg_mmcs_open writes them: every matrix's row in the given
 * order, then log2_ceil(max height) sibling digests) hashes to `root`, else 0.
 * The matrices are given by heights and widths, in commit order.  ZKG_EINVAL
 * for what p3 refuses with an error (heights that differ but round up to the
 * same power of two, an index at or past the tallest height) and for
 * non-canonical values. */
int zkg_mmcs_verify_batch(zkg_ctx *ctx, int field, int p2_width, const void *root, uint32_t n_mats,
                          const uint64_t *heights, const uint32_t *widths, uint32_t n, const uint64_t *indices,
                          const void *rows, const void *paths, uint8_t *ok);

/* ---- PCS opening: DEEP quotient --------------------------------------------
 * Replaces the reduced-opening step of p3-fri 0.4.2 TwoAdicFriPcs::open and
 * p3-interpolation interpolate_coset (external; the reference's node is
 * ProverOp::FriQueryOpen, zkf-runtime/src/graph.rs:105-108).  `lde` is a
 * committed LDE as zkg_commit/zkg_coset_lde_batch leave it: height H = 2^k
 * rows in bit-reversed order on the coset shift * H_H, so row i holds
 * p(x_i), x_i = shift * w_H^rev(i).  Challenge-field values (z, alpha, opened
 * values, ro entries):
 *   ZKG_BABYBEAR   EF4 = x^4 - 11, 4 canonical u32 coefficients, low
 *                  first (base-field challenges, the reference's
 *                  Challenge<F> = F of zkf-backends/src/plonky3.rs:55,
 *                  are EF4 values with zero upper coefficients);
 *   ZKG_GOLDILOCKS Goldilocks itself, one canonical u64 (Challenge<F> = F).
 * z must not lie on the domain.
 *
 * Opened values: ys_out[k] = p_k(z) for every column k, by barycentric
 * interpolation over the first H >> log_blowup rows. */
int zkg_deep_open_values(zkg_ctx *ctx, int field, const zkg_buf *lde, uint64_t height, uint32_t width,
                         uint32_t log_blowup, uint64_t shift, const void *z, void *ys_out);
/* Reduced opening: ro[i] += alpha^alpha_offset * sum_k alpha^k (p_k(x_i) - ys[k]) / (x_i - z)
 * for every row (ro: H challenge-field values, canonical, accumulated in place). */
int zkg_deep_reduce(zkg_ctx *ctx, int field, const zkg_buf *lde, uint64_t height, uint32_t width, uint64_t shift,
                    const void *z, const void *alpha, uint64_t alpha_offset, const void *ys, zkg_buf *ro);
/* Both, sharing one pass of inverse denominators: ys_out as above, then ro. */
int zkg_deep_open(zkg_ctx *ctx, int field, const zkg_buf *lde, uint64_t height, uint32_t width, uint32_t log_blowup,
                  uint64_t shift, const void *z, const void *alpha, uint64_t alpha_offset, void *ys_out, zkg_buf *ro);

/* ---- polynomial operations (Goldilocks) -------------------------------------
 * Replace PolyOpsAccelerator (zkf-core/src/acceleration.rs:173-192; Metal
 * zkf-metal/src/poly.rs:37-246).  Host memory in and out, canonical values
 * (ZKG_EINVAL otherwise); each call returns when `out` is written.  Unlike the
 * Metal accelerator there is no minimum size (it declines fewer than 1024
 * points or evaluations).
 * batch_eval: out[i] = sum_k coeffs[k] points[i]^k (Horner). */
int zkg_poly_eval_goldilocks(zkg_ctx *ctx, const uint64_t *coeffs, uint64_t n_coeffs, const uint64_t *points,
                             uint64_t n_points, uint64_t *out);
/* coset_eval: out[i] = sum_k coeffs[k] (shift w^i)^k for i < 2^log_n, w the
 * two-adic generator of order 2^log_n (natural order; 2^log_n coefficients,
 * log_n <= 22). */
int zkg_poly_coset_eval_goldilocks(zkg_ctx *ctx, const uint64_t *coeffs, uint32_t log_n, uint64_t shift,
                                   uint64_t *out);
/* quotient: out[i] = (evals[i] - f_z) / (w^i - z) for i < n (a power of two,
 * w of order n), and 0 where w^i = z. */
int zkg_poly_quotient_goldilocks(zkg_ctx *ctx, const uint64_t *evals, uint64_t n, uint64_t z, uint64_t f_z,
                                 uint64_t *out);

/* ---- constraint evaluation ---------------------------------------------------
 * Replaces ConstraintEvalAccelerator::eval_trace_goldilocks (zkf-core/src/
 * acceleration.rs:219-242; Metal body zkf-metal/src/constraint_eval.rs:115-190
 * and shaders/constraint_eval.metal:1-74).  The bytecode is the reference's:
 * word = op << 24 | operand (24 bits), op 0 CONST (push constants[operand]),
 * 1 LOAD (push trace[row][operand]), 2 ADD, 3 SUB, 4 MUL (pop b, pop a, push
 * a op b), 5 DUP, 6 EMIT (pop into out[row][operand]); other opcodes do
 * nothing.  Every row of the row-major Goldilocks trace (n_rows x width,
 * canonical) runs the whole program; out is n_rows x n_constraints and
 * constraints the program never emits read zero.  Unlike the reference,
 * programs are checked before they run: a stack underflow, a stack deeper
 * than the reference's 32 entries, or a column, constant or constraint index
 * out of range is ZKG_EINVAL (the reference would read or write out of
 * bounds), and so is a non-canonical constant.  Constant and constraint
 * indices are limited to 2^22, column indices to the reference's 2^24. */
typedef struct zkg_constraint_program zkg_constraint_program;
/* Check and translate a program once, upload it to the context's device. */
int zkg_constraint_program_create(zkg_ctx *ctx, uint32_t width, const uint32_t *bytecode, uint32_t n_instructions,
                                  const uint64_t *constants, uint32_t n_constants, uint32_t n_constraints,
                                  zkg_constraint_program **out);
/* Evaluate over n_rows >= 1 rows of a device trace into a device output;
 * stream-ordered on the context's stream (zkg_buf_download waits for it). */
int zkg_constraint_program_eval(zkg_constraint_program *prog, const zkg_buf *trace, uint64_t n_rows, zkg_buf *out);
/* info: device words, panel slots, panel stages per wave, LDS stack entries,
 * LDS bytes per wave, columns per panel, rows per lane, LDS bytes per panel
 * column. */
int zkg_constraint_program_info(const zkg_constraint_program *prog, uint32_t info[8]);
/* Quotient values over a committed Goldilocks LDE (height H rows, row i the
 * trace polynomials at x_i = shift * w_H^rev(i), as zkg_coset_lde_batch /
 * zkg_commit leave it): q_out[i] = (sum_k coef[k] C_k(row i)) / (x_i^N - 1),
 * N = H >> log_blowup, C_k the constraint values above, coef the caller's
 * combination (canonical, n_constraints of them; a prover passes powers of its
 * challenge alpha in its own order).  Dividing by the trace domain's vanishing
 * polynomial Z_H is the quotient step of the STARK prover that precedes the
 * quotient commitment.  A constraint emitted more than once counts with its
 * last value.  ZKG_EINVAL if the coset meets the trace domain.  Returns when
 * q_out is written. */
int zkg_constraint_program_quotient(zkg_constraint_program *prog, const zkg_buf *lde, uint64_t height,
                                    uint32_t log_blowup, uint64_t shift, const uint64_t *coef, zkg_buf *q_out);
/* Compile the program once into straight-line kernels for gfx950 (hiprtc,
 * about a second each, evaluation and quotient); later evals and quotients of
 * this handle run them instead of the interpreter, with identical results.
 * ZKG_EDEVICE carries the compiler log. */
int zkg_constraint_program_compile(zkg_constraint_program *prog);
void zkg_constraint_program_free(zkg_constraint_program *prog);
/* One-shot form of the calls above; returns when out is written.  The
 * context remembers the programs it is given (by content), so a program passed
 * again skips the check and upload, and is compiled on its second use (up to
 * 16384 instructions; longer ones stay interpreted unless compiled through a
 * handle); results are the same either way. */
int zkg_constraint_eval_gl(zkg_ctx *ctx, const zkg_buf *trace, uint64_t n_rows, uint32_t width,
                           const uint32_t *bytecode, uint32_t n_instructions, const uint64_t *constants,
                           uint32_t n_constants, uint32_t n_constraints, zkg_buf *out);
/* The context's program memory for zkg_constraint_eval_gl: keep up to
 * max_programs (least recently used dropped; 0 turns it off) and compile a
 * program on its compile_after_uses-th use (0: never).  Default 16 and 2.
 * Reconfiguring drops the remembered programs. */
int zkg_constraint_cache_config(zkg_ctx *ctx, uint32_t max_programs, uint32_t compile_after_uses);
/* info: programs remembered, of those compiled, total uses. */
int zkg_constraint_cache_info(const zkg_ctx *ctx, uint32_t info[3]);
/* The translation alone, on the host with no device (for evaluation, or for
 * the quotient when quotient != 0): writes up to cap device words and the
 * info array of zkg_constraint_program_info (info[0] = words needed;
 * ZKG_EINVAL if cap is smaller). */
int zkg_constraint_translate(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions, uint32_t n_constants,
                             uint32_t n_constraints, int quotient, uint32_t *words, uint32_t cap, uint32_t info[8]);
/* The HIP source zkg_constraint_program_compile would compile (*size bytes
 * including the terminator; up to cap written). */
int zkg_constraint_jit_source(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions,
                              const uint64_t *constants, uint32_t n_constants, uint32_t n_constraints, char *src,
                              uint64_t cap, uint64_t *size);
/* Translate and compile both kernels on the host with no device (the
 * compiler's log, if any, into log). */
int zkg_constraint_compile_check(uint32_t width, const uint32_t *bytecode, uint32_t n_instructions,
                                 const uint64_t *constants, uint32_t n_constants, uint32_t n_constraints, char *log,
                                 uint32_t log_cap);

/* Fused commit of one trace: coset LDE at log_blowup with the field's
 * generator shift (p3 Val::GENERATOR: 7 or 31), row hash, tree.  lde_out
 * (height << log_blowup rows) receives the LDE and backs the tree's openings. */
int zkg_commit(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *trace, uint64_t height, uint32_t width,
               uint32_t log_blowup, zkg_buf *lde_out, zkg_tree **out);
/* The same commit through the unfused, reference-structured path: one radix-2
 * global pass per NTT stage, per-element shift powers, a bit-reversal pass,
 * one launch per Merkle level, a stream synchronisation after every launch
 * (zkf-metal/src/ntt/p3_adapter.rs:224-258, merkle.rs:324-443). */
int zkg_commit_plain(zkg_ctx *ctx, int field, int p2_width, const zkg_buf *trace, uint64_t height,
                     uint32_t width, uint32_t log_blowup, zkg_buf *lde_out, zkg_tree **out);

/* ---- FRI ------------------------------------------------------------------ */
/* FriAccelerator::fold_goldilocks semantics (zkf-core/src/acceleration.rs:
 * 196-217): out[i] = (e[2i] + e[2i+1])/2 + alpha (e[2i] - e[2i+1]) inv_tw[i]. */
int zkg_fri_fold(zkg_ctx *ctx, int field, const zkg_buf *evals, uint64_t n_out, uint64_t alpha,
                 const zkg_buf *inv_twiddles, zkg_buf *out);
/* p3-fri TwoAdicFriFolding::fold_matrix, arity 2: evals holds 2 n_out values
 * in bit-reversed order; ext = 1 for BabyBear EF4 (beta: 4 uint32_t), 0 for
 * the base field (beta: one element as uint64_t). */
int zkg_fri_fold_matrix(zkg_ctx *ctx, int ext, int field, const zkg_buf *evals, uint64_t n_out,
                        const void *beta, zkg_buf *out);
/* FRI commit phase over EF4 (BabyBear Poseidon2-16 MMCS): every layer of
 * length 2^log_len ... 2^(log_final+1) is committed as a width-2 matrix and
 * folded with betas[4 r .. 4 r + 3].  The codeword buffer is read, not kept. */
int zkg_fri_commit_phase(zkg_ctx *ctx, const zkg_buf *codeword, uint32_t log_len, uint32_t log_final,
                         const uint32_t *betas, zkg_fri **out);
/* FRI commit phase over a base field (ZKG_GOLDILOCKS, or ZKG_BABYBEAR with
 * base-field challenges): p3-fri's commit phase with Challenge = Val, the
 * reference's configuration (`type Challenge<F> = F`, zkf-backends/src/
 * plonky3.rs:55, with create_test_fri_params at :597).  Each layer of
 * 2^log_len ... 2^(log_final+1) canonical field elements is committed as a
 * width-2 matrix under the field's Poseidon2-16 MMCS and folded (arity 2)
 * with betas[r], a canonical element.  The codeword buffer is read, not kept. */
int zkg_fri_commit_phase_base(zkg_ctx *ctx, int field, const zkg_buf *codeword, uint32_t log_len,
                              uint32_t log_final, const uint64_t *betas, zkg_fri **out);
/* The commit phase one layer at a time, in p3-fri's transcript order (p3-fri
 * 0.4.2 commit_phase: commit layer r, observe its root, sample beta_r, fold),
 * which the two calls above cannot follow because they take every beta at
 * once.  zkg_fri_begin commits layer 0 of the codeword (ext = 1: EF4 over
 * ZKG_BABYBEAR; ext = 0: the base field `field`); each zkg_fri_fold_next folds
 * the newest layer with beta (ext = 1: 4 canonical uint32_t; ext = 0: one
 * canonical element as uint64_t, as zkg_fri_fold_matrix takes it) and commits
 * the layer it produces, until log_len - log_final calls leave the final
 * codeword.  Layer r's root can be read once r folds are done; zkg_fri_final
 * and zkg_fri_open need the whole phase.  Refused (ZKG_ESTATE) after the
 * Poseidon2 constants are reloaded mid-phase.  The result equals the one-call
 * commit phase with the same betas. */
int zkg_fri_begin(zkg_ctx *ctx, int ext, int field, const zkg_buf *codeword, uint32_t log_len, uint32_t log_final,
                  zkg_fri **out);
int zkg_fri_fold_next(zkg_fri *fri, const void *beta);
int zkg_fri_num_layers(const zkg_fri *fri, uint32_t *n);
/* layer root: 8 field elements (uint32_t for BabyBear, uint64_t for
 * Goldilocks; root8 points to 32 or 64 bytes) */
int zkg_fri_layer_root(const zkg_fri *fri, uint32_t layer, uint32_t *root8);
/* final codeword: 2^log_final elements (EF4: 4 uint32_t each; base field:
 * one uint32_t or uint64_t each) */
int zkg_fri_final(const zkg_fri *fri, uint32_t *out);
/* Query openings as p3-fri's answer_query makes them: for a query index q of
 * the first codeword (q < 2^log_len), layer r opens leaf (q >> (r+1)) of its
 * width-2 matrix.  rows_out gets, for each query and layer, the opened pair
 * (EF4: 8 uint32_t; Goldilocks: 2 uint64_t; base-field BabyBear: 2
 * uint32_t); paths_out gets, per query, the sibling paths of layers 0, 1, ...
 * concatenated (layer r has log_len - 1 - r digests of 8 field elements). */
int zkg_fri_open(const zkg_fri *fri, const uint64_t *indices, uint32_t n, uint32_t *rows_out,
                 uint32_t *paths_out);
int zkg_fri_free(zkg_fri *fri);

/* ---- batched commit ------------------------------------------------------ */
typedef struct zkg_job {
    int field;            /* ZKG_GOLDILOCKS or ZKG_BABYBEAR */
    int p2_width;         /* 16 (8 for Goldilocks width-8) */
    const zkg_buf *trace; /* height x width, canonical */
    uint64_t height;
    uint32_t width;
    uint32_t log_blowup;
    zkg_buf *lde_out;     /* (height << log_blowup) x width */
} zkg_job;

/* Commits n traces together (replaces BatchProver::prove_all's thread per
 * job, zkf-metal/src/batch_prover.rs:167-274).  The coset LDEs run on parallel
 * streams; then, per Poseidon2 instance, ONE persistent row-hash launch walks a
 * device table of every job's rows (a grid sized to the chip takes 256-row
 * tiles, heaviest first, from a device counter), and each round of Merkle
 * levels is one launch over all the trees.  The whole pipeline is captured
 * once as a hipGraph per job set (shapes and buffers) and replayed; with
 * timing on (zkg_ctx_set_timing) the same launches run straight from the
 * streams and zkg_ctx_stage_ms reports them.  trees_out[i] receives job i's
 * tree (a copy: later batches never write it). */
int zkg_commit_batch(zkg_ctx *ctx, const zkg_job *jobs, uint32_t n, zkg_tree **trees_out);

/* ---- serving -------------------------------------------------------------- */
/* Continuous-batching request loop: submitted jobs are admitted while the HBM
 * budget and in-flight slots allow, packed into shared launches with the
 * other admitted jobs, and run on one of `slots` streams so batches overlap.
 * Each result is identical to a lone zkg_commit of the same job. */
int zkg_serve_start(zkg_ctx *ctx, uint32_t slots, uint32_t max_batch);
/* Copies the job description; the trace and lde_out buffers must stay alive
 * until the ticket is polled done. */
int zkg_serve_submit(zkg_ctx *ctx, const zkg_job *job, uint64_t *ticket);
/* ZKG_OK and *tree set when done, ZKG_EBUSY while queued or running. */
int zkg_serve_poll(zkg_ctx *ctx, uint64_t ticket, zkg_tree **tree, double *latency_ms);
int zkg_serve_stop(zkg_ctx *ctx);

/* ---- runtime dispatch driver ------------------------------------------------
 * The UMPG GPU lane over libzkg: zkf-runtime's GpuDispatchDriver
 * (zkf-runtime/src/metal_driver.rs:55-93) executing commit-path ProverOps
 * (graph.rs:69-108) with the node payloads of execution.rs:51-106, and the
 * BufferBridge slots allocated as device buffers (GpuBufferAllocator::
 * alloc_shared / free_shared, buffer_bridge.rs:77-84).  Each arm keeps the
 * semantics of the Metal driver's arm for that op (metal_dispatch_macos.rs:
 * 305-779); LDE and the BabyBear Poseidon batch and Merkle layer are payloads
 * the reference lacks, with the same conventions. */
typedef struct zkg_bridge zkg_bridge;
int zkg_bridge_create(zkg_ctx *ctx, zkg_bridge **out);
void zkg_bridge_destroy(zkg_bridge *bridge);
/* alloc_shared: slot (nonzero; 0 means "no buffer") gets `bytes` of HBM,
 * replacing what it held; free_shared releases it. */
int zkg_bridge_alloc(zkg_bridge *bridge, uint32_t slot, uint64_t bytes);
int zkg_bridge_free(zkg_bridge *bridge, uint32_t slot);
/* Host bytes into a slot (sized to them) / a slot's bytes out (*bytes = its
 * size; nothing copied when cap is 0). */
int zkg_bridge_write(zkg_bridge *bridge, uint32_t slot, const void *src, uint64_t bytes);
int zkg_bridge_read(const zkg_bridge *bridge, uint32_t slot, void *dst, uint64_t cap, uint64_t *bytes);
/* The device buffer behind a slot, for the other entry points (NULL if none). */
zkg_buf *zkg_bridge_buffer(zkg_bridge *bridge, uint32_t slot);

enum zkg_op_kind {
    ZKG_OP_NTT = 1,          /* ProverOp::Ntt, NttGoldilocks { values_slot } */
    ZKG_OP_LDE = 2,          /* ProverOp::Lde, Lde{Goldilocks,BabyBear} { values_slot, blowup } */
    ZKG_OP_POSEIDON_BATCH = 3, /* ProverOp::PoseidonBatch, Poseidon{Goldilocks,BabyBear} */
    ZKG_OP_MERKLE_LAYER = 4, /* ProverOp::MerkleLayer, Merkle{Goldilocks,BabyBear} (Poseidon2) */
    ZKG_OP_FRI_FOLD = 5,     /* ProverOp::FriFold, FriFold{Goldilocks,BabyBear} */
};
typedef struct zkg_node {
    uint32_t op;        /* zkg_op_kind */
    int32_t field;      /* payload variant: ZKG_GOLDILOCKS or ZKG_BABYBEAR */
    uint64_t size;      /* Ntt/FriFold size, Lde rows, PoseidonBatch count, MerkleLayer leaf_count */
    uint32_t width;     /* PoseidonBatch width; columns of an Ntt or Lde (1 for a vector) */
    uint32_t blowup;    /* Lde */
    uint32_t inverse;   /* Ntt */
    uint32_t n_ext, n_int; /* PoseidonBatch rounds (with constants in aux_slot) */
    uint64_t seed;      /* MerkleLayer: the runtime's Poseidon seed */
    uint32_t in_slot;   /* values / states / leaves / evals */
    uint32_t aux_slot;  /* PoseidonBatch round constants (0: seed 42); FriFold alpha */
    uint32_t aux2_slot; /* FriFold inverse twiddles */
    uint32_t out_slot;  /* the primary output */
    uint64_t out_bytes; /* MerkleLayer: the output buffer's size, which sets the digest width */
} zkg_node;
typedef struct zkg_telemetry {
    char kernel[48];
    double ms;
    uint64_t input_bytes, output_bytes;
} zkg_telemetry;
/* GpuDispatchDriver::execute: ZKG_OK with telemetry; ZKG_EFALLBACK (the
 * FallbackSignal, reason in zkg_last_error) when this lane does not take the
 * node; another negative code for a hard failure.  Returns when the output
 * slot is written. */
int zkg_driver_execute(zkg_ctx *ctx, const zkg_node *node, zkg_bridge *bridge, zkg_telemetry *telemetry);
/* should_try_on_either: 1 if the node is large enough to be worth a launch. */
int zkg_driver_should_try(const zkg_node *node);

/* ---- tuning ------------------------------------------------------------------
 * After the reference's device tuning (zkf-metal/src/tuning.rs:84-111,276-365:
 * a runtime override beats learned values, which beat the static device
 * profile; learned values from zkf-runtime/src/adaptive_tuning.rs).  Two
 * process-wide settings resolve that way:
 *   - the tile plan of every column transform (NTT, coset LDE), per shape
 *     (field, log2 height, width class = ceil(log2 width), log2 blowup, 0 for
 *     a plain NTT).  Plans differ in radix split, tile store unrolling and
 *     pass-B occupancy; all give identical values.  Static layer: the table
 *     ziros_amd/tune.py measured on gfx950, nearest measured shape; then plan 0;
 *   - the work-form splits of the row hash and Merkle levels per (field,
 *     Poseidon2 width): levels of at most 2^merkle_lanes_log2 digests use W
 *     lanes per compression, Goldilocks levels of at most
 *     2^merkle_chunks_log2 one 4-chunk of the state per lane, Goldilocks row
 *     hashes of at most 2^hash_chunks_log2 rows a 4-chunk per lane, and
 *     fused_levels levels per one-compression-per-lane launch.  The
 *     environment variables ZKG_MERKLE_LANES_LOG2, ZKG_MERKLE_CHUNKS_LOG2,
 *     ZKG_HASH_CHUNKS_LOG2 and ZKG_MERKLE_FUSED_K seed the override layer.
 * Override and learned entries match exactly; a field of -1 matches anything
 * (split values of -1 fall through to the next layer).  Setting a layer
 * replaces it; n = 0 clears it.  A change applies to launches made after it
 * (a cached batch graph keeps the plans it was captured with). */
enum zkg_tune_layer { ZKG_TUNE_OVERRIDE = 0, ZKG_TUNE_LEARNED = 1 };
typedef struct zkg_plan_entry {
    int32_t field, log_height, width_class, log_blowup;
    int32_t plan;  /* 0 .. zkg_lde_plan_count() - 1 */
} zkg_plan_entry;
typedef struct zkg_split_entry {
    int32_t field, p2_width;
    int32_t merkle_lanes_log2, merkle_chunks_log2, hash_chunks_log2, fused_levels;
} zkg_split_entry;
int zkg_lde_plan_count(void);
const char *zkg_lde_plan_name(int plan);
int zkg_lde_plan_set(int layer, const zkg_plan_entry *entries, uint32_t n);
/* The plan a transform of this shape runs now; *source = 0 override,
 * 1 learned, 2 static table, 3 built-in default (source may be NULL). */
int zkg_lde_plan_for(int field, uint64_t height, uint32_t width, uint32_t log_blowup, int *plan, int *source);
/* Times every plan on this shape on the context's device (random input,
 * median of three) and records the fastest in the learned layer. */
int zkg_lde_plan_learn(zkg_ctx *ctx, int field, uint64_t height, uint32_t width, uint32_t log_blowup, int *plan);
int zkg_split_set(int layer, const zkg_split_entry *entries, uint32_t n);
/* The effective split values for (field, p2_width). */
int zkg_split_get(int field, int p2_width, zkg_split_entry *out);

/* ---- telemetry ------------------------------------------------------------- */
/* Per-stage device time of the most recent commit / FRI commit phase / batch,
 * from HIP events on the streams the kernels run on.  Commit: 0 LDE, 1 row
 * hash, 2 Merkle; FRI: 3 the whole phase; batch: 0 LDE (start to the last LDE
 * stream's end), 1 row hash and 2 Merkle (summed over Poseidon2 instances),
 * 3 the whole batch. */
int zkg_ctx_set_timing(zkg_ctx *ctx, int enable);
int zkg_ctx_stage_ms(const zkg_ctx *ctx, double *ms4);

#ifdef __cplusplus
}
#endif
#endif
