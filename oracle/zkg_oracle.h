/*
 * zkg_oracle.h -- CPU restatement of ZirOS's STARK trace-commitment path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the bit-exactness checker for
 * the HIP implementation (libzkg) and the CPU baseline timed beside it.  The
 * product path never links, loads or calls it.
 *
 * Every function restates the reference algorithm it names (file:line under
 * the reference tree, or the Plonky3 0.4.2 / rand 0.9.2 semantics those files
 * call, which are not vendored in the reference).  Field elements of both
 * fields are passed as uint64_t; BabyBear values occupy the low 31 bits.
 *
 * Pins (checked by tests/test_oracle_*.py):
 *   - Goldilocks Poseidon2-16, RNG, constant order: golden vector of
 *     zkf-backends/src/wrapping/poseidon2_goldilocks.rs:546-587.
 *   - Field arithmetic: Python big-int brute force.
 *   - DFT / coset LDE: O(n^2) definition and closed form on small sizes.
 *   - Merkle: p3 layer layout (vendor/p3-merkle-tree-gpu/src/merkle_tree.rs)
 *     checked with the XOR compressor tests of that file, plus verify().
 *   - FRI fold: degree halving + even/odd identity.
 *   - BabyBear Poseidon2-16 round constants and Goldilocks Poseidon2-8 diag:
 *     parity unpinned (no in-tree vector; see the constant functions).
 */
#ifndef ZKG_ORACLE_H
#define ZKG_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_GOLDILOCKS = 0, OR_BABYBEAR = 1 };

#define OR_GL_P 0xFFFFFFFF00000001ULL
#define OR_BB_P 0x78000001u

/* ---- field arithmetic ---- */
uint64_t or_gl_add(uint64_t a, uint64_t b);
uint64_t or_gl_sub(uint64_t a, uint64_t b);
uint64_t or_gl_mul(uint64_t a, uint64_t b);
uint64_t or_gl_pow(uint64_t a, uint64_t e);
uint64_t or_bb_mul(uint64_t a, uint64_t b);
uint64_t or_bb_pow(uint64_t a, uint64_t e);
uint64_t or_f_add(int field, uint64_t a, uint64_t b);
uint64_t or_f_sub(int field, uint64_t a, uint64_t b);
uint64_t or_f_mul(int field, uint64_t a, uint64_t b);
uint64_t or_f_pow(int field, uint64_t a, uint64_t e);
uint64_t or_f_inv(int field, uint64_t a);
uint64_t or_f_two_adic_generator(int field, uint32_t log_n);
uint64_t or_f_generator(int field);
/* EF4 = BabyBear[x]/(x^4 - 11); elements are 4 coefficients, low first. */
void or_ef4_mul(const uint64_t *a, const uint64_t *b, uint64_t *out);

/* ---- rand 0.9.2 SmallRng (Xoshiro256++, SplitMix64 seeding) ---- */
typedef struct { uint64_t s[4]; } or_rng;
void or_rng_seed(or_rng *r, uint64_t seed);
uint64_t or_rng_next_u64(or_rng *r);
uint32_t or_rng_next_u32(or_rng *r);

/* ---- synthetic inputs (this build's own generator) ---- */
uint64_t or_gen_elem(int field, uint64_t seed, uint64_t index);
void or_gen_fill(int field, uint64_t seed, uint64_t start, uint64_t n, uint64_t *out);

/* ---- Poseidon2 ---- */
typedef struct {
    int field;       /* OR_GOLDILOCKS or OR_BABYBEAR */
    int width;       /* 8 or 16 */
    int rounds_f;    /* total external rounds */
    int rounds_p;    /* internal rounds */
    uint64_t diag[16];
    /* flat [initial rf/2*W][internal rp][terminal rf/2*W] */
    uint64_t rc[8 * 16 + 32];
} or_p2;

/* bb_rc_mode: 0 = Montgomery interpretation of the sampled u31 (p3
 * MontyField31 StandardUniform, the default), 1 = canonical interpretation. */
int or_p2_init(or_p2 *p, int field, int width, uint64_t seed, int bb_rc_mode);
void or_p2_permute(const or_p2 *p, uint64_t *state);
/* PaddingFreeSponge<P, W, rate=W/2, out=W/2> over `len` elements. */
void or_p2_hash(const or_p2 *p, const uint64_t *data, size_t len, uint64_t *digest);
/* TruncatedPermutation<P, 2, W/2, W>. */
void or_p2_compress(const or_p2 *p, const uint64_t *l, const uint64_t *r, uint64_t *out);
/* Flat round-constant export in the reference's GPU layout. */
size_t or_p2_flat_constants(const or_p2 *p, uint64_t *out);

/* ---- DFT / LDE (column-wise on row-major height x width) ---- */
void or_dft_batch(int field, uint64_t *mat, uint64_t height, uint32_t width, int inverse);
void or_coset_lde_batch(int field, const uint64_t *in, uint64_t height, uint32_t width,
                        uint32_t log_blowup, uint64_t shift, uint64_t *out);
void or_bit_reverse_rows(uint64_t *mat, uint64_t height, uint32_t width);

/* ---- Merkle tree (p3 MerkleTree::new layout) ---- */
typedef struct {
    int digest;          /* elements per digest */
    int nlayers;
    uint64_t *layer_len; /* nlayers */
    uint64_t **layers;   /* layer i: layer_len[i] * digest elements */
} or_tree;

/* Matrices given tallest-first or in any order; heights must satisfy the p3
 * rule (equal when they round up to the same power of two). */
int or_merkle_build(const or_p2 *p, int nmats, const uint64_t *const *mats,
                    const uint64_t *heights, const uint32_t *widths, or_tree *out);
void or_merkle_free(or_tree *t);
void or_merkle_root(const or_tree *t, uint64_t *root);
/* Sibling path for `index`: log2_ceil(max_height) digests. */
int or_merkle_path(const or_tree *t, uint64_t index, uint64_t *path);
/* Verify one opening of single or multiple matrices (p3 verify_batch). */
int or_merkle_verify(const or_p2 *p, const uint64_t *root, int nmats, const uint64_t *heights,
                     const uint32_t *widths, uint64_t index, const uint64_t *const *rows,
                     const uint64_t *path, int path_len);
/* Convenience for ctypes: build a single-matrix tree and copy the layers. */
int or_merkle_single(const or_p2 *p, const uint64_t *mat, uint64_t height, uint32_t width,
                     uint64_t *root, uint64_t *layers_flat, uint64_t layers_cap);

/* Full commit: coset LDE (bit-reversed rows) + row hash + tree.  Returns the
 * root; `lde_out` (height<<log_blowup rows) may be NULL. */
int or_commit(const or_p2 *p, const uint64_t *trace, uint64_t height, uint32_t width,
              uint32_t log_blowup, uint64_t shift, uint64_t *lde_out, uint64_t *root);

/* ---- FRI ---- */
/* Reference FriAccelerator::fold_* (base field, explicit inverse twiddles). */
void or_fri_fold_base(int field, const uint64_t *evals, uint64_t n_out, uint64_t alpha,
                      const uint64_t *inv_twiddles, uint64_t *out);
/* p3-fri TwoAdicFriFolding::fold_matrix for arity 2 over EF4 (ext=1) or the
 * base field (ext=0): input 2*n_out elements in bit-reversed order. */
void or_fri_fold_matrix(int ext, int field, const uint64_t *evals, uint64_t n_out,
                        const uint64_t *beta, uint64_t *out);
/* Commit phase over EF4: layers of length 2^log_len ... 2^(log_final+1) are
 * committed as width-2 matrices (8 base elements per leaf), folded with
 * betas[r] (4 coefficients each).  layer_roots: nfolds * 8; final_out:
 * 2^log_final * 4. Codeword is consumed (overwritten). */
int or_fri_commit_phase_ef4(const or_p2 *p, uint64_t *codeword, uint32_t log_len,
                            uint32_t log_final, const uint64_t *betas, uint64_t *layer_roots,
                            uint64_t *final_out, or_tree *trees_out);
void or_gen_ef4_betas(uint64_t seed, int n, uint64_t *betas);
/* Commit phase over the base field of p (Poseidon2 width 16): layers committed
 * as width-2 matrices, folded with betas[r] (one element each).  layer_roots:
 * nfolds * digest; final_out: 2^log_final.  Codeword is consumed. */
int or_fri_commit_phase_base(const or_p2 *p, uint64_t *codeword, uint32_t log_len, uint32_t log_final,
                             const uint64_t *betas, uint64_t *layer_roots, uint64_t *final_out);

/* ---- PCS opening: DEEP quotient over BabyBear with EF4 challenges ----
 * Restates p3-fri 0.4.2 TwoAdicFriPcs::open (external to the reference; the
 * reference's node is ProverOp::FriQueryOpen, zkf-runtime/src/graph.rs:105-108)
 * and p3-interpolation 0.4.2 interpolate_coset.  An LDE of height H = B*N holds
 * rows in bit-reversed order on the coset g*H_H: row i is p(x_i),
 * x_i = g * w_H^rev(i).  EF4 values are 4 canonical coefficients. */
void or_ef4_inv(const uint64_t *a, uint64_t *out);
/* p_k(z) for each column k: barycentric interpolation over the first N rows
 * (p on g*H_N): (z^N - g^N) / (N g^N) * sum_j e_j x_j / (z - x_j). */
void or_deep_open_values(const uint64_t *lde, uint64_t height, uint32_t width, uint32_t log_blowup,
                         uint64_t shift, const uint64_t *z, uint64_t *ys);
/* ro[i] += alpha^off * sum_k alpha^k (p_k(x_i) - ys_k) / (x_i - z), for i < height. */
void or_deep_reduce(const uint64_t *lde, uint64_t height, uint32_t width, uint64_t shift, const uint64_t *z,
                    const uint64_t *alpha, uint64_t alpha_off, const uint64_t *ys, uint64_t *ro);

/* Goldilocks with base-field challenges (the reference's Challenge<F> = F,
 * zkf-backends/src/plonky3.rs:55): the same two steps, z, alpha, ys and ro in
 * Goldilocks. */
void or_deep_gl_open_values(const uint64_t *lde, uint64_t height, uint32_t width, uint32_t log_blowup,
                            uint64_t shift, uint64_t z, uint64_t *ys);
void or_deep_gl_reduce(const uint64_t *lde, uint64_t height, uint32_t width, uint64_t shift, uint64_t z,
                       uint64_t alpha, uint64_t alpha_off, const uint64_t *ys, uint64_t *ro);

/* ---- Constraint evaluation (ConstraintEvalAccelerator::eval_trace_goldilocks,
 * zkf-core/src/acceleration.rs:219-242; interpreter semantics of
 * zkf-metal/src/shaders/constraint_eval.metal:1-74) ----
 * Each instruction word is opcode << 24 | operand (24 bits).  Opcodes:
 * 0 CONST (push constants[operand]), 1 LOAD (push trace[row][operand]),
 * 2 ADD, 3 SUB, 4 MUL (pop b, pop a, push a op b), 5 DUP, 6 EMIT (pop into
 * out[row][operand]); any other opcode does nothing.  Every row runs the
 * whole program; out (n_rows x n_constraints) is zeroed first.  Returns -1 if
 * the program underflows or overflows the 32-entry stack or indexes out of
 * range (the reference would read or write out of bounds). */
int or_constraint_eval_gl(const uint64_t *trace, uint64_t n_rows, uint32_t width, const uint32_t *bytecode,
                          uint32_t n_instr, const uint64_t *constants, uint32_t n_constants,
                          uint32_t n_constraints, uint64_t *out);
/* Quotient values over a committed Goldilocks LDE (height H rows, row i holding
 * the trace polynomials at x_i = shift * w_H^rev(i), as or_coset_lde_batch
 * leaves them): q_i = (sum_k coef[k] C_k(row i)) / (x_i^N - 1), N = H >> log_blowup,
 * with C_k the constraint values above.  The division by the vanishing
 * polynomial of the trace domain is the quotient step of a STARK prover.
 * Returns -1 for an invalid program or a vanishing value of zero. */
int or_constraint_quotient_gl(const uint64_t *lde, uint64_t height, uint32_t width, uint32_t log_blowup,
                              uint64_t shift, const uint32_t *bytecode, uint32_t n_instr, const uint64_t *constants,
                              uint32_t n_constants, uint32_t n_constraints, const uint64_t *coef, uint64_t *q);

/* ---- PolyOpsAccelerator, Goldilocks (zkf-core/src/acceleration.rs:173-192;
 * Metal semantics zkf-metal/src/poly.rs:37-246, shaders/poly_ops.metal) ----
 * eval: out[i] = sum_k coeffs[k] points[i]^k by Horner (poly_eval_goldilocks).
 * coset_eval: coeffs[k] *= shift^k, then the forward DFT in natural order
 * (n = 2^log_n coefficients; MetalPolyOps::coset_eval_goldilocks).
 * quotient: out[i] = (evals[i] - f_z) / (g^i - z), g = two_adic_generator
 * (log2 n), 0 where the denominator vanishes (poly_quotient_goldilocks). */
void or_poly_eval_gl(const uint64_t *coeffs, uint64_t n_coeffs, const uint64_t *points, uint64_t n_points,
                     uint64_t *out);
void or_poly_coset_eval_gl(const uint64_t *coeffs, uint32_t log_n, uint64_t shift, uint64_t *out);
void or_poly_quotient_gl(const uint64_t *evals, uint64_t n, uint64_t z, uint64_t f_z, uint64_t *out);

/* ---- threading ---- */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
