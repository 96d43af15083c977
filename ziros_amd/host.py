"""Host-side mirror of the reference's accelerator interfaces over libzkg.

Names follow the reference so that call sites read the same:

* ``HipDft`` -- p3 ``TwoAdicSubgroupDft`` as MetalDft implements it
  (zkf-metal/src/ntt/p3_adapter.rs:35-459): ``dft_batch``, ``idft_batch``,
  ``coset_lde_batch`` (bit-reversed rows, as TwoAdicFriPcs commits them).
* ``HipPoseidon2`` -- ``Poseidon2Accelerator`` (zkf-core/src/acceleration.rs:
  106-140): ``batch_permute_goldilocks`` / ``batch_permute_babybear``.
* ``HipMerkleTreeMmcs`` -- ``Mmcs::{commit, open_batch}`` of MetalMerkleTreeMmcs
  (zkf-metal/src/merkle.rs:55-138).
* ``HipFri`` -- ``FriAccelerator::fold_*`` (zkf-core/src/acceleration.rs:196-217)
  plus the FRI commit phase over the BabyBear degree-4 extension.
* ``HipConstraintEval`` -- ``ConstraintEvalAccelerator::eval_trace_goldilocks``
  (zkf-core/src/acceleration.rs:219-242) with the reference's
  ``ConstraintCompiler`` bytecode builder; ``ConstraintProgram`` keeps a
  translated program resident for device traces.

Arrays are numpy (uint64 for Goldilocks, uint32 for BabyBear), row-major.
Everything runs on the GPU through libzkg; errors raise ``ZkgError``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import Job, ZkgError, check

GOLDILOCKS = 0
BABYBEAR = 1
GENERATOR = {GOLDILOCKS: 7, BABYBEAR: 31}
DTYPE = {GOLDILOCKS: np.uint64, BABYBEAR: np.uint32}


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One device context (zkg_ctx): resident twiddles and round constants."""

    def __init__(self, device=0, hbm_budget=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.zkg_ctx_create(device, hbm_budget, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.zkg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        check(self.lib.zkg_ctx_synchronize(self.h))

    def hbm_used(self):
        v = ctypes.c_uint64()
        check(self.lib.zkg_ctx_hbm_used(self.h, ctypes.byref(v)))
        return v.value

    def load_constants(self, field, width=16, seed=42):
        """Constants of Poseidon2<field, width> derived from SmallRng(seed)."""
        check(self.lib.zkg_constants_load(self.h, field, width, seed))

    def set_constants(self, field, width, flat, n_external_rounds=8, n_internal_rounds=None):
        """An explicit flat [initial][internal][terminal] table, as p3's
        flatten_round_constants hands it over (zkf-metal/src/poseidon2/
        babybear.rs:76-111, goldilocks.rs:42-83)."""
        n_int = n_internal_rounds if n_internal_rounds is not None else (22 if field == GOLDILOCKS else 13)
        a = np.ascontiguousarray(flat, dtype=DTYPE[field])
        check(self.lib.zkg_constants_set(self.h, field, width, _ptr(a), n_external_rounds, n_int))

    def set_timing(self, on=True):
        check(self.lib.zkg_ctx_set_timing(self.h, 1 if on else 0))

    def stage_ms(self):
        a = (ctypes.c_double * 4)()
        check(self.lib.zkg_ctx_stage_ms(self.h, a))
        return list(a)

    # buffers
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, max(arr.nbytes, 1))
        b.upload(arr)
        return b

    def random(self, field, n, seed, start=0):
        b = DeviceBuffer(self, max(n * np.dtype(DTYPE[field]).itemsize, 1))
        check(self.lib.zkg_fill_random(self.h, field, b.h, seed, start, n))
        return b


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = nbytes
        h = ctypes.c_void_p()
        check(ctx.lib.zkg_buf_alloc(ctx.h, nbytes, ctypes.byref(h)))
        self.h = h

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        check(self.ctx.lib.zkg_buf_upload(self.h, offset, _ptr(arr), arr.nbytes))

    def download(self, dtype, count, offset=0):
        out = np.empty(count, dtype=dtype)
        check(self.ctx.lib.zkg_buf_download(self.h, offset, _ptr(out), out.nbytes))
        return out

    def free(self):
        if self.h:
            self.ctx.lib.zkg_buf_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class HipDft:
    """p3 TwoAdicSubgroupDft over libzkg (MetalDft counterpart)."""

    def __init__(self, ctx, field):
        self.ctx, self.field = ctx, field
        self.dtype = DTYPE[field]

    def _run(self, mat, inverse):
        mat = np.ascontiguousarray(mat, dtype=self.dtype)
        h, w = mat.shape
        b = self.ctx.upload(mat)
        check(self.ctx.lib.zkg_ntt_batch(self.ctx.h, self.field, b.h, h, w, 1 if inverse else 0))
        out = b.download(self.dtype, h * w).reshape(h, w)
        b.free()
        return out

    def dft_batch(self, mat):
        return self._run(mat, False)

    def idft_batch(self, mat):
        return self._run(mat, True)

    def coset_lde_batch(self, mat, log_blowup, shift=None):
        """bit_reverse_rows(coset_lde_batch(mat, log_blowup, shift))"""
        mat = np.ascontiguousarray(mat, dtype=self.dtype)
        h, w = mat.shape
        shift = GENERATOR[self.field] if shift is None else shift
        bi = self.ctx.upload(mat)
        bo = self.ctx.alloc(max((h << log_blowup) * w * mat.itemsize, 1))
        check(self.ctx.lib.zkg_coset_lde_batch(self.ctx.h, self.field, bi.h, h, w, log_blowup, shift, bo.h))
        out = bo.download(self.dtype, (h << log_blowup) * w).reshape(h << log_blowup, w)
        bi.free()
        bo.free()
        return out


class HipPoseidon2:
    """Poseidon2Accelerator counterpart (width 16, explicit constants)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def name(self):
        return "hip-gfx950-poseidon2"

    def is_available(self):
        return True

    def batch_permute_goldilocks(self, states, round_constants, n_external_rounds, n_internal_rounds):
        s = np.ascontiguousarray(states, dtype=np.uint64)
        rc = np.ascontiguousarray(round_constants, dtype=np.uint64)
        check(self.ctx.lib.zkg_poseidon2_batch_permute_goldilocks(
            self.ctx.h, _ptr(s), s.size // 16, _ptr(rc), n_external_rounds, n_internal_rounds))
        return s

    def batch_permute_babybear(self, states, round_constants, n_external_rounds, n_internal_rounds):
        s = np.ascontiguousarray(states, dtype=np.uint32)
        rc = np.ascontiguousarray(round_constants, dtype=np.uint32)
        check(self.ctx.lib.zkg_poseidon2_batch_permute_babybear(
            self.ctx.h, _ptr(s), s.size // 16, _ptr(rc), n_external_rounds, n_internal_rounds))
        return s

    def permute(self, field, width, states):
        """Permute with the context's resident constants."""
        s = np.ascontiguousarray(states, dtype=DTYPE[field])
        b = self.ctx.upload(s)
        check(self.ctx.lib.zkg_poseidon2_permute(self.ctx.h, field, width, b.h, s.size // width))
        out = b.download(DTYPE[field], s.size).reshape(s.shape)
        b.free()
        return out


class HipPolyOps:
    """PolyOpsAccelerator (zkf-core/src/acceleration.rs:173-192) over libzkg:
    batch_eval_goldilocks, coset_eval_goldilocks, quotient_goldilocks."""

    def __init__(self, ctx):
        self.ctx = ctx

    def name(self):
        return "hip-gfx950-poly-ops"

    def is_available(self):
        return True

    def min_batch_size(self):
        return 1024

    def batch_eval_goldilocks(self, coeffs, points):
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        x = np.ascontiguousarray(points, dtype=np.uint64)
        out = np.empty(x.size, dtype=np.uint64)
        check(self.ctx.lib.zkg_poly_eval_goldilocks(self.ctx.h, _ptr(c) if c.size else None, c.size,
                                                    _ptr(x) if x.size else None, x.size,
                                                    _ptr(out) if out.size else None))
        return out

    def coset_eval_goldilocks(self, coeffs, shift, log_n):
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        if c.size != 1 << log_n:
            raise ValueError("coset_eval needs 2^log_n coefficients")
        out = np.empty(c.size, dtype=np.uint64)
        check(self.ctx.lib.zkg_poly_coset_eval_goldilocks(self.ctx.h, _ptr(c), log_n, int(shift), _ptr(out)))
        return out

    def quotient_goldilocks(self, evals, z, f_z):
        e = np.ascontiguousarray(evals, dtype=np.uint64)
        out = np.empty(e.size, dtype=np.uint64)
        check(self.ctx.lib.zkg_poly_quotient_goldilocks(self.ctx.h, _ptr(e), e.size, int(z), int(f_z), _ptr(out)))
        return out


class MerkleTree:
    """Prover data of one commitment (zkg_tree)."""

    def __init__(self, ctx, handle, field, keep=()):
        self.ctx, self.h, self.field = ctx, handle, field
        self._keep = keep  # buffers that back the opened rows
        n = ctypes.c_uint32()
        check(ctx.lib.zkg_tree_num_layers(handle, ctypes.byref(n)))
        self.num_layers = n.value
        check(ctx.lib.zkg_tree_digest_elems(handle, ctypes.byref(n)))
        self.digest = n.value

    def root(self):
        out = np.empty(self.digest, dtype=DTYPE[self.field])
        check(self.ctx.lib.zkg_tree_root(self.h, _ptr(out)))
        return out

    def layer_len(self, i):
        n = ctypes.c_uint64()
        check(self.ctx.lib.zkg_tree_layer_len(self.h, i, ctypes.byref(n)))
        return n.value

    def layer(self, i, n_leaves=None):
        n = self.layer_len(i)
        out = np.empty(n * self.digest, dtype=DTYPE[self.field])
        check(self.ctx.lib.zkg_tree_layer(self.h, i, _ptr(out), out.nbytes))
        return out.reshape(n, self.digest)

    def open_batch(self, indices, width):
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        n = idx.size
        rows = np.empty(n * width, dtype=DTYPE[self.field])
        paths = np.empty(n * (self.num_layers - 1) * self.digest, dtype=DTYPE[self.field])
        check(self.ctx.lib.zkg_tree_open(self.h, _ptr(idx), n, _ptr(rows), _ptr(paths)))
        return rows.reshape(n, width), paths.reshape(n, self.num_layers - 1, self.digest)

    def open_mmcs(self, indices, total_width):
        """Mmcs::open_batch for each index: the rows of every matrix (commit
        order, concatenated) and log2_ceil(max height) sibling digests."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        n = idx.size
        rows = np.empty(max(n * total_width, 1), dtype=DTYPE[self.field])
        paths = np.empty(max(n * (self.num_layers - 1) * self.digest, 1), dtype=DTYPE[self.field])
        check(self.ctx.lib.zkg_mmcs_open(self.h, _ptr(idx), n, _ptr(rows), _ptr(paths)))
        return (rows[: n * total_width].reshape(n, total_width),
                paths[: n * (self.num_layers - 1) * self.digest].reshape(n, self.num_layers - 1, self.digest))

    def free(self):
        if self.h:
            self.ctx.lib.zkg_tree_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class HipMerkleTreeMmcs:
    """Mmcs over PaddingFreeSponge / TruncatedPermutation of Poseidon2."""

    def __init__(self, ctx, field, p2_width=16):
        self.ctx, self.field, self.p2_width = ctx, field, p2_width

    def commit_matrix(self, mat):
        mat = np.ascontiguousarray(mat, dtype=DTYPE[self.field])
        h, w = mat.shape
        b = self.ctx.upload(mat)
        t = ctypes.c_void_p()
        check(self.ctx.lib.zkg_merkle_commit(self.ctx.h, self.field, self.p2_width, b.h, h, w, ctypes.byref(t)))
        tree = MerkleTree(self.ctx, t, self.field, keep=(b,))
        return tree.root(), tree

    def commit(self, mats):
        """Mmcs::commit over matrices of different heights (compress_and_inject)."""
        mats = [np.ascontiguousarray(m, dtype=DTYPE[self.field]) for m in mats]
        bufs = [self.ctx.upload(m) for m in mats]
        n = len(mats)
        handles = (ctypes.c_void_p * n)(*[b.h for b in bufs])
        heights = np.array([m.shape[0] for m in mats], dtype=np.uint64)
        widths = np.array([m.shape[1] for m in mats], dtype=np.uint32)
        t = ctypes.c_void_p()
        check(self.ctx.lib.zkg_mmcs_commit(self.ctx.h, self.field, self.p2_width, n, handles, _ptr(heights),
                                           _ptr(widths), ctypes.byref(t)))
        tree = MerkleTree(self.ctx, t, self.field, keep=tuple(bufs))
        return tree.root(), tree

    def open_batch(self, index, tree, width):
        rows, paths = tree.open_batch([index], width)
        return rows[0], paths[0]

    def verify_batch(self, root, heights, widths, indices, rows, paths):
        """Mmcs::verify_batch for many openings on the device: rows (n, sum of
        widths) and paths (n, log2_ceil(max height), digest) as open_mmcs
        returns them; a bool per opening."""
        dt = DTYPE[self.field]
        hs = np.ascontiguousarray(heights, dtype=np.uint64)
        ws = np.ascontiguousarray(widths, dtype=np.uint32)
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        r = np.ascontiguousarray(rows, dtype=dt)
        p = np.ascontiguousarray(paths, dtype=dt)
        rt = np.ascontiguousarray(root, dtype=dt)
        ok = np.zeros(idx.size, dtype=np.uint8)
        check(self.ctx.lib.zkg_mmcs_verify_batch(self.ctx.h, self.field, self.p2_width, _ptr(rt), hs.size, _ptr(hs),
                                                 _ptr(ws), idx.size, _ptr(idx), _ptr(r), _ptr(p), _ptr(ok)))
        return ok.astype(bool)


class HipDeep:
    """DEEP quotient of TwoAdicFriPcs::open over a device-resident LDE
    (bit-reversed rows on the coset shift * H).  Challenge values: BabyBear ->
    EF4, 4 canonical u32 (shape (..., 4)); Goldilocks -> the base field, one
    canonical u64."""

    def __init__(self, ctx, field, lde_buf, height, width, log_blowup, shift=None):
        self.ctx, self.field, self.lde = ctx, field, lde_buf
        self.height, self.width, self.log_blowup = height, width, log_blowup
        self.shift = GENERATOR[field] if shift is None else shift
        self.ext = (4,) if field == BABYBEAR else ()

    def _val(self, v):
        return np.ascontiguousarray(np.asarray(v).reshape(self.ext) if self.ext else np.asarray([v]).reshape(1),
                                    dtype=DTYPE[self.field])

    def open_values(self, z):
        z = self._val(z)
        ys = np.empty((self.width,) + self.ext, dtype=DTYPE[self.field])
        check(self.ctx.lib.zkg_deep_open_values(self.ctx.h, self.field, self.lde.h, self.height, self.width,
                                                self.log_blowup, self.shift, _ptr(z), _ptr(ys)))
        return ys

    def reduce(self, z, alpha, alpha_offset, ys, ro_buf):
        z, alpha = self._val(z), self._val(alpha)
        ys = np.ascontiguousarray(ys, dtype=DTYPE[self.field])
        check(self.ctx.lib.zkg_deep_reduce(self.ctx.h, self.field, self.lde.h, self.height, self.width, self.shift,
                                           _ptr(z), _ptr(alpha), alpha_offset, _ptr(ys), ro_buf.h))

    def open(self, z, alpha, alpha_offset, ro_buf):
        z, alpha = self._val(z), self._val(alpha)
        ys = np.empty((self.width,) + self.ext, dtype=DTYPE[self.field])
        check(self.ctx.lib.zkg_deep_open(self.ctx.h, self.field, self.lde.h, self.height, self.width, self.log_blowup,
                                         self.shift, _ptr(z), _ptr(alpha), alpha_offset, _ptr(ys), ro_buf.h))
        return ys


# ---------------------------------------------------------------- constraint evaluation
OP_CONST, OP_LOAD, OP_ADD, OP_SUB, OP_MUL, OP_DUP, OP_EMIT = range(7)


def encode_instruction(op, operand):
    """Opcode in the top 8 bits, operand in the low 24 (zkf-metal/src/constraint_eval.rs:24-27)."""
    return ((op & 0xFF) << 24) | (operand & 0x00FFFFFF)


class ConstraintCompiler:
    """Bytecode builder of zkf-metal/src/constraint_eval.rs:29-98: push_const,
    load_column, add, sub, mul, dup, emit; finish() -> (bytecode, constants,
    n_constraints)."""

    def __init__(self):
        self.bytecode, self.constants, self.n_constraints = [], [], 0

    def push_const(self, value):
        self.bytecode.append(encode_instruction(OP_CONST, len(self.constants)))
        self.constants.append(value)

    def load_column(self, col):
        self.bytecode.append(encode_instruction(OP_LOAD, col))

    def add(self):
        self.bytecode.append(encode_instruction(OP_ADD, 0))

    def sub(self):
        self.bytecode.append(encode_instruction(OP_SUB, 0))

    def mul(self):
        self.bytecode.append(encode_instruction(OP_MUL, 0))

    def dup(self):
        self.bytecode.append(encode_instruction(OP_DUP, 0))

    def emit(self, constraint_idx):
        self.bytecode.append(encode_instruction(OP_EMIT, constraint_idx))
        self.n_constraints = max(self.n_constraints, constraint_idx + 1)

    def finish(self):
        return self.bytecode, self.constants, self.n_constraints


_CE_INFO = ("words", "slots", "stages", "stack_slots", "wave_bytes", "panel", "rows_per_lane", "col_stride")


def translate_constraints(width, bytecode, n_constants, n_constraints, quotient=False):
    """The host translation libzkg runs before evaluation (or, quotient=True,
    before a quotient) with no device: (device words, info dict)."""
    lib = _lib.load()
    bc = np.ascontiguousarray(bytecode, dtype=np.uint32)
    info = np.zeros(8, dtype=np.uint32)
    q = 1 if quotient else 0
    rc = lib.zkg_constraint_translate(width, _ptr(bc), bc.size, n_constants, n_constraints, q, None, 0, _ptr(info))
    if rc != 0 and info[0] == 0:
        check(rc)
    words = np.zeros(int(info[0]), dtype=np.uint32)
    check(lib.zkg_constraint_translate(width, _ptr(bc), bc.size, n_constants, n_constraints, q, _ptr(words),
                                       words.size, _ptr(info)))
    return words, dict(zip(_CE_INFO, (int(v) for v in info)))


def compile_check_constraints(width, bytecode, constants, n_constraints):
    """Translate and compile a program on the host (no device); returns the
    compiler log, raises ZkgError if it does not compile."""
    lib = _lib.load()
    bc = np.ascontiguousarray(bytecode, dtype=np.uint32)
    cs = np.ascontiguousarray(constants, dtype=np.uint64) if len(constants) else np.zeros(1, np.uint64)
    log = ctypes.create_string_buffer(4096)
    check(lib.zkg_constraint_compile_check(width, _ptr(bc), bc.size, _ptr(cs), len(constants), n_constraints, log,
                                           len(log)))
    return log.value.decode()


def constraint_source(width, bytecode, constants, n_constraints):
    """The HIP source a compiled program is built from."""
    lib = _lib.load()
    bc = np.ascontiguousarray(bytecode, dtype=np.uint32)
    cs = np.ascontiguousarray(constants, dtype=np.uint64) if len(constants) else np.zeros(1, np.uint64)
    n = ctypes.c_uint64()
    check(lib.zkg_constraint_jit_source(width, _ptr(bc), bc.size, _ptr(cs), len(constants), n_constraints, None, 0,
                                        ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    check(lib.zkg_constraint_jit_source(width, _ptr(bc), bc.size, _ptr(cs), len(constants), n_constraints, buf,
                                        n.value, ctypes.byref(n)))
    return buf.value.decode()


class ConstraintProgram:
    """A checked, translated program resident on the device; evaluates any
    number of device traces of its width.  compiled=True also compiles it to
    a straight-line kernel (once, about a second) that later evals run."""

    def __init__(self, ctx, width, bytecode, constants, n_constraints, compiled=False):
        self.ctx, self.width, self.n_constraints = ctx, width, n_constraints
        bc = np.ascontiguousarray(bytecode, dtype=np.uint32)
        cs = np.ascontiguousarray(constants, dtype=np.uint64) if len(constants) else np.zeros(1, np.uint64)
        h = ctypes.c_void_p()
        check(ctx.lib.zkg_constraint_program_create(ctx.h, width, _ptr(bc), bc.size, _ptr(cs), len(constants),
                                                    n_constraints, ctypes.byref(h)))
        self.h = h
        if compiled:
            self.compile()

    def compile(self):
        check(self.ctx.lib.zkg_constraint_program_compile(self.h))

    def info(self):
        a = np.zeros(8, dtype=np.uint32)
        check(self.ctx.lib.zkg_constraint_program_info(self.h, _ptr(a)))
        return dict(zip(_CE_INFO, (int(v) for v in a)))

    def eval(self, trace_buf, n_rows, out_buf):
        check(self.ctx.lib.zkg_constraint_program_eval(self.h, trace_buf.h, n_rows, out_buf.h))

    def quotient(self, lde_buf, height, log_blowup, shift, coef, q_buf):
        """q[i] = sum_k coef[k] C_k(row i) / Z_H(x_i) over a committed bit-reversed coset LDE."""
        cf = np.ascontiguousarray(coef, dtype=np.uint64)
        if cf.size != self.n_constraints:
            raise ZkgError(-1, "%d coefficients for %d constraints" % (cf.size, self.n_constraints))
        check(self.ctx.lib.zkg_constraint_program_quotient(self.h, lde_buf.h, height, log_blowup, shift, _ptr(cf),
                                                           q_buf.h))

    def free(self):
        if self.h:
            self.ctx.lib.zkg_constraint_program_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class HipConstraintEval:
    """ConstraintEvalAccelerator counterpart (zkf-core/src/acceleration.rs:
    219-242, MetalConstraintEvalAccelerator of zkf-metal/src/registry.rs:
    213-237): host trace in, host constraint values out (n_rows x
    n_constraints, flat like the reference's Vec<u64>).  Refuses the inputs
    the reference's dispatch refuses (zkf-metal/src/constraint_eval.rs:
    123-130): width 0, an empty or ragged trace, fewer than min_rows() rows,
    empty bytecode, zero constraints."""

    MIN_ROWS = 1024

    def __init__(self, ctx):
        self.ctx = ctx

    def name(self):
        return "hip-gfx950-constraint-eval"

    def is_available(self):
        return True

    def min_rows(self):
        return self.MIN_ROWS

    def cache_config(self, max_programs=16, compile_after_uses=2):
        """Program memory of the context (zkg_constraint_cache_config)."""
        check(self.ctx.lib.zkg_constraint_cache_config(self.ctx.h, max_programs, compile_after_uses))

    def cache_info(self):
        a = np.zeros(3, dtype=np.uint32)
        check(self.ctx.lib.zkg_constraint_cache_info(self.ctx.h, _ptr(a)))
        return dict(programs=int(a[0]), compiled=int(a[1]), uses=int(a[2]))

    def eval_trace_goldilocks(self, trace, width, bytecode, constants, n_constraints):
        t = np.ascontiguousarray(trace, dtype=np.uint64).reshape(-1)
        if width == 0 or t.size == 0 or t.size % width:
            raise ZkgError(-1, "trace of %d elements is not a non-empty multiple of width %d" % (t.size, width))
        n_rows = t.size // width
        if n_rows < self.MIN_ROWS or len(bytecode) == 0 or n_constraints == 0:
            raise ZkgError(-1, "constraint eval dispatch refused (%d rows, %d instructions, %d constraints)"
                           % (n_rows, len(bytecode), n_constraints))
        tb = self.ctx.upload(t)
        ob = self.ctx.alloc(n_rows * n_constraints * 8)
        try:
            bc = np.ascontiguousarray(bytecode, dtype=np.uint32)
            cs = np.ascontiguousarray(constants, dtype=np.uint64) if len(constants) else np.zeros(1, np.uint64)
            check(self.ctx.lib.zkg_constraint_eval_gl(self.ctx.h, tb.h, n_rows, width, _ptr(bc), bc.size, _ptr(cs),
                                                      len(constants), n_constraints, ob.h))
            return ob.download(np.uint64, n_rows * n_constraints)
        finally:
            tb.free()
            ob.free()


# ---------------------------------------------------------------- runtime dispatch driver
OP_NTT, OP_LDE, OP_POSEIDON_BATCH, OP_MERKLE_LAYER, OP_FRI_FOLD = 1, 2, 3, 4, 5


class FallbackSignal(ZkgError):
    """The driver does not take the node; the runtime may run it on the CPU
    (zkf-runtime/src/metal_driver.rs:40-48)."""


class BufferBridge:
    """BufferBridge slots held in HBM (GpuBufferAllocator over zkg buffers)."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(ctx.lib.zkg_bridge_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def alloc_shared(self, slot, nbytes):
        check(self.ctx.lib.zkg_bridge_alloc(self.h, slot, nbytes))

    def free_shared(self, slot):
        check(self.ctx.lib.zkg_bridge_free(self.h, slot))

    def write(self, slot, arr):
        a = np.ascontiguousarray(arr)
        check(self.ctx.lib.zkg_bridge_write(self.h, slot, _ptr(a) if a.nbytes else None, a.nbytes))

    def read(self, slot, dtype):
        n = ctypes.c_uint64()
        check(self.ctx.lib.zkg_bridge_read(self.h, slot, None, 0, ctypes.byref(n)))
        out = np.empty(n.value // np.dtype(dtype).itemsize, dtype=dtype)
        check(self.ctx.lib.zkg_bridge_read(self.h, slot, _ptr(out) if out.nbytes else None, out.nbytes,
                                           ctypes.byref(n)))
        return out

    def destroy(self):
        if self.h:
            self.ctx.lib.zkg_bridge_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.destroy()
        except Exception:
            pass


class HipDispatchDriver:
    """GpuDispatchDriver (zkf-runtime/src/metal_driver.rs:55-93) over libzkg.
    execute(node, bridge) runs a commit-path ProverOp whose payload slots live
    in the bridge; it returns the telemetry, or raises FallbackSignal when
    this lane does not take the node."""

    def __init__(self, ctx):
        self.ctx = ctx

    def is_available(self):
        return True

    def should_try_on_either(self, **node):
        return bool(self.ctx.lib.zkg_driver_should_try(ctypes.byref(_lib.Node(**node))))

    def execute(self, bridge, **node):
        nd = _lib.Node(**node)
        tel = _lib.Telemetry()
        rc = self.ctx.lib.zkg_driver_execute(self.ctx.h, ctypes.byref(nd), bridge.h, ctypes.byref(tel))
        if rc == _lib.EFALLBACK:
            raise FallbackSignal(rc, (self.ctx.lib.zkg_last_error() or b"").decode())
        check(rc)
        return dict(kernel=tel.kernel.decode(), ms=tel.ms, input_bytes=tel.input_bytes,
                    output_bytes=tel.output_bytes)


# ---------------------------------------------------------------- tuning layers
TUNE_OVERRIDE, TUNE_LEARNED = 0, 1
PLAN_SOURCES = ("override", "learned", "static", "default")


def lde_plans():
    """Names of the transform tile plans, by index."""
    lib = _lib.load()
    return [lib.zkg_lde_plan_name(i).decode() for i in range(lib.zkg_lde_plan_count())]


def set_lde_plans(layer, entries):
    """Replace a tuning layer's plan entries: (field, log_height, width_class,
    log_blowup, plan) tuples, -1 matching anything; [] clears the layer."""
    arr = (_lib.PlanEntry * max(len(entries), 1))(*[_lib.PlanEntry(*e) for e in entries])
    check(_lib.load().zkg_lde_plan_set(layer, arr, len(entries)))


def lde_plan_for(field, height, width, log_blowup):
    """(plan, source) a transform of this shape runs now."""
    p, src = ctypes.c_int(), ctypes.c_int()
    check(_lib.load().zkg_lde_plan_for(field, height, width, log_blowup, ctypes.byref(p), ctypes.byref(src)))
    return p.value, PLAN_SOURCES[src.value]


def learn_lde_plan(ctx, field, height, width, log_blowup):
    """Time every plan on this shape and record the fastest (learned layer)."""
    p = ctypes.c_int()
    check(ctx.lib.zkg_lde_plan_learn(ctx.h, field, height, width, log_blowup, ctypes.byref(p)))
    return p.value


def set_splits(layer, entries):
    """Replace a tuning layer's split entries: (field, p2_width, merkle_lanes_log2,
    merkle_chunks_log2, hash_chunks_log2, fused_levels), -1 falling through."""
    arr = (_lib.SplitEntry * max(len(entries), 1))(*[_lib.SplitEntry(*e) for e in entries])
    check(_lib.load().zkg_split_set(layer, arr, len(entries)))


def splits(field, p2_width):
    e = _lib.SplitEntry()
    check(_lib.load().zkg_split_get(field, p2_width, ctypes.byref(e)))
    return dict(merkle_lanes_log2=e.merkle_lanes_log2, merkle_chunks_log2=e.merkle_chunks_log2,
                hash_chunks_log2=e.hash_chunks_log2, fused_levels=e.fused_levels)


def commit_trace(ctx, field, trace_buf, height, width, log_blowup, p2_width=16, plain=False):
    """Fused (or plain reference-structured) commit of a device-resident trace."""
    es = np.dtype(DTYPE[field]).itemsize
    lde = ctx.alloc(max((height << log_blowup) * width * es, 1))
    t = ctypes.c_void_p()
    fn = ctx.lib.zkg_commit_plain if plain else ctx.lib.zkg_commit
    check(fn(ctx.h, field, p2_width, trace_buf.h, height, width, log_blowup, lde.h, ctypes.byref(t)))
    return MerkleTree(ctx, t, field, keep=(lde,)), lde


def make_job(field, trace_buf, height, width, log_blowup, lde_buf, p2_width=16):
    return Job(field, p2_width, trace_buf.h, height, width, log_blowup, lde_buf.h)


def commit_batch(ctx, jobs):
    arr = (Job * len(jobs))(*jobs)
    outs = (ctypes.c_void_p * len(jobs))()
    check(ctx.lib.zkg_commit_batch(ctx.h, arr, len(jobs), outs))
    return [MerkleTree(ctx, ctypes.c_void_p(outs[i]), jobs[i].field) for i in range(len(jobs))]


class HipFri:
    """FriAccelerator counterpart plus the EF4 and base-field commit phases."""

    def __init__(self, ctx):
        self.ctx = ctx

    def name(self):
        return "hip-gfx950-fri"

    def is_available(self):
        return True

    def _fold(self, field, evals, alpha, inv_twiddles):
        dt = DTYPE[field]
        ev = np.ascontiguousarray(evals, dtype=dt)
        tw = np.ascontiguousarray(inv_twiddles, dtype=dt)
        n = ev.size // 2
        be, bt, bo = self.ctx.upload(ev), self.ctx.upload(tw), self.ctx.alloc(max(n * ev.itemsize, 1))
        check(self.ctx.lib.zkg_fri_fold(self.ctx.h, field, be.h, n, int(alpha), bt.h, bo.h))
        return bo.download(dt, n)

    def fold_goldilocks(self, evals, alpha, inv_twiddles):
        return self._fold(GOLDILOCKS, evals, alpha, inv_twiddles)

    def fold_babybear(self, evals, alpha, inv_twiddles):
        return self._fold(BABYBEAR, evals, alpha, inv_twiddles)

    def fold_matrix_ef4(self, evals, beta):
        """p3 fold_matrix, arity 2, over EF4: evals (2n, 4) uint32 -> (n, 4)."""
        ev = np.ascontiguousarray(evals, dtype=np.uint32)
        n = ev.shape[0] // 2
        be, bo = self.ctx.upload(ev), self.ctx.alloc(max(n * 16, 1))
        b = np.ascontiguousarray(beta, dtype=np.uint32)
        check(self.ctx.lib.zkg_fri_fold_matrix(self.ctx.h, 1, BABYBEAR, be.h, n, _ptr(b), bo.h))
        return bo.download(np.uint32, n * 4).reshape(n, 4)

    def fold_matrix_base(self, field, evals, beta):
        dt = DTYPE[field]
        ev = np.ascontiguousarray(evals, dtype=dt)
        n = ev.size // 2
        be, bo = self.ctx.upload(ev), self.ctx.alloc(max(n * ev.itemsize, 1))
        b = np.array([beta], dtype=np.uint64)
        check(self.ctx.lib.zkg_fri_fold_matrix(self.ctx.h, 0, field, be.h, n, _ptr(b), bo.h))
        return bo.download(dt, n)

    def commit_phase(self, codeword_buf, log_len, log_final, betas):
        """EF4 commit phase: codeword of 2^log_len EF4 values, betas (n, 4)."""
        b = np.ascontiguousarray(betas, dtype=np.uint32)
        h = ctypes.c_void_p()
        check(self.ctx.lib.zkg_fri_commit_phase(self.ctx.h, codeword_buf.h, log_len, log_final, _ptr(b),
                                                ctypes.byref(h)))
        return FriCommitment(self.ctx, h, log_len, log_final)

    def commit_phase_base(self, field, codeword_buf, log_len, log_final, betas):
        """Base-field commit phase (Challenge = Val): codeword of 2^log_len field
        elements, one canonical beta per fold."""
        b = np.ascontiguousarray(betas, dtype=np.uint64)
        h = ctypes.c_void_p()
        check(self.ctx.lib.zkg_fri_commit_phase_base(self.ctx.h, field, codeword_buf.h, log_len, log_final, _ptr(b),
                                                     ctypes.byref(h)))
        return FriCommitment(self.ctx, h, log_len, log_final, field=field)

    def begin(self, codeword_buf, log_len, log_final, field=None):
        """The commit phase one layer at a time, in transcript order: EF4
        (field None) or a base field.  Layer 0 is committed on return; each
        FriCommitment.fold_next(beta) folds and commits the next layer."""
        h = ctypes.c_void_p()
        check(self.ctx.lib.zkg_fri_begin(self.ctx.h, 1 if field is None else 0,
                                         BABYBEAR if field is None else field, codeword_buf.h, log_len,
                                         log_final, ctypes.byref(h)))
        return FriCommitment(self.ctx, h, log_len, log_final, field=field)


class FriCommitment:
    """Layers of a FRI commit phase: EF4 (field None) or a base field."""

    def __init__(self, ctx, h, log_len, log_final, field=None):
        self.ctx, self.h, self.log_len, self.log_final = ctx, h, log_len, log_final
        self.field = field
        # element dtype and words per element of the codeword, digest dtype
        self.dt = np.uint32 if field is None else DTYPE[field]
        self.per = 4 if field is None else 1
        n = ctypes.c_uint32()
        check(ctx.lib.zkg_fri_num_layers(h, ctypes.byref(n)))
        self.num_layers = n.value

    def layer_root(self, r):
        out = np.empty(8, dtype=self.dt)
        check(self.ctx.lib.zkg_fri_layer_root(self.h, r, _ptr(out)))
        return out

    def fold_next(self, beta):
        """Fold the newest layer with beta (EF4: 4 canonical words; base field:
        one canonical element) and commit the layer it produces."""
        if self.field is None:
            b = np.ascontiguousarray(np.asarray(beta, dtype=np.uint64).reshape(4), dtype=np.uint32)
        else:
            b = np.array([int(beta)], dtype=np.uint64)
        check(self.ctx.lib.zkg_fri_fold_next(self.h, _ptr(b)))

    def layer_roots(self):
        out = np.empty((self.num_layers, 8), dtype=self.dt)
        for r in range(self.num_layers):
            check(self.ctx.lib.zkg_fri_layer_root(self.h, r, _ptr(out[r])))
        return out

    def final(self):
        out = np.empty((1 << self.log_final) * self.per, dtype=self.dt)
        check(self.ctx.lib.zkg_fri_final(self.h, _ptr(out)))
        return out.reshape(-1, self.per) if self.field is None else out

    def open(self, indices):
        """(rows, paths): rows (n, layers, 2 * per) opened pairs, paths (n, sum of
        path lengths, 8) sibling digests."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        n = idx.size
        nl = self.num_layers
        plen = sum(self.log_len - 1 - r for r in range(nl))
        rows = np.empty(n * nl * 2 * self.per, dtype=self.dt)
        paths = np.empty(n * plen * 8, dtype=self.dt)
        check(self.ctx.lib.zkg_fri_open(self.h, _ptr(idx), n, _ptr(rows), _ptr(paths)))
        return rows.reshape(n, nl, 2 * self.per), paths.reshape(n, plen, 8)

    def free(self):
        if self.h:
            self.ctx.lib.zkg_fri_free(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.free()
        except Exception:
            pass


class Server:
    """Continuous-batching commit server (zkg_serve_*)."""

    def __init__(self, ctx, slots=2, max_batch=16):
        self.ctx = ctx
        check(ctx.lib.zkg_serve_start(ctx.h, slots, max_batch))

    def submit(self, job):
        t = ctypes.c_uint64()
        check(self.ctx.lib.zkg_serve_submit(self.ctx.h, ctypes.byref(job), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        """(tree handle, latency_ms) when done, None while pending."""
        tr = ctypes.c_void_p()
        lat = ctypes.c_double()
        rc = self.ctx.lib.zkg_serve_poll(self.ctx.h, ticket, ctypes.byref(tr), ctypes.byref(lat))
        if rc == _lib.EBUSY:
            return None
        check(rc)
        return tr, lat.value

    def stop(self):
        check(self.ctx.lib.zkg_serve_stop(self.ctx.h))
