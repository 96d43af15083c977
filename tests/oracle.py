"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests, __graft_entry__.smoke() and bench.py's CPU
baseline leg use it, and only as the checker.  All values are uint64 numpy
arrays (BabyBear values in the low 31 bits).
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

GL, BB = 0, 1
GL_P = 0xFFFFFFFF00000001
BB_P = 0x78000001
PRIME = {GL: GL_P, BB: BB_P}

c_u64 = ctypes.c_uint64
c_u32 = ctypes.c_uint32
c_int = ctypes.c_int
c_vp = ctypes.c_void_p


class P2(ctypes.Structure):
    _fields_ = [
        ("field", c_int),
        ("width", c_int),
        ("rounds_f", c_int),
        ("rounds_p", c_int),
        ("diag", c_u64 * 16),
        ("rc", c_u64 * (8 * 16 + 32)),
    ]


class Tree(ctypes.Structure):
    _fields_ = [
        ("digest", c_int),
        ("nlayers", c_int),
        ("layer_len", ctypes.POINTER(c_u64)),
        ("layers", ctypes.POINTER(ctypes.POINTER(c_u64))),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        for name in ("or_gl_add", "or_gl_sub", "or_gl_mul", "or_gl_pow", "or_bb_mul", "or_bb_pow"):
            getattr(L, name).restype = c_u64
            getattr(L, name).argtypes = [c_u64, c_u64]
        for name in ("or_f_add", "or_f_sub", "or_f_mul", "or_f_pow"):
            getattr(L, name).restype = c_u64
            getattr(L, name).argtypes = [c_int, c_u64, c_u64]
        L.or_f_inv.restype = c_u64
        L.or_f_inv.argtypes = [c_int, c_u64]
        L.or_f_two_adic_generator.restype = c_u64
        L.or_f_two_adic_generator.argtypes = [c_int, c_u32]
        L.or_f_generator.restype = c_u64
        L.or_f_generator.argtypes = [c_int]
        L.or_ef4_mul.argtypes = [c_vp, c_vp, c_vp]
        L.or_gen_elem.restype = c_u64
        L.or_gen_elem.argtypes = [c_int, c_u64, c_u64]
        L.or_gen_fill.argtypes = [c_int, c_u64, c_u64, c_u64, c_vp]
        L.or_p2_init.restype = c_int
        L.or_p2_init.argtypes = [ctypes.POINTER(P2), c_int, c_int, c_u64, c_int]
        L.or_p2_permute.argtypes = [ctypes.POINTER(P2), c_vp]
        L.or_p2_hash.argtypes = [ctypes.POINTER(P2), c_vp, ctypes.c_size_t, c_vp]
        L.or_p2_compress.argtypes = [ctypes.POINTER(P2), c_vp, c_vp, c_vp]
        L.or_p2_flat_constants.restype = ctypes.c_size_t
        L.or_p2_flat_constants.argtypes = [ctypes.POINTER(P2), c_vp]
        L.or_dft_batch.argtypes = [c_int, c_vp, c_u64, c_u32, c_int]
        L.or_coset_lde_batch.argtypes = [c_int, c_vp, c_u64, c_u32, c_u32, c_u64, c_vp]
        L.or_bit_reverse_rows.argtypes = [c_vp, c_u64, c_u32]
        L.or_merkle_build.restype = c_int
        L.or_merkle_build.argtypes = [ctypes.POINTER(P2), c_int, c_vp, c_vp, c_vp, ctypes.POINTER(Tree)]
        L.or_merkle_free.argtypes = [ctypes.POINTER(Tree)]
        L.or_merkle_path.restype = c_int
        L.or_merkle_path.argtypes = [ctypes.POINTER(Tree), c_u64, c_vp]
        L.or_merkle_verify.restype = c_int
        L.or_merkle_verify.argtypes = [ctypes.POINTER(P2), c_vp, c_int, c_vp, c_vp, c_u64, c_vp, c_vp, c_int]
        L.or_commit.restype = c_int
        L.or_commit.argtypes = [ctypes.POINTER(P2), c_vp, c_u64, c_u32, c_u32, c_u64, c_vp, c_vp]
        L.or_fri_fold_base.argtypes = [c_int, c_vp, c_u64, c_u64, c_vp, c_vp]
        L.or_fri_fold_matrix.argtypes = [c_int, c_int, c_vp, c_u64, c_vp, c_vp]
        L.or_fri_commit_phase_base.restype = c_int
        L.or_fri_commit_phase_base.argtypes = [ctypes.POINTER(P2), c_vp, c_u32, c_u32, c_vp, c_vp, c_vp]
        L.or_fri_commit_phase_ef4.restype = c_int
        L.or_fri_commit_phase_ef4.argtypes = [ctypes.POINTER(P2), c_vp, c_u32, c_u32, c_vp, c_vp, c_vp, c_vp]
        L.or_gen_ef4_betas.argtypes = [c_u64, c_int, c_vp]
        L.or_ef4_inv.argtypes = [c_vp, c_vp]
        L.or_deep_open_values.argtypes = [c_vp, c_u64, c_u32, c_u32, c_u64, c_vp, c_vp]
        L.or_deep_reduce.argtypes = [c_vp, c_u64, c_u32, c_u64, c_vp, c_vp, c_u64, c_vp, c_vp]
        L.or_deep_gl_open_values.argtypes = [c_vp, c_u64, c_u32, c_u32, c_u64, c_u64, c_vp]
        L.or_deep_gl_reduce.argtypes = [c_vp, c_u64, c_u32, c_u64, c_u64, c_u64, c_u64, c_vp, c_vp]
        L.or_constraint_eval_gl.restype = c_int
        L.or_constraint_eval_gl.argtypes = [c_vp, c_u64, c_u32, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp]
        L.or_constraint_quotient_gl.restype = c_int
        L.or_constraint_quotient_gl.argtypes = [c_vp, c_u64, c_u32, c_u32, c_u64, c_vp, c_u32, c_vp, c_u32, c_u32,
                                                c_vp, c_vp]
        L.or_poly_eval_gl.argtypes = [c_vp, c_u64, c_vp, c_u64, c_vp]
        L.or_poly_coset_eval_gl.argtypes = [c_vp, c_u32, c_u64, c_vp]
        L.or_poly_quotient_gl.argtypes = [c_vp, c_u64, c_u64, c_u64, c_vp]
        L.or_set_threads.argtypes = [c_int]
        L.or_get_threads.restype = c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(c_vp)


def u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


# ---------------------------------------------------------------- fields
def mul(f, a, b):
    return lib().or_f_mul(f, a, b)


def add(f, a, b):
    return lib().or_f_add(f, a, b)


def sub(f, a, b):
    return lib().or_f_sub(f, a, b)


def pw(f, a, e):
    return lib().or_f_pow(f, a, e)


def inv(f, a):
    return lib().or_f_inv(f, a)


def two_adic_generator(f, log_n):
    return lib().or_f_two_adic_generator(f, log_n)


def ef4_mul(a, b):
    a, b = u64(a), u64(b)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_ef4_mul(_p(a), _p(b), _p(out))
    return out


def ef4_inv(a):
    a = u64(a)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_ef4_inv(_p(a), _p(out))
    return out


# ---------------------------------------------------------------- DEEP quotient
def deep_open_values(lde, log_blowup, shift, z):
    """p_k(z) of every column of a bit-reversed coset LDE (EF4, shape (width, 4))."""
    m = u64(lde)
    h, w = m.shape
    ys = np.zeros((w, 4), dtype=np.uint64)
    lib().or_deep_open_values(_p(m), h, w, log_blowup, shift, _p(u64(z)), _p(ys))
    return ys


def deep_reduce(lde, shift, z, alpha, alpha_off, ys, ro=None):
    """ro[i] += alpha^off * sum_k alpha^k (p_k(x_i) - ys_k) / (x_i - z); returns ro (height, 4)."""
    m = u64(lde)
    h, w = m.shape
    out = np.zeros((h, 4), dtype=np.uint64) if ro is None else u64(ro).copy()
    lib().or_deep_reduce(_p(m), h, w, shift, _p(u64(z)), _p(u64(alpha)), alpha_off, _p(u64(ys)), _p(out))
    return out


def deep_gl_open_values(lde, log_blowup, shift, z):
    """Goldilocks, base-field z: p_k(z) for every column (shape (width,))."""
    m = u64(lde)
    h, w = m.shape
    ys = np.zeros(w, dtype=np.uint64)
    lib().or_deep_gl_open_values(_p(m), h, w, log_blowup, shift, z, _p(ys))
    return ys


def deep_gl_reduce(lde, shift, z, alpha, alpha_off, ys, ro=None):
    m = u64(lde)
    h, w = m.shape
    out = np.zeros(h, dtype=np.uint64) if ro is None else u64(ro).copy()
    lib().or_deep_gl_reduce(_p(m), h, w, shift, z, alpha, alpha_off, _p(u64(ys)), _p(out))
    return out


# ---------------------------------------------------------------- constraint evaluation
OP_CONST, OP_LOAD, OP_ADD, OP_SUB, OP_MUL, OP_DUP, OP_EMIT = range(7)


def instr(op, operand=0):
    return (op << 24) | (operand & 0xFFFFFF)


def constraint_eval_gl(trace, bytecode, constants, n_constraints):
    """Goldilocks stack-machine evaluation of every row; None if the program is invalid."""
    t = u64(trace)
    rows, width = t.shape
    bc = np.ascontiguousarray(bytecode, dtype=np.uint32)
    cs = u64(constants) if len(constants) else np.zeros(1, dtype=np.uint64)
    out = np.zeros((rows, max(n_constraints, 1)), dtype=np.uint64)
    rc = lib().or_constraint_eval_gl(_p(t), rows, width, _p(bc), bc.size, _p(cs), len(constants), n_constraints,
                                     _p(out))
    return None if rc else out[:, :n_constraints]


def constraint_quotient_gl(lde, log_blowup, shift, bytecode, constants, n_constraints, coef):
    """q_i = sum_k coef[k] C_k(row i) / (x_i^N - 1) over a bit-reversed coset LDE; None if invalid."""
    t = u64(lde)
    rows, width = t.shape
    bc = np.ascontiguousarray(bytecode, dtype=np.uint32)
    cs = u64(constants) if len(constants) else np.zeros(1, dtype=np.uint64)
    cf = u64(coef)
    q = np.zeros(rows, dtype=np.uint64)
    rc = lib().or_constraint_quotient_gl(_p(t), rows, width, log_blowup, shift, _p(bc), bc.size, _p(cs),
                                         len(constants), n_constraints, _p(cf), _p(q))
    return None if rc else q


# ---------------------------------------------------------------- PolyOps (Goldilocks)
def poly_eval_gl(coeffs, points):
    c, x = u64(coeffs), u64(points)
    out = np.empty(x.size, dtype=np.uint64)
    lib().or_poly_eval_gl(_p(c) if c.size else None, c.size, _p(x), x.size, _p(out))
    return out


def poly_coset_eval_gl(coeffs, shift, log_n):
    c = u64(coeffs)
    assert c.size == 1 << log_n
    out = np.empty(c.size, dtype=np.uint64)
    lib().or_poly_coset_eval_gl(_p(c), log_n, shift, _p(out))
    return out


def poly_quotient_gl(evals, z, f_z):
    e = u64(evals)
    out = np.empty(e.size, dtype=np.uint64)
    lib().or_poly_quotient_gl(_p(e), e.size, z, f_z, _p(out))
    return out


# ---------------------------------------------------------------- generator
def gen(f, seed, n, start=0):
    out = np.empty(n, dtype=np.uint64)
    lib().or_gen_fill(f, seed, start, n, _p(out))
    return out


# ---------------------------------------------------------------- Poseidon2
class Poseidon2:
    def __init__(self, field, width=16, seed=42, bb_rc_mode=0):
        self.p = P2()
        rc = lib().or_p2_init(ctypes.byref(self.p), field, width, seed, bb_rc_mode)
        assert rc == 0, "unsupported Poseidon2 instance"
        self.field, self.width = field, width
        self.digest = width // 2

    def permute(self, state):
        s = u64(state).copy()
        assert s.size == self.width
        lib().or_p2_permute(ctypes.byref(self.p), _p(s))
        return s

    def hash(self, data):
        d = u64(data)
        out = np.zeros(self.digest, dtype=np.uint64)
        lib().or_p2_hash(ctypes.byref(self.p), _p(d) if d.size else None, d.size, _p(out))
        return out

    def compress(self, left, right):
        l, r = u64(left), u64(right)
        out = np.zeros(self.digest, dtype=np.uint64)
        lib().or_p2_compress(ctypes.byref(self.p), _p(l), _p(r), _p(out))
        return out

    def flat_constants(self):
        n = lib().or_p2_flat_constants(ctypes.byref(self.p), None)
        out = np.zeros(n, dtype=np.uint64)
        lib().or_p2_flat_constants(ctypes.byref(self.p), _p(out))
        return out

    @property
    def diag(self):
        return np.array(self.p.diag[: self.width], dtype=np.uint64)


# ---------------------------------------------------------------- DFT / LDE
def dft_batch(f, mat, inverse=False):
    m = u64(mat).copy()
    h, w = m.shape
    lib().or_dft_batch(f, _p(m), h, w, 1 if inverse else 0)
    return m


def coset_lde_batch(f, mat, log_blowup, shift):
    m = u64(mat)
    h, w = m.shape
    out = np.empty((h << log_blowup, w), dtype=np.uint64)
    lib().or_coset_lde_batch(f, _p(m), h, w, log_blowup, shift, _p(out))
    return out


# ---------------------------------------------------------------- Merkle
class MerkleTree:
    """p3 MerkleTree over one or more matrices (oracle)."""

    def __init__(self, p2, mats):
        self.p2 = p2
        self.mats = [u64(m) for m in mats]
        n = len(self.mats)
        ptrs = (c_vp * n)(*[_p(m) for m in self.mats])
        self.heights = np.array([m.shape[0] for m in self.mats], dtype=np.uint64)
        self.widths = np.array([m.shape[1] for m in self.mats], dtype=np.uint32)
        self.t = Tree()
        rc = lib().or_merkle_build(ctypes.byref(p2.p), n, ptrs, _p(self.heights), _p(self.widths), ctypes.byref(self.t))
        if rc != 0:
            raise ValueError("merkle build failed: %d" % rc)
        d = self.t.digest
        self.layers = []
        for i in range(self.t.nlayers):
            ln = self.t.layer_len[i]
            arr = np.ctypeslib.as_array(self.t.layers[i], shape=(ln * d,)).copy().reshape(ln, d)
            self.layers.append(arr)
        lib().or_merkle_free(ctypes.byref(self.t))

    def root(self):
        return self.layers[-1][0]

    def path(self, index):
        return np.array([self.layers[i][(index >> i) ^ 1] for i in range(len(self.layers) - 1)])

    def verify(self, root, index, rows, path):
        n = len(self.mats)
        rws = [u64(r) for r in rows]
        rptr = (c_vp * n)(*[_p(r) for r in rws])
        pa = u64(path).reshape(-1)
        return bool(lib().or_merkle_verify(ctypes.byref(self.p2.p), _p(u64(root)), n, _p(self.heights),
                                           _p(self.widths), index, rptr, _p(pa) if pa.size else None,
                                           len(path)))


def commit(p2, trace, log_blowup, shift=None):
    """root, lde (bit-reversed rows) of the full commit."""
    m = u64(trace)
    h, w = m.shape
    shift = lib().or_f_generator(p2.field) if shift is None else shift
    lde = np.empty((h << log_blowup, w), dtype=np.uint64)
    root = np.zeros(p2.digest, dtype=np.uint64)
    rc = lib().or_commit(ctypes.byref(p2.p), _p(m), h, w, log_blowup, shift, _p(lde), _p(root))
    assert rc == 0
    return root, lde


# ---------------------------------------------------------------- FRI
def fri_fold_base(f, evals, alpha, inv_twiddles):
    e, t = u64(evals), u64(inv_twiddles)
    n = e.size // 2
    out = np.empty(n, dtype=np.uint64)
    lib().or_fri_fold_base(f, _p(e), n, alpha, _p(t), _p(out))
    return out


def fri_fold_matrix_ef4(evals, beta):
    e = u64(evals).reshape(-1)
    n = e.size // 8
    out = np.empty(n * 4, dtype=np.uint64)
    lib().or_fri_fold_matrix(1, BB, _p(e), n, _p(u64(beta)), _p(out))
    return out.reshape(n, 4)


def fri_fold_matrix_base(f, evals, beta):
    e = u64(evals)
    n = e.size // 2
    out = np.empty(n, dtype=np.uint64)
    lib().or_fri_fold_matrix(0, f, _p(e), n, _p(u64([beta])), _p(out))
    return out


def fri_betas(seed, n):
    out = np.empty(4 * n, dtype=np.uint64)
    lib().or_gen_ef4_betas(seed, n, _p(out))
    return out.reshape(n, 4)


def fri_commit_phase_ef4(p2, codeword, log_len, log_final, betas):
    cw = u64(codeword).reshape(-1).copy()
    nl = log_len - log_final
    roots = np.empty(nl * 8, dtype=np.uint64)
    final = np.empty((1 << log_final) * 4, dtype=np.uint64)
    r = lib().or_fri_commit_phase_ef4(ctypes.byref(p2.p), _p(cw), log_len, log_final, _p(u64(betas).reshape(-1)),
                                      _p(roots), _p(final), None)
    assert r == nl, r
    return roots.reshape(nl, 8), final.reshape(-1, 4)


def fri_commit_phase_base(p2, codeword, log_len, log_final, betas):
    """Base-field commit phase: (layer roots (nfolds, digest), final codeword)."""
    cw = u64(codeword).reshape(-1).copy()
    nl = log_len - log_final
    roots = np.empty(nl * p2.digest, dtype=np.uint64)
    final = np.empty(1 << log_final, dtype=np.uint64)
    r = lib().or_fri_commit_phase_base(ctypes.byref(p2.p), _p(cw), log_len, log_final, _p(u64(betas)), _p(roots),
                                       _p(final))
    assert r == nl, r
    return roots.reshape(nl, p2.digest), final


def set_threads(n):
    lib().or_set_threads(n)


def get_threads():
    return lib().or_get_threads()


def verify(p2, root, heights, widths, index, rows, path):
    """p3 verify_batch of one opening (matrices given by heights/widths)."""
    n = len(heights)
    hs = np.ascontiguousarray(heights, dtype=np.uint64)
    ws = np.ascontiguousarray(widths, dtype=np.uint32)
    rws = [u64(r) for r in rows]
    rptr = (c_vp * n)(*[_p(r) for r in rws])
    pa = u64(path).reshape(-1)
    return bool(lib().or_merkle_verify(ctypes.byref(p2.p), _p(u64(root)), n, _p(hs), _p(ws), index, rptr,
                                       _p(pa) if pa.size else None, len(path)))
