"""ctypes binding of libzkg (include/zkg.h).

The library is built in-tree (``ziros_amd/libzkg.so``).  Loading fails loudly
when it is missing: there is no CPU fallback anywhere in this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# ZKG_LIBRARY selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("ZKG_LIBRARY") or os.path.join(_HERE, "libzkg.so")

c_u64 = ctypes.c_uint64
c_u32 = ctypes.c_uint32
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_dbl = ctypes.c_double
P = ctypes.POINTER


class Node(ctypes.Structure):
    """zkg_node"""

    _fields_ = [
        ("op", c_u32),
        ("field", ctypes.c_int32),
        ("size", c_u64),
        ("width", c_u32),
        ("blowup", c_u32),
        ("inverse", c_u32),
        ("n_ext", c_u32),
        ("n_int", c_u32),
        ("seed", c_u64),
        ("in_slot", c_u32),
        ("aux_slot", c_u32),
        ("aux2_slot", c_u32),
        ("out_slot", c_u32),
        ("out_bytes", c_u64),
    ]


class Telemetry(ctypes.Structure):
    """zkg_telemetry"""

    _fields_ = [("kernel", ctypes.c_char * 48), ("ms", c_dbl), ("input_bytes", c_u64), ("output_bytes", c_u64)]


class Job(ctypes.Structure):
    """zkg_job"""

    _fields_ = [
        ("field", c_int),
        ("p2_width", c_int),
        ("trace", c_vp),
        ("height", c_u64),
        ("width", c_u32),
        ("log_blowup", c_u32),
        ("lde_out", c_vp),
    ]


class PlanEntry(ctypes.Structure):
    """zkg_plan_entry"""

    _fields_ = [("field", ctypes.c_int32), ("log_height", ctypes.c_int32), ("width_class", ctypes.c_int32),
                ("log_blowup", ctypes.c_int32), ("plan", ctypes.c_int32)]


class SplitEntry(ctypes.Structure):
    """zkg_split_entry"""

    _fields_ = [("field", ctypes.c_int32), ("p2_width", ctypes.c_int32), ("merkle_lanes_log2", ctypes.c_int32),
                ("merkle_chunks_log2", ctypes.c_int32), ("hash_chunks_log2", ctypes.c_int32),
                ("fused_levels", ctypes.c_int32)]


# name -> (restype, argtypes); every entry point declared in include/zkg.h
SIGNATURES = {
    "zkg_last_error": (ctypes.c_char_p, []),
    "zkg_abi_version": (c_int, []),
    "zkg_ctx_create": (c_int, [c_int, c_u64, P(c_vp)]),
    "zkg_ctx_destroy": (c_int, [c_vp]),
    "zkg_ctx_synchronize": (c_int, [c_vp]),
    "zkg_ctx_hbm_used": (c_int, [c_vp, P(c_u64)]),
    "zkg_buf_alloc": (c_int, [c_vp, c_u64, P(c_vp)]),
    "zkg_buf_free": (c_int, [c_vp]),
    "zkg_buf_upload": (c_int, [c_vp, c_u64, c_vp, c_u64]),
    "zkg_buf_download": (c_int, [c_vp, c_u64, c_vp, c_u64]),
    "zkg_buf_size": (c_int, [c_vp, P(c_u64)]),
    "zkg_fill_random": (c_int, [c_vp, c_int, c_vp, c_u64, c_u64, c_u64]),
    "zkg_constants_load": (c_int, [c_vp, c_int, c_int, c_u64]),
    "zkg_constants_set": (c_int, [c_vp, c_int, c_int, c_vp, c_u32, c_u32]),
    "zkg_ntt_batch": (c_int, [c_vp, c_int, c_vp, c_u64, c_u32, c_int]),
    "zkg_ntt_batch_noalloc": (c_int, [c_vp, c_int, c_vp, c_u64, c_u32, c_int]),
    "zkg_lde_plan_count": (c_int, []),
    "zkg_lde_plan_name": (ctypes.c_char_p, [c_int]),
    "zkg_lde_plan_set": (c_int, [c_int, c_vp, c_u32]),
    "zkg_lde_plan_for": (c_int, [c_int, c_u64, c_u32, c_u32, P(c_int), P(c_int)]),
    "zkg_lde_plan_learn": (c_int, [c_vp, c_int, c_u64, c_u32, c_u32, P(c_int)]),
    "zkg_split_set": (c_int, [c_int, c_vp, c_u32]),
    "zkg_split_get": (c_int, [c_int, c_int, c_vp]),
    "zkg_poly_eval_goldilocks": (c_int, [c_vp, c_vp, c_u64, c_vp, c_u64, c_vp]),
    "zkg_poly_coset_eval_goldilocks": (c_int, [c_vp, c_vp, c_u32, c_u64, c_vp]),
    "zkg_poly_quotient_goldilocks": (c_int, [c_vp, c_vp, c_u64, c_u64, c_u64, c_vp]),
    "zkg_coset_lde_batch": (c_int, [c_vp, c_int, c_vp, c_u64, c_u32, c_u32, c_u64, c_vp]),
    "zkg_poseidon2_permute": (c_int, [c_vp, c_int, c_int, c_vp, c_u64]),
    "zkg_poseidon2_batch_permute_goldilocks": (c_int, [c_vp, c_vp, c_u64, c_vp, c_u32, c_u32]),
    "zkg_poseidon2_batch_permute_babybear": (c_int, [c_vp, c_vp, c_u64, c_vp, c_u32, c_u32]),
    "zkg_hash_rows": (c_int, [c_vp, c_int, c_int, c_vp, c_u64, c_u32, c_vp]),
    "zkg_merkle_commit": (c_int, [c_vp, c_int, c_int, c_vp, c_u64, c_u32, P(c_vp)]),
    "zkg_tree_root": (c_int, [c_vp, c_vp]),
    "zkg_tree_num_layers": (c_int, [c_vp, P(c_u32)]),
    "zkg_tree_digest_elems": (c_int, [c_vp, P(c_u32)]),
    "zkg_tree_layer": (c_int, [c_vp, c_u32, c_vp, c_u64]),
    "zkg_tree_open": (c_int, [c_vp, c_vp, c_u32, c_vp, c_vp]),
    "zkg_tree_layer_len": (c_int, [c_vp, c_u32, P(c_u64)]),
    "zkg_mmcs_commit": (c_int, [c_vp, c_int, c_int, c_u32, c_vp, c_vp, c_vp, c_vp]),
    "zkg_mmcs_open": (c_int, [c_vp, c_vp, c_u32, c_vp, c_vp]),
    "zkg_mmcs_verify_batch": (c_int, [c_vp, c_int, c_int, c_vp, c_u32, c_vp, c_vp, c_u32, c_vp, c_vp, c_vp, c_vp]),
    "zkg_deep_open_values": (c_int, [c_vp, c_int, c_vp, c_u64, c_u32, c_u32, c_u64, c_vp, c_vp]),
    "zkg_deep_reduce": (c_int, [c_vp, c_int, c_vp, c_u64, c_u32, c_u64, c_vp, c_vp, c_u64, c_vp, c_vp]),
    "zkg_deep_open": (c_int, [c_vp, c_int, c_vp, c_u64, c_u32, c_u32, c_u64, c_vp, c_vp, c_u64, c_vp, c_vp]),
    "zkg_constraint_program_create": (c_int, [c_vp, c_u32, c_vp, c_u32, c_vp, c_u32, c_u32, P(c_vp)]),
    "zkg_constraint_program_eval": (c_int, [c_vp, c_vp, c_u64, c_vp]),
    "zkg_constraint_program_info": (c_int, [c_vp, c_vp]),
    "zkg_constraint_program_compile": (c_int, [c_vp]),
    "zkg_constraint_program_free": (None, [c_vp]),
    "zkg_constraint_jit_source": (c_int, [c_u32, c_vp, c_u32, c_vp, c_u32, c_u32, ctypes.c_char_p, c_u64, P(c_u64)]),
    "zkg_constraint_compile_check": (c_int, [c_u32, c_vp, c_u32, c_vp, c_u32, c_u32, ctypes.c_char_p, c_u32]),
    "zkg_constraint_eval_gl": (c_int, [c_vp, c_vp, c_u64, c_u32, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp]),
    "zkg_constraint_cache_config": (c_int, [c_vp, c_u32, c_u32]),
    "zkg_constraint_cache_info": (c_int, [c_vp, c_vp]),
    "zkg_constraint_translate": (c_int, [c_u32, c_vp, c_u32, c_u32, c_u32, c_int, c_vp, c_u32, c_vp]),
    "zkg_constraint_program_quotient": (c_int, [c_vp, c_vp, c_u64, c_u32, c_u64, c_vp, c_vp]),
    "zkg_tree_free": (c_int, [c_vp]),
    "zkg_commit": (c_int, [c_vp, c_int, c_int, c_vp, c_u64, c_u32, c_u32, c_vp, P(c_vp)]),
    "zkg_commit_plain": (c_int, [c_vp, c_int, c_int, c_vp, c_u64, c_u32, c_u32, c_vp, P(c_vp)]),
    "zkg_fri_fold": (c_int, [c_vp, c_int, c_vp, c_u64, c_u64, c_vp, c_vp]),
    "zkg_fri_fold_matrix": (c_int, [c_vp, c_int, c_int, c_vp, c_u64, c_vp, c_vp]),
    "zkg_fri_commit_phase": (c_int, [c_vp, c_vp, c_u32, c_u32, c_vp, P(c_vp)]),
    "zkg_fri_commit_phase_base": (c_int, [c_vp, c_int, c_vp, c_u32, c_u32, c_vp, P(c_vp)]),
    "zkg_fri_begin": (c_int, [c_vp, c_int, c_int, c_vp, c_u32, c_u32, P(c_vp)]),
    "zkg_fri_fold_next": (c_int, [c_vp, c_vp]),
    "zkg_fri_num_layers": (c_int, [c_vp, P(c_u32)]),
    "zkg_fri_layer_root": (c_int, [c_vp, c_u32, c_vp]),
    "zkg_fri_final": (c_int, [c_vp, c_vp]),
    "zkg_fri_open": (c_int, [c_vp, c_vp, c_u32, c_vp, c_vp]),
    "zkg_fri_free": (c_int, [c_vp]),
    "zkg_commit_batch": (c_int, [c_vp, P(Job), c_u32, P(c_vp)]),
    "zkg_serve_start": (c_int, [c_vp, c_u32, c_u32]),
    "zkg_serve_submit": (c_int, [c_vp, P(Job), P(c_u64)]),
    "zkg_serve_poll": (c_int, [c_vp, c_u64, P(c_vp), P(c_dbl)]),
    "zkg_serve_stop": (c_int, [c_vp]),
    "zkg_bridge_create": (c_int, [c_vp, P(c_vp)]),
    "zkg_bridge_destroy": (None, [c_vp]),
    "zkg_bridge_alloc": (c_int, [c_vp, c_u32, c_u64]),
    "zkg_bridge_free": (c_int, [c_vp, c_u32]),
    "zkg_bridge_write": (c_int, [c_vp, c_u32, c_vp, c_u64]),
    "zkg_bridge_read": (c_int, [c_vp, c_u32, c_vp, c_u64, P(c_u64)]),
    "zkg_bridge_buffer": (c_vp, [c_vp, c_u32]),
    "zkg_driver_execute": (c_int, [c_vp, P(Node), c_vp, P(Telemetry)]),
    "zkg_driver_should_try": (c_int, [P(Node)]),
    "zkg_ctx_set_timing": (c_int, [c_vp, c_int]),
    "zkg_ctx_stage_ms": (c_int, [c_vp, P(c_dbl)]),
}

_lib = None


def load():
    """Load libzkg.so and bind every signature; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libzkg.so is not built (%s); run `python -m ziros_amd.build`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class ZkgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("zkg error %d: %s" % (code, msg))
        self.code = code


EBUSY = -5
EFALLBACK = -6


def check(rc):
    if rc != 0:
        raise ZkgError(rc, (load().zkg_last_error() or b"").decode())
    return rc
