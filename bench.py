"""Benchmark of the STARK trace-commitment path on one MI355X.

Headline workload: the batched commit.  48 BabyBear traces (a seeded draw of
heights 2^U{12..20} and widths U{8, 16, ..., 256}, 1.59e9 cells) each get a
coset LDE at blowup 2, Poseidon2-16 row hashes and a Merkle tree, in one
zkg_commit_batch call (a replayed hipGraph: LDEs on parallel streams, one
persistent row-hash launch over all traces, one Merkle launch per round of
levels).  A step is one such batch with the traces already in HBM; value =
trace cells (rows x columns before the LDE) committed per second.

Also reported on the same line:
  batched_vs_plain  one plain call per trace (reference-structured: one
                global pass per NTT stage, per-level Merkle launches, a
                synchronisation after every launch) on the same traces, and
                batched_vs_single_fused against one fused zkg_commit per trace;
  roofline      the dominant kernel, the batch's row-hash launch (VALU bound):
                profiled VALU wave-instructions per launch over its live
                HIP-event duration against the issue ceiling, and 772 BabyBear
                products per permutation against the measured product ceiling;
  c2            one BabyBear 2^20 x 64 commit, with the row-hash and LDE
                rooflines (the LDE against 8 TB/s);
  c1, c4_fri, pcs_open_deep, mmcs, constraint_eval  the other configurations
                and the opening-side kernels; gl_fri the Goldilocks (base-field)
                FRI commit phase;
  served        continuous-batching stream (60% BabyBear / 40% Goldilocks,
                heights 2^U{14..22}, widths U{16..128}, blowup 2 or 4,
                Poisson arrivals): p50/p99 latency and cells/s vs one plain
                call per request;
  cpu_baseline  the CPU oracle (C restatement of the reference CPU path) on a
                bounded sample of the batched draw, on this GPU's share of the
                host cores and on one core.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import ziros_amd as Z  # noqa: E402
from ziros_amd import workloads  # noqa: E402
from ziros_amd._lib import check  # noqa: E402

HBM_PEAK = 8.0e12
BB_PERM_MULS = 772
GL16_PERM_MULS = 952
SEED_C1 = 0x5EED0001
SEED_C2 = 0x5EED0002
SEED_C3 = workloads.SEED_BATCH
SEED_C5 = workloads.SEED_SERVE


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def probe_peak(kind=0, iters=4096):
    lib = ctypes.CDLL(os.path.join(ROOT, "ziros_amd", "libzkg_probe.so"))
    lib.zkg_probe_rate.restype = ctypes.c_int
    lib.zkg_probe_rate.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    v = ctypes.c_double()
    rc = lib.zkg_probe_rate(kind, iters, ctypes.byref(v))
    if rc:
        raise RuntimeError("probe failed: %d" % rc)
    return v.value


def probe_copy(nbytes=1 << 30):
    lib = ctypes.CDLL(os.path.join(ROOT, "ziros_amd", "libzkg_probe.so"))
    lib.zkg_probe_copy.restype = ctypes.c_int
    lib.zkg_probe_copy.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    v = ctypes.c_double()
    rc = lib.zkg_probe_copy(nbytes, ctypes.byref(v))
    if rc:
        raise RuntimeError("copy probe failed: %d" % rc)
    return v.value


# VALU issue ceiling: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles each, 2.4 GHz peak engine clock.
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4


def valu_issue(insts, ms):
    if not insts or not ms:
        return None
    rate = insts / (ms * 1e-3)
    return {"achieved": rate / 1e9, "peak": VALU_ISSUE_PEAK / 1e9, "unit": "G wave-instr/s",
            "frac": rate / VALU_ISSUE_PEAK, "insts_per_launch": insts}


BATCH_COUNTERS = "batch_bb_48traces_counters.md"
C2_COUNTERS = "commit_bb_2p20x64_traffic.md"


def profiled_counters(name):
    """Per-launch figures measured with rocprofv3 and committed under profiles/:
    HBM bytes (FETCH_SIZE / WRITE_SIZE, calibrated against a 1 GiB copy) and VALU
    wave-instructions (SQ_INSTS_VALU), as `key: integer` lines."""
    out = {}
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            k, _, v = line.partition(":")
            if k.startswith(("traffic_", "valu_")) and v.strip().isdigit():
                out[k] = int(v)
    except OSError:
        pass
    return out


def free_tree(ctx, h):
    ctx.lib.zkg_tree_free(h)


def commit_once(ctx, fn, field, p2w, trace, n, w, lb, lde):
    t = ctypes.c_void_p()
    check(fn(ctx.h, field, p2w, trace.h, n, w, lb, lde.h, ctypes.byref(t)))
    return t


def bench_c2(ctx, steps, warmup):
    n, w, lb = 1 << 20, 64, 1
    trace = ctx.random(Z.BABYBEAR, n * w, SEED_C2)
    lde = ctx.alloc((n << lb) * w * 4)
    ctx.set_timing(True)
    for _ in range(warmup):
        free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit, Z.BABYBEAR, 16, trace, n, w, lb, lde))
    ctx.synchronize()
    stages = np.zeros(4)
    t0 = time.perf_counter()
    for _ in range(steps):
        t = commit_once(ctx, ctx.lib.zkg_commit, Z.BABYBEAR, 16, trace, n, w, lb, lde)
        stages += ctx.stage_ms()
        free_tree(ctx, t)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    stages /= steps
    # plain (reference-structured) path
    t0 = time.perf_counter()
    plain_steps = max(1, min(steps, 3))
    for _ in range(plain_steps):
        free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit_plain, Z.BABYBEAR, 16, trace, n, w, lb, lde))
    ctx.synchronize()
    dt_plain = (time.perf_counter() - t0) / plain_steps
    ctx.set_timing(False)
    trace.free()
    lde.free()
    return dict(n=n, w=w, lb=lb, sec_per_step=dt / steps, stage_ms=stages.tolist(), plain_sec=dt_plain)


def bench_c1(ctx, steps):
    """Config 1: Goldilocks 2^14 x 8, blowup 2, Poseidon2 width 8 (rate 4,
    digest 4) as configured, and width 16 (rate 8, digest 8) as the reference
    STARK uses.  Small: latency of one synchronous commit."""
    n, w, lb = 1 << 14, 8, 1
    trace = ctx.random(Z.GOLDILOCKS, n * w, SEED_C1)
    lde = ctx.alloc((n << lb) * w * 8)
    out = {}
    for p2w in (8, 16):
        for _ in range(3):
            free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit, Z.GOLDILOCKS, p2w, trace, n, w, lb, lde))
        t0 = time.perf_counter()
        for _ in range(steps):
            free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit, Z.GOLDILOCKS, p2w, trace, n, w, lb, lde))
        dt = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        for _ in range(3):
            free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit_plain, Z.GOLDILOCKS, p2w, trace, n, w, lb, lde))
        dtp = (time.perf_counter() - t0) / 3
        out["poseidon2_w%d" % p2w] = dict(ms_per_commit=dt * 1e3, cells_per_s=n * w / dt, plain_ms=dtp * 1e3)
    trace.free()
    lde.free()
    return out


def bench_gl_row_hash(ctx, steps):
    """The Goldilocks row hash of one fused commit: 2^20 x 97 trace, blowup 2,
    so 2^21 LDE rows of 13 permutations each, Poseidon2 width 16 and width 8
    (the served mix's Goldilocks traces run the same kernel).  Device time of
    the row-hash stage from the library's HIP events, median over `steps`."""
    n, w, lb = 1 << 20, 97, 1
    trace = ctx.random(Z.GOLDILOCKS, n * w, 0x5EED0097)
    lde = ctx.alloc((n << lb) * w * 8)
    out = {}
    ctx.set_timing(True)
    try:
        for p2w in (16, 8):
            rate = 8 if p2w == 16 else 4
            perms = (n << lb) * -(-w // rate)
            ms = []
            for i in range(steps + 1):
                free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit, Z.GOLDILOCKS, p2w, trace, n, w, lb, lde))
                if i:
                    ms.append(ctx.stage_ms()[1])
            med = sorted(ms)[len(ms) // 2]
            out["poseidon2_w%d" % p2w] = dict(row_hash_ms=med, permutations=perms, permutations_per_s=perms / (med * 1e-3))
    finally:
        ctx.set_timing(False)
        trace.free()
        lde.free()
    return out


def bench_c4(ctx, steps):
    """Config 4: FRI commit phase over EF4, codeword 2^22 (a 2^20-coefficient
    polynomial at blowup 4) folded to 2^8, one Poseidon2-16 tree per layer."""
    import numpy as np  # noqa: F811

    log_len, log_final = 22, 8
    coeff = ctx.random(Z.BABYBEAR, (1 << 20) * 4, 0x5EED0004)
    cw = ctx.alloc((1 << log_len) * 16)
    check(ctx.lib.zkg_coset_lde_batch(ctx.h, Z.BABYBEAR, coeff.h, 1 << 20, 4, 2, 31, cw.h))
    rng = np.random.default_rng(0x5EED0004)
    betas = rng.integers(0, 0x78000001, (log_len - log_final) * 4).astype(np.uint32)
    fri = Z.HipFri(ctx)
    ctx.set_timing(True)
    for _ in range(2):
        fri.commit_phase(cw, log_len, log_final, betas).free()
    t0 = time.perf_counter()
    dev = 0.0
    for _ in range(steps):
        f = fri.commit_phase(cw, log_len, log_final, betas)
        dev += ctx.stage_ms()[3]
        f.free()
    dt = (time.perf_counter() - t0) / steps
    ctx.set_timing(False)
    # transcript order: each layer's root read back before its fold (zkg_fri_begin / zkg_fri_fold_next)
    t0 = time.perf_counter()
    for _ in range(steps):
        f = fri.begin(cw, log_len, log_final)
        for r in range(log_len - log_final):
            f.layer_root(r)
            f.fold_next(betas[4 * r:4 * r + 4])
        f.final()
        f.free()
    dt_step = (time.perf_counter() - t0) / steps
    opening = fri_open_times(fri.commit_phase(cw, log_len, log_final, betas), log_len)
    coeff.free()
    cw.free()
    return dict(codeword=1 << log_len, layers=log_len - log_final, ms_per_commit_phase=dt * 1e3,
                device_ms=dev / steps, codeword_elems_per_s=(1 << log_len) / dt,
                ms_per_commit_phase_transcript_order=dt_step * 1e3, query_openings=opening)


def fri_open_times(f, log_len, n=80):
    """zkg_fri_open of n seeded query indices (rows and sibling paths of every
    layer), all in one call and one call per query."""
    q = np.random.default_rng(0x5EED0080).integers(0, 1 << log_len, n).astype(np.uint64)
    f.open(q)
    t0 = time.perf_counter()
    f.open(q)
    t_all = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k in range(n):
        f.open(q[k:k + 1])
    t_one = time.perf_counter() - t0
    f.free()
    return dict(queries=n, ms_all_in_one_call=t_all * 1e3, ms_one_call_per_query=t_one * 1e3)


def bench_fri_gl(ctx, steps):
    """The Goldilocks STARK's FRI commit phase (Challenge = Val): a codeword of
    2^22 base-field values (a 2^20-coefficient polynomial at blowup 4, one
    column) folded to 2^8, one Poseidon2-16 tree per layer."""
    log_len, log_final = 22, 8
    coeff = ctx.random(Z.GOLDILOCKS, 1 << 20, 0x5EED0044)
    cw = ctx.alloc((1 << log_len) * 8)
    check(ctx.lib.zkg_coset_lde_batch(ctx.h, Z.GOLDILOCKS, coeff.h, 1 << 20, 1, 2, 7, cw.h))
    rng = np.random.default_rng(0x5EED0044)
    betas = rng.integers(0, 0xFFFFFFFF00000001, log_len - log_final, dtype=np.uint64)
    fri = Z.HipFri(ctx)
    ctx.set_timing(True)
    for _ in range(2):
        fri.commit_phase_base(Z.GOLDILOCKS, cw, log_len, log_final, betas).free()
    t0 = time.perf_counter()
    dev = 0.0
    for _ in range(steps):
        f = fri.commit_phase_base(Z.GOLDILOCKS, cw, log_len, log_final, betas)
        dev += ctx.stage_ms()[3]
        f.free()
    dt = (time.perf_counter() - t0) / steps
    ctx.set_timing(False)
    # transcript order: each layer's root read back before its fold (zkg_fri_begin / zkg_fri_fold_next)
    t0 = time.perf_counter()
    for _ in range(steps):
        f = fri.begin(cw, log_len, log_final, field=Z.GOLDILOCKS)
        for r in range(log_len - log_final):
            f.layer_root(r)
            f.fold_next(int(betas[r]))
        f.final()
        f.free()
    dt_step = (time.perf_counter() - t0) / steps
    opening = fri_open_times(fri.commit_phase_base(Z.GOLDILOCKS, cw, log_len, log_final, betas), log_len)
    coeff.free()
    cw.free()
    return dict(codeword=1 << log_len, layers=log_len - log_final, ms_per_commit_phase=dt * 1e3,
                device_ms=dev / steps, codeword_elems_per_s=(1 << log_len) / dt,
                ms_per_commit_phase_transcript_order=dt_step * 1e3, query_openings=opening)


def bench_mmcs(ctx, steps):
    """Mixed-height MMCS commit (p3 MerkleTree::new over several matrices,
    compress_and_inject) on device-resident matrices: a BabyBear 2^20 x 64
    trace with a 2^18 x 32 matrix injected two layers up, two BabyBear
    2^20 x 32 matrices hashed as one row stream, and a Goldilocks 2^18 x 16
    trace with a 2^15 x 8 matrix injected."""
    import numpy as np  # noqa: F811

    out = {}
    for name, field, p2w, shapes in (("babybear_injected", Z.BABYBEAR, 16, [(1 << 20, 64), (1 << 18, 32)]),
                                     ("babybear_pair", Z.BABYBEAR, 16, [(1 << 20, 32), (1 << 20, 32)]),
                                     ("goldilocks_injected", Z.GOLDILOCKS, 16, [(1 << 18, 16), (1 << 15, 8)])):
        bufs = [ctx.random(field, h * w, 0x5EED0100 + k) for k, (h, w) in enumerate(shapes)]
        n = len(bufs)
        handles = (ctypes.c_void_p * n)(*[b.h for b in bufs])
        hs = np.array([h for h, _ in shapes], dtype=np.uint64)
        ws = np.array([w for _, w in shapes], dtype=np.uint32)

        def once():
            t = ctypes.c_void_p()
            check(ctx.lib.zkg_mmcs_commit(ctx.h, field, p2w, n, handles, hs.ctypes.data_as(ctypes.c_void_p),
                                          ws.ctypes.data_as(ctypes.c_void_p), ctypes.byref(t)))
            return t

        for _ in range(2):
            free_tree(ctx, once())
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            free_tree(ctx, once())
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / steps
        cells = sum(h * w for h, w in shapes)
        out[name] = dict(shapes=[list(x) for x in shapes], ms_per_commit=dt * 1e3, cells_per_s=cells / dt)
        # p3 open_batch of 80 seeded query indices: every matrix's row and the
        # sibling path, one zkg_mmcs_open call
        t = once()
        q = np.random.default_rng(0x5EED0081).integers(0, shapes[0][0], 80).astype(np.uint64)
        es = 8 if field == Z.GOLDILOCKS else 4
        rows = np.empty(80 * sum(w for _, w in shapes) * es, dtype=np.uint8)
        paths = np.empty(80 * (shapes[0][0].bit_length() - 1) * (p2w // 2) * es, dtype=np.uint8)
        qp, rp, pp = (a.ctypes.data_as(ctypes.c_void_p) for a in (q, rows, paths))
        check(ctx.lib.zkg_mmcs_open(t, qp, 80, rp, pp))
        t0 = time.perf_counter()
        check(ctx.lib.zkg_mmcs_open(t, qp, 80, rp, pp))
        out[name]["ms_open_80_queries"] = (time.perf_counter() - t0) * 1e3
        free_tree(ctx, t)
        for b in bufs:
            b.free()
    return out


def bench_deep(ctx, steps):
    """PCS opening, DEEP quotient over the config-2 LDE (BabyBear 2^21 x 64 rows,
    bit-reversed on the coset 31*H): opened values at one EF4 point, then the
    reduced codeword.  Algorithmic bytes: the LDE once (reduce) plus its first
    half (interpolation), and the EF4 codeword read and written."""
    n, w, lb = 1 << 20, 64, 1
    h = n << lb
    tr = ctx.random(Z.BABYBEAR, n * w, SEED_C2)
    lde = ctx.alloc(h * w * 4)
    check(ctx.lib.zkg_coset_lde_batch(ctx.h, Z.BABYBEAR, tr.h, n, w, lb, 31, lde.h))
    ro = ctx.upload(np.zeros(h * 4, dtype=np.uint32))
    rng = np.random.default_rng(SEED_C2 + 1)
    z = rng.integers(0, 0x78000001, 4).astype(np.uint32)
    alpha = rng.integers(0, 0x78000001, 4).astype(np.uint32)
    deep = Z.HipDeep(ctx, Z.BABYBEAR, lde, h, w, lb, 31)
    for _ in range(2):
        deep.open(z, alpha, 0, ro)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        deep.open(z, alpha, 0, ro)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    alg = h * w * 4 + n * w * 4 + h * 32
    for b in (tr, lde, ro):
        b.free()
    out = dict(lde_rows=h, width=w, ms_per_open=dt * 1e3, algorithmic_bytes=alg,
               achieved_GBps=alg / dt / 1e9, hbm_frac=alg / dt / HBM_PEAK)
    # the reference STARK's configuration: Goldilocks with base-field challenges,
    # a 2^20 x 64 trace at blowup 2 (2^21 x 64 LDE rows of u64)
    tr = ctx.random(Z.GOLDILOCKS, n * w, SEED_C2 + 2)
    lde = ctx.alloc(h * w * 8)
    check(ctx.lib.zkg_coset_lde_batch(ctx.h, Z.GOLDILOCKS, tr.h, n, w, lb, 7, lde.h))
    ro = ctx.upload(np.zeros(h, dtype=np.uint64))
    zg = int(rng.integers(0, 0xFFFFFFFF00000001, dtype=np.uint64))
    ag = int(rng.integers(0, 0xFFFFFFFF00000001, dtype=np.uint64))
    deep = Z.HipDeep(ctx, Z.GOLDILOCKS, lde, h, w, lb, 7)
    for _ in range(2):
        deep.open(zg, ag, 0, ro)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        deep.open(zg, ag, 0, ro)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    alg = h * w * 8 + n * w * 8 + h * 16
    for b in (tr, lde, ro):
        b.free()
    out["goldilocks"] = dict(lde_rows=h, width=w, ms_per_open=dt * 1e3, algorithmic_bytes=alg,
                             achieved_GBps=alg / dt / 1e9, hbm_frac=alg / dt / HBM_PEAK)
    return out


def air_program(width, n_c, seed):
    """A degree-3 round-function AIR: constraint k is x^3 + a_k y + b_k - u v over
    columns 2k .. 2k+3 (mod width), built with the reference's bytecode builder."""
    rng = np.random.default_rng(seed)
    p = 0xFFFFFFFF00000001
    cc = Z.ConstraintCompiler()
    for k in range(n_c):
        x, y, u, v = (2 * k) % width, (2 * k + 1) % width, (2 * k + 2) % width, (2 * k + 3) % width
        cc.load_column(x); cc.dup(); cc.dup(); cc.mul(); cc.mul()
        cc.push_const(int(rng.integers(0, 2**63, dtype=np.uint64)) % p); cc.load_column(y); cc.mul(); cc.add()
        cc.push_const(int(rng.integers(0, 2**63, dtype=np.uint64)) % p); cc.add()
        cc.load_column(u); cc.load_column(v); cc.mul(); cc.sub()
        cc.emit(k)
    return cc.finish()


CE_SHAPE = (1 << 20, 64, 1, 32)  # trace rows, width, log blowup, constraints


def bench_constraints(ctx, steps, peak_gl):
    """Constraint evaluation over a committed Goldilocks LDE (2^20 x 64 trace at
    blowup 2: 2^21 rows, bit-reversed on the coset 7*H), 32 degree-3
    constraints, interpreted and compiled.  Algorithmic bytes per evaluation:
    every LDE row read once and every constraint value written once,
    rows x (64 + 32) x 8; field products: 4 per constraint per row.  The
    quotient (constraints combined with 32 coefficients and divided by the
    vanishing polynomial, one value per row) moves rows x (64 + 1) x 8 bytes."""
    n, w, lb, n_c = CE_SHAPE
    h = n << lb
    bc, consts, n_c = air_program(w, n_c, SEED_C2 + 7)
    tr = ctx.random(Z.GOLDILOCKS, n * w, SEED_C2 + 8)
    lde = ctx.alloc(h * w * 8)
    check(ctx.lib.zkg_coset_lde_batch(ctx.h, Z.GOLDILOCKS, tr.h, n, w, lb, 7, lde.h))
    out = ctx.alloc(h * n_c * 8)
    alg = h * (w + n_c) * 8
    products = h * n_c * 4
    res = dict(lde_rows=h, width=w, constraints=n_c, instructions=len(bc), algorithmic_bytes=alg,
               gl_products=products)
    q = ctx.alloc(h * 8)
    rng = np.random.default_rng(SEED_C2 + 10)
    coef = [int(v) % 0xFFFFFFFF00000001 for v in rng.integers(0, 2**63, n_c, dtype=np.uint64)]
    alg_q = h * (w + 1) * 8
    for mode in ("interpreted", "compiled"):
        t0 = time.perf_counter()
        prog = Z.ConstraintProgram(ctx, w, bc, consts, n_c, compiled=(mode == "compiled"))
        prep = time.perf_counter() - t0
        for _ in range(2):
            prog.eval(lde, h, out)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            prog.eval(lde, h, out)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / steps
        # quotient: the same constraints combined with coefficients and divided
        # by Z_H, one value per row (each call returns when q is written)
        prog.quotient(lde, h, lb, 7, coef, q)
        t0 = time.perf_counter()
        for _ in range(steps):
            prog.quotient(lde, h, lb, 7, coef, q)
        dq = (time.perf_counter() - t0) / steps
        info = prog.info()
        prog.free()
        res[mode] = dict(ms=dt * 1e3, rows_per_s=h / dt, achieved_GBps=alg / dt / 1e9, hbm_frac=alg / dt / HBM_PEAK,
                         gl_products_per_s=products / dt, gl_product_frac=products / dt / peak_gl,
                         prepare_s=prep, shape="%d rows/lane, %d-column panels" % (info["rows_per_lane"],
                                                                                   info["panel"]),
                         quotient_ms=dq * 1e3, quotient_rows_per_s=h / dq, quotient_GBps=alg_q / dq / 1e9)
    for b in (tr, lde, out, q):
        b.free()
    res["quotient_algorithmic_bytes"] = alg_q
    res["compiled_vs_interpreted"] = res["interpreted"]["ms"] / res["compiled"]["ms"]
    return res


def c3_shapes():
    return workloads.batch_shapes(48, SEED_C3)


def bench_c3(ctx, steps, warmup):
    """The batched workload: 48 BabyBear traces committed by one zkg_commit_batch
    call (graph replay).  A step is one whole batch.  Also: one plain
    (reference-structured) call per trace, one fused zkg_commit per trace, and
    a timed pass (the same launches straight from the streams, bracketed by HIP
    events) that gives each stage's device time."""
    shapes = c3_shapes()
    bufs, jobs, cells = [], [], 0
    for k, (lg, w) in enumerate(shapes):
        n = 1 << lg
        tr = ctx.random(Z.BABYBEAR, n * w, SEED_C3 + k)
        lde = ctx.alloc((n << 1) * w * 4)
        bufs.append((tr, lde))
        jobs.append(Z.make_job(Z.BABYBEAR, tr, n, w, 1, lde))
        cells += n * w
    arr = (Z._lib.Job * len(jobs))(*jobs)
    outs = (ctypes.c_void_p * len(jobs))()

    def run():
        check(ctx.lib.zkg_commit_batch(ctx.h, arr, len(jobs), outs))
        for i in range(len(jobs)):
            free_tree(ctx, ctypes.c_void_p(outs[i]))

    for _ in range(max(1, warmup)):
        run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # stage device times: the same launches, timed with events on their streams
    ctx.set_timing(True)
    stages = np.zeros(4)
    reps = 3
    for _ in range(reps):
        run()
        stages += ctx.stage_ms()
    ctx.set_timing(False)
    stages /= reps
    # plain: one synchronous reference-structured call per trace
    t0 = time.perf_counter()
    for (tr, lde), (lg, w) in zip(bufs, shapes):
        free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit_plain, Z.BABYBEAR, 16, tr, 1 << lg, w, 1, lde))
    ctx.synchronize()
    dt_plain = time.perf_counter() - t0
    # fused single calls, one per trace (isolates the batching gain)
    t0 = time.perf_counter()
    for (tr, lde), (lg, w) in zip(bufs, shapes):
        free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit, Z.BABYBEAR, 16, tr, 1 << lg, w, 1, lde))
    ctx.synchronize()
    dt_single = time.perf_counter() - t0
    for tr, lde in bufs:
        tr.free()
        lde.free()
    perms = sum((1 << (lg + 1)) * ((w + 7) // 8) for lg, w in shapes)
    lde_bytes = sum((1 << lg) * w * 4 * 3 for lg, w in shapes)
    return dict(traces=len(jobs), cells=cells, sec_per_step=dt, batched_cells_per_s=cells / dt,
                plain_cells_per_s=cells / dt_plain, single_fused_cells_per_s=cells / dt_single,
                batched_vs_plain=dt_plain / dt, batched_vs_single_fused=dt_single / dt, batch_ms=dt * 1e3,
                plain_ms=dt_plain * 1e3, single_fused_ms=dt_single * 1e3,
                stage_ms=dict(lde=stages[0], row_hash=stages[1], merkle=stages[2], timed_batch=stages[3]),
                row_hash_permutations=perms, lde_algorithmic_bytes=lde_bytes)


def bench_c5(ctx, n_req, load, slots=2, max_batch=8):
    """Served stream vs the plain path on the same Poisson arrival trace.

    The plain path is a single FIFO server making one synchronous
    reference-structured call per request: its service times are measured
    back to back and its latencies follow from the queue recursion
    done_i = max(arrival_i, done_{i-1}) + service_i.  The arrival rate is
    `load` times the plain path's capacity (requests/s), so load > 1 is a
    stream the plain path cannot keep up with.
    """
    reqs, unit_gaps = workloads.serve_requests(n_req, SEED_C5)
    bufs, jobs, cells = [], [], 0
    for k, (field, lg, w, lb) in enumerate(reqs):
        n = 1 << lg
        es = 8 if field == Z.GOLDILOCKS else 4
        tr = ctx.random(field, n * w, SEED_C5 + k)
        lde = ctx.alloc((n << lb) * w * es)
        bufs.append((tr, lde))
        jobs.append(Z.make_job(field, tr, n, w, lb, lde))
        cells += n * w
    ctx.synchronize()
    # plain service times (one warm pass, then measured)
    for rep in range(2):
        service = []
        for (tr, lde), (field, lg, w, lb) in zip(bufs, reqs):
            a = time.perf_counter()
            free_tree(ctx, commit_once(ctx, ctx.lib.zkg_commit_plain, field, 16, tr, 1 << lg, w, lb, lde))
            service.append(time.perf_counter() - a)
    service = np.array(service)
    rate = load * n_req / service.sum()
    arrivals = np.cumsum(unit_gaps / rate)
    done, prev = np.empty(n_req), 0.0
    for i in range(n_req):
        prev = max(arrivals[i], prev) + service[i]
        done[i] = prev
    plain_lat = (done - arrivals) * 1e3
    plain_span = done[-1] - arrivals[0]

    def serve(offsets):
        """Submits job k at offsets[k] seconds; returns (latencies ms, span s)."""
        srv = Z.Server(ctx, slots=slots, max_batch=max_batch)
        lat, pending = [], {}
        t0 = time.perf_counter()
        last_done = t0
        k = 0
        while k < n_req or pending:
            now = time.perf_counter() - t0
            while k < n_req and now >= offsets[k]:
                pending[srv.submit(jobs[k])] = k
                k += 1
            for t in list(pending):
                r = srv.poll(t)
                if r is not None:
                    free_tree(ctx, r[0])
                    lat.append(r[1])
                    last_done = time.perf_counter()
                    del pending[t]
            time.sleep(0.0001)
        srv.stop()
        return np.array(lat), last_done - t0

    # capacity: every request queued at t = 0 (plain: back-to-back service).
    # One warm pass first, as the plain path has: it builds the coset tables of
    # the request shapes (resident from then on) and fills the allocator cache.
    serve(np.zeros(n_req))
    _, sat_span = serve(np.zeros(n_req))
    # latency: the Poisson trace above, replayed into the serving loop
    lat, span = serve(arrivals - arrivals[0])
    # where the served ratio is capped: each field's requests committed as one
    # timed batch give that field's row-hash device time; with the LDE and the
    # Merkle levels free, served time could not drop below the two row hashes
    cap = {}
    ctx.set_timing(True)
    for field, name in ((Z.GOLDILOCKS, "goldilocks"), (Z.BABYBEAR, "babybear")):
        idx = [k for k, r in enumerate(reqs) if r[0] == field]
        arr = (Z._lib.Job * len(idx))(*[jobs[k] for k in idx])
        outs = (ctypes.c_void_p * len(idx))()
        st = np.zeros(4)
        for rep in range(3):
            check(ctx.lib.zkg_commit_batch(ctx.h, arr, len(idx), outs))
            for i in range(len(idx)):
                free_tree(ctx, ctypes.c_void_p(outs[i]))
            if rep:
                st += ctx.stage_ms()
        st /= 2
        cap[name] = dict(requests=len(idx), plain_ms=float(service[idx].sum() * 1e3), batch_lde_ms=st[0],
                         batch_row_hash_ms=st[1], batch_merkle_ms=st[2], batch_ms=st[3])
    ctx.set_timing(False)
    hash_ms = sum(c["batch_row_hash_ms"] for c in cap.values())
    cap["served_vs_plain_if_only_row_hashes_cost"] = service.sum() * 1e3 / hash_ms
    for tr, lde in bufs:
        tr.free()
        lde.free()
    return dict(requests=n_req, cells=cells,
                capacity=dict(served_cells_per_s=cells / sat_span, plain_cells_per_s=cells / service.sum(),
                              served_vs_plain=service.sum() / sat_span),
                poisson=dict(arrival_rate_hz=rate, load_vs_plain_capacity=load,
                             served_cells_per_s=cells / span, plain_cells_per_s=cells / plain_span,
                             p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                             plain_p50_ms=float(np.percentile(plain_lat, 50)),
                             plain_p99_ms=float(np.percentile(plain_lat, 99))),
                slots=slots, max_batch=max_batch, row_hash_cap=cap)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(max_log_multi=16, max_log_one=13):
    """The oracle (C restatement of the p3 CPU path: coset LDE, Poseidon2-16
    sponge rows, Merkle) on a bounded sample of the batched workload: the draw's
    traces of at most 2^max_log_multi rows, on every core this process may run
    on, and those of at most 2^max_log_one rows on one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O

    p2 = O.Poseidon2(O.BB, 16, 42)
    shapes = c3_shapes()

    def run(max_log, threads):
        O.set_threads(threads)
        sample = [(k, lg, w) for k, (lg, w) in enumerate(shapes) if lg <= max_log]
        cells, secs = 0, 0.0
        for k, lg, w in sample:
            n = 1 << lg
            tr = O.gen(O.BB, SEED_C3 + k, n * w).reshape(n, w)
            t0 = time.perf_counter()
            O.commit(p2, tr, 1)
            secs += time.perf_counter() - t0
            cells += n * w
        return cells, secs, len(sample), O.get_threads()

    # every core of this GPU's share of the host: the affinity set, capped at
    # the box's per-GPU CPU share (OMP_NUM_THREADS there)
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS") or affinity)
    cores = max(1, min(affinity, share))
    cells, secs, nt, used = run(max_log_multi, cores)
    cells1, secs1, nt1, _ = run(max_log_one, 1)
    # constraint evaluation: the oracle's stack machine on 1/8 of the LDE rows
    O.set_threads(cores)
    ce_rows = (CE_SHAPE[0] << CE_SHAPE[2]) // 8
    bc, consts, n_c = air_program(CE_SHAPE[1], CE_SHAPE[3], SEED_C2 + 7)
    ce_tr = O.gen(O.GL, SEED_C2 + 9, ce_rows * CE_SHAPE[1]).reshape(ce_rows, CE_SHAPE[1])
    t1 = time.perf_counter()
    O.constraint_eval_gl(ce_tr, bc, consts, n_c)
    ce_dt = time.perf_counter() - t1
    # the single Goldilocks commit (2^14 x 8, blowup 2, Poseidon2-8): the
    # reference's CPU-runnable case, whole, on every core and on one
    p2_gl8 = O.Poseidon2(O.GL, 8, 42)
    tr_c1 = O.gen(O.GL, SEED_C1, (1 << 14) * 8).reshape(1 << 14, 8)
    c1_ms = {}
    for threads in (cores, 1):
        O.set_threads(threads)
        O.commit(p2_gl8, tr_c1, 1)
        t1 = time.perf_counter()
        for _ in range(3):
            O.commit(p2_gl8, tr_c1, 1)
        c1_ms[threads] = (time.perf_counter() - t1) / 3 * 1e3
    O.set_threads(cores)
    cpu = _cpu_model()
    return dict(value=cells / secs, unit="cells/s", cores=used, kind="port",
                sample="the %d traces of at most 2^%d rows of the 48-trace batched draw (%d cells, blowup 2, "
                       "Poseidon2-16 rows, Merkle), committed one after another by the oracle on %d threads; "
                       "%s; %.1f s wall" % (nt, max_log_multi, cells, used, cpu, secs),
                one_core=dict(value=cells1 / secs1, unit="cells/s", cores=1,
                              sample="the %d traces of at most 2^%d rows (%d cells) on one thread; %.1f s wall"
                                     % (nt1, max_log_one, cells1, secs1)),
                host_threads_visible=os.cpu_count(), affinity=affinity,
                c1=dict(workload="Goldilocks 2^14 x 8 trace, coset LDE blowup 2, Poseidon2-8 rows, Merkle",
                        ms_per_commit=c1_ms[cores], cores=cores, ms_per_commit_one_core=c1_ms[1], kind="port"),
                constraint_eval=dict(rows_per_s=ce_rows / ce_dt, cores=used, kind="port",
                                     sample="%d of the %d LDE rows x %d columns, %d constraints (%d instructions); "
                                            "%.2f s wall" % (ce_rows, CE_SHAPE[0] << CE_SHAPE[2], CE_SHAPE[1], n_c,
                                                             len(bc), ce_dt)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batches (one batch = the 48-trace commit)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-legs", action="store_true",
                    help="only the batched headline (with its plain ratio), no single-trace / FRI / DEEP / MMCS / "
                         "constraint legs")
    ap.add_argument("--only-constraints", action="store_true",
                    help="run only the constraint-evaluation leg (for profiling) and print its JSON")
    ap.add_argument("--only-c1", action="store_true",
                    help="run only the Goldilocks 2^14 x 8 single-commit leg (for profiling) and print its JSON")
    ap.add_argument("--only-gl-hash", action="store_true",
                    help="run only the Goldilocks 2^21-row x 97 row-hash leg (for profiling) and print its JSON")
    ap.add_argument("--only-c4", action="store_true",
                    help="run only the FRI commit-phase leg (for profiling) and print its JSON")
    ap.add_argument("--only-c2", action="store_true",
                    help="run only the BabyBear 2^20 x 64 single-commit leg (for profiling) and print its JSON")
    ap.add_argument("--only-served", action="store_true",
                    help="run only the served-stream leg (for profiling) and print its JSON")
    ap.add_argument("--only-deep", action="store_true",
                    help="run only the PCS-opening (DEEP quotient) leg and print its JSON")
    ap.add_argument("--only-mmcs", action="store_true",
                    help="run only the mixed-height MMCS leg (commits and openings) and print its JSON")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-load", type=float, default=1.5,
                    help="Poisson arrival rate as a multiple of the plain path's capacity")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see --gpus 1)", file=sys.stderr)
        sys.exit(2)

    ctx = Z.Context(0)
    for field, width in ((Z.BABYBEAR, 16), (Z.GOLDILOCKS, 16), (Z.GOLDILOCKS, 8)):
        ctx.load_constants(field, width, 42)
    log("probe: VALU ceilings")
    peak_bb = probe_peak(0)
    peak_gl = probe_peak(1)
    peak_mul32 = probe_peak(2)
    peak_bbf = probe_peak(4)
    copy_gbps = probe_copy()
    log("probe: %.3e BabyBear Montgomery products/s, %.3e BabyBear FP64 products/s, %.3e Goldilocks products/s, "
        "%.3e mul_lo_u32/s, copy %.0f GB/s" % (peak_bb, peak_bbf, peak_gl, peak_mul32, copy_gbps))
    peak_bb = max(peak_bb, peak_bbf)
    if args.only_constraints:
        print(json.dumps({"constraint_eval": bench_constraints(ctx, max(2, args.steps), peak_gl)}))
        ctx.close()
        return
    if args.only_gl_hash:
        print(json.dumps({"gl_row_hash": bench_gl_row_hash(ctx, max(3, args.steps))}))
        ctx.close()
        return
    if args.only_c4:
        print(json.dumps({"c4_fri": bench_c4(ctx, args.steps), "gl_fri": bench_fri_gl(ctx, args.steps)}))
        ctx.close()
        return
    if args.only_deep:
        print(json.dumps({"pcs_open_deep": bench_deep(ctx, max(5, args.steps))}))
        ctx.close()
        return
    if args.only_mmcs:
        print(json.dumps({"mmcs": bench_mmcs(ctx, max(5, args.steps // 2))}))
        ctx.close()
        return
    if args.only_c1:
        print(json.dumps({"c1": bench_c1(ctx, max(5, args.steps))}))
        ctx.close()
        return
    if args.only_c2:
        print(json.dumps({"c2": c2_leg(ctx, args.steps, args.warmup, peak_bb)}))
        ctx.close()
        return
    if args.only_served:
        print(json.dumps({"served": bench_c5(ctx, args.serve_requests, args.serve_load)}, default=float))
        ctx.close()
        return

    log("batched: 48 BabyBear traces, %d warmup + %d timed batches" % (args.warmup, args.steps))
    c3 = bench_c3(ctx, args.steps, args.warmup)
    log("batched: %s" % c3)
    counters = profiled_counters(BATCH_COUNTERS)
    hash_s = c3["stage_ms"]["row_hash"] * 1e-3
    products = c3["row_hash_permutations"] * BB_PERM_MULS
    insts = counters.get("valu_insts_hash_jobs")
    roof = {"kernel": "k_hash_jobs<Bb,16>", "bound": "valu",
            "achieved": insts / hash_s / 1e9 if insts else None, "peak": VALU_ISSUE_PEAK / 1e9,
            "unit": "G VALU wave-instructions/s", "frac": insts / hash_s / VALU_ISSUE_PEAK if insts else None,
            "traffic": counters.get("traffic_hash_jobs_bytes"),
            "launch_ms": hash_s * 1e3, "valu_insts_per_launch": insts,
            "product_ceiling": {"achieved": products / hash_s / 1e9, "peak": peak_bb / 1e9, "unit": "Gmodmul/s",
                                "frac": products / hash_s / peak_bb},
            "note": "the batch's one row-hash launch (%d Poseidon2-16 permutations over the 48 LDEs); achieved = "
                    "profiled VALU wave-instructions per launch over the launch's HIP-event duration in this run, "
                    "peak = 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction; product_ceiling: 772 "
                    "BabyBear products per permutation against the measured product rate (exact FP64 or "
                    "Montgomery, 8 independent chains per lane over the whole chip)" % c3["row_hash_permutations"]}
    out = {
        "metric": "trace cells committed/sec, batched vs plain per-trace GPU; p99 commit latency",
        "value": c3["batched_cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c3["batch_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "batched commit of 48 BabyBear traces, heights 2^U{12..20}, widths U{8,16..256} "
                               "(seeded draw, %d cells): coset LDE blowup 2 (shift 31), Poseidon2-16 row sponge "
                               "(8+13 rounds, rate 8), Merkle to one root each; one zkg_commit_batch call "
                               "(hipGraph replay) per step" % c3["cells"],
                   "traces": c3["traces"], "cells": c3["cells"], "log_blowup": 1, "field": "babybear",
                   "poseidon2_width": 16},
        "batched_vs_plain": c3["batched_vs_plain"],
        "batched_vs_single_fused": c3["batched_vs_single_fused"],
        "plain_cells_per_s": c3["plain_cells_per_s"],
        "single_fused_cells_per_s": c3["single_fused_cells_per_s"],
        "stage_ms": c3["stage_ms"],
        "roofline": roof,
        "probe": {"bb_product_ceiling_per_s": peak_bb, "bb_fp64_products_per_s": peak_bbf,
                  "gl_products_per_s": peak_gl, "mul_lo_u32_per_s": peak_mul32, "copy_1GiB_GBps": copy_gbps},
    }
    if not args.skip_legs:
        log("c2: single BabyBear 2^20 x 64 commit")
        out["c2"] = c2_leg(ctx, max(5, args.steps // 2), args.warmup, peak_bb)
        log("c2: %s" % out["c2"])
        log("c1: goldilocks 2^14 x 8")
        out["c1"] = bench_c1(ctx, max(5, args.steps))
        log("c1: %s" % out["c1"])
        out["gl_row_hash"] = bench_gl_row_hash(ctx, 5)
        log("gl_row_hash: %s" % out["gl_row_hash"])
        log("c4: fri commit phase 2^22 -> 2^8")
        out["c4_fri"] = bench_c4(ctx, max(3, args.steps // 2))
        log("c4: %s" % out["c4_fri"])
        out["gl_fri"] = bench_fri_gl(ctx, max(3, args.steps // 2))
        log("gl_fri: %s" % out["gl_fri"])
        log("deep: PCS opening DEEP quotient over the config-2 LDE")
        out["pcs_open_deep"] = bench_deep(ctx, max(5, args.steps))
        log("deep: %s" % out["pcs_open_deep"])
        log("mmcs: mixed-height commits")
        out["mmcs"] = bench_mmcs(ctx, max(5, args.steps // 2))
        log("mmcs: %s" % out["mmcs"])
        log("constraints: 32 degree-3 constraints over a Goldilocks 2^21 x 64 LDE")
        out["constraint_eval"] = bench_constraints(ctx, max(5, args.steps), peak_gl)
        log("constraints: %s" % out["constraint_eval"])
    if not args.skip_serve:
        log("served: continuous batching")
        out["served"] = bench_c5(ctx, args.serve_requests, args.serve_load)
        log("served: %s" % out["served"])
    ctx.close()
    if not args.skip_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
        log("cpu baseline: %s" % out["cpu_baseline"])
    print(json.dumps(out))


def c2_leg(ctx, steps, warmup, peak_bb):
    """Single large commit (BabyBear 2^20 x 64, blowup 2) with its row-hash and
    LDE rooflines."""
    c2 = bench_c2(ctx, steps, warmup)
    n, w, lb = c2["n"], c2["w"], c2["lb"]
    lde_ms, hash_ms, merkle_ms = c2["stage_ms"][0], c2["stage_ms"][1], c2["stage_ms"][2]
    leaves = n << lb
    hash_muls = leaves * ((w + 7) // 8) * BB_PERM_MULS
    lde_bytes = n * w * 4 * (1 + (1 << lb))
    counters = profiled_counters(C2_COUNTERS)
    return {
        "workload": "BabyBear commit: 2^20 x 64 trace, coset LDE blowup 2 (2^21 rows, shift 31), Poseidon2-16 "
                    "row sponge, Merkle to one root",
        "cells_per_s": n * w / c2["sec_per_step"], "ms_per_commit": c2["sec_per_step"] * 1e3,
        "stage_ms": {"lde": lde_ms, "row_hash": hash_ms, "merkle": merkle_ms},
        "plain": {"cells_per_s": n * w / c2["plain_sec"], "ms_per_commit": c2["plain_sec"] * 1e3,
                  "fused_vs_plain": c2["plain_sec"] / c2["sec_per_step"]},
        "roofline_row_hash": {"kernel": "k_hash_rows<Bb,16>", "bound": "valu",
                              "valu_issue": valu_issue(counters.get("valu_insts_hash_rows"), hash_ms),
                              "product_ceiling": {"achieved": hash_muls / (hash_ms * 1e-3) / 1e9,
                                                  "peak": peak_bb / 1e9, "unit": "Gmodmul/s",
                                                  "frac": hash_muls / (hash_ms * 1e-3) / peak_bb},
                              "traffic": counters.get("traffic_hash_rows_bytes")},
        "roofline_lde": {"kernels": "k_fs_pass1 + k_lde_mid + k_lde_final", "bound": "hbm",
                         "achieved": lde_bytes / (lde_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": lde_bytes / (lde_ms * 1e-3) / HBM_PEAK, "algorithmic_bytes": lde_bytes,
                         "traffic": counters.get("traffic_lde_bytes"),
                         "valu_issue": valu_issue(counters.get("valu_insts_lde"), lde_ms),
                         "note": "N*W*4*(1+B) algorithmic bytes over the summed duration of the three passes"},
    }


if __name__ == "__main__":
    main()
