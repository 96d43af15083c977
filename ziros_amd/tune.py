"""Autotuner for libzkg: writes the static tuning layer (csrc/zkg_tuned.h).

Two settings are tuned on the device, both chosen at run time through the
tuning layers of zkg_tuning.hip (override > learned > static > built-in):

* the tile plan of the column transforms per shape.  Every plan of
  csrc/zkg_plans.h is compiled into the library; for each shape of a grid
  covering the single, batched and served workloads (both fields, heights
  2^12..2^22, widths 8..256, blowup 2 and 4, and plain NTTs) the tuner forces
  each plan through the override layer, times the transform (median of 3 after
  a warm-up) and keeps the fastest if it beats plan 0 by more than
  PLAN_MARGIN.  At run time a shape takes the entry of the nearest measured
  shape of its field.
* the work-form splits of the row hash and Merkle levels per (field,
  Poseidon2 width): for each candidate value the tuner times fused commits
  of several heights and keeps the value with the least total if it beats the
  current value by more than SPLIT_MARGIN.

    python -m ziros_amd.tune --measure   # on the GPU: time, write zkg_tuned.h
    python -m ziros_amd.build            # rebuild with the new static layer
"""
import argparse
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED_H = os.path.join(HERE, "csrc", "zkg_tuned.h")

GL, BB = 0, 1
# (field, log2 height, width, log_blowup); blowup 0 = plain NTT
LDE_GRID = [(f, lg, w, lb) for f in (BB, GL) for lg in (12, 14, 16, 18, 20, 22) for w in (8, 32, 128, 256)
            for lb in (1, 2) if not (lg == 22 and w >= 128) and not (f == GL and lg == 22 and w == 32 and lb == 2)]
LDE_GRID += [(f, lg, w, 0) for f in (BB, GL) for lg in (14, 18, 22) for w in (8, 64)]

# split candidates and the commits (log2 height, width, log_blowup) that time them
SPLIT_CANDIDATES = {
    "merkle_lanes_log2": {(BB, 16): (13, 14, 15, 16, 17), (GL, 16): (11, 12, 13, 14, 15), (GL, 8): (12, 13, 14, 15, 16)},
    "fused_levels": {(BB, 16): (2, 3, 4), (GL, 16): (2, 3, 4), (GL, 8): (2, 3, 4)},
    "merkle_chunks_log2": {(GL, 16): (14, 15, 16, 17), (GL, 8): (14, 15, 16, 17)},
    "hash_chunks_log2": {(GL, 16): (13, 14, 15, 16), (GL, 8): (13, 14, 15, 16)},
}
SPLIT_COMMITS = [(12, 16, 1), (14, 8, 1), (16, 16, 1), (18, 16, 1), (20, 16, 1)]
SPLIT_ORDER = ("merkle_lanes_log2", "merkle_chunks_log2", "hash_chunks_log2", "fused_levels")


# A measured winner replaces the default only if it is faster by more than
# this fraction: closer than that, the order of two candidates is run-to-run
# noise (wall-clock timing of calls of 0.05 .. 20 ms).
PLAN_MARGIN = 0.02
SPLIT_MARGIN = 0.01


def _median(ts):
    return sorted(ts)[len(ts) // 2]


def choose(times, default, margin):
    """Index of the fastest entry of `times`, unless entry `default` is within
    `margin` of it."""
    best = min(range(len(times)), key=lambda i: times[i])
    return default if times[default] <= (1 + margin) * times[best] else best


def measure_plans(ctx, log=print):
    import ziros_amd as Z
    from ziros_amd import host as H
    from ziros_amd._lib import check

    n_plans = len(H.lde_plans())
    table = []
    for field, lg, w, lb in LDE_GRID:
        n = 1 << lg
        es = 8 if field == GL else 4
        tr = ctx.random(field, n * w, 1000 + lg)
        out = ctx.alloc((n << lb) * w * es) if lb else None
        times = []
        for plan in range(n_plans):
            H.set_lde_plans(H.TUNE_OVERRIDE, [(-1, -1, -1, -1, plan)])

            def once():
                if lb:
                    check(ctx.lib.zkg_coset_lde_batch(ctx.h, field, tr.h, n, w, lb, Z.GENERATOR[field], out.h))
                else:
                    check(ctx.lib.zkg_ntt_batch(ctx.h, field, tr.h, n, w, 0))

            once()
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                once()
                ts.append(time.perf_counter() - t0)
            times.append(_median(ts) * 1e3)
        H.set_lde_plans(H.TUNE_OVERRIDE, [])
        best = choose(times, 0, PLAN_MARGIN)
        table.append(((field, lg, w, lb), times, best))
        log("plan %d/%d/%d/%d %s -> %d" % (field, lg, w, lb, " ".join("%.3f" % t for t in times), best))
        tr.free()
        if out:
            out.free()
    return table


def measure_splits(ctx, log=print):
    import ziros_amd as Z
    from ziros_amd import host as H
    from ziros_amd._lib import check
    import ctypes

    bufs = {}
    for field in (BB, GL):
        for lg, w, lb in SPLIT_COMMITS:
            n = 1 << lg
            es = 8 if field == GL else 4
            bufs[(field, lg, w, lb)] = (ctx.random(field, n * w, 2000 + lg), ctx.alloc((n << lb) * w * es))

    def total(field, p2w):
        s = 0.0
        for lg, w, lb in SPLIT_COMMITS:
            tr, lde = bufs[(field, lg, w, lb)]
            ts = []
            for rep in range(4):
                t = ctypes.c_void_p()
                t0 = time.perf_counter()
                check(ctx.lib.zkg_commit(ctx.h, field, p2w, tr.h, 1 << lg, w, lb, lde.h, ctypes.byref(t)))
                dt = time.perf_counter() - t0
                ctx.lib.zkg_tree_free(t)
                if rep:
                    ts.append(dt)
            s += _median(ts)
        return s * 1e3

    chosen = {}
    for key in ((BB, 16), (GL, 16), (GL, 8)):
        field, p2w = key
        cur = H.splits(field, p2w)
        rows = []
        for name in SPLIT_ORDER:
            cands = SPLIT_CANDIDATES[name].get(key)
            if not cands:
                continue
            res = []
            for v in cands:
                trial = dict(cur, **{name: v})
                H.set_splits(H.TUNE_OVERRIDE, [(field, p2w, trial["merkle_lanes_log2"], trial["merkle_chunks_log2"],
                                                trial["hash_chunks_log2"], trial["fused_levels"])])
                res.append((total(field, p2w), v))
            H.set_splits(H.TUNE_OVERRIDE, [])
            values = [v for _, v in res]
            best_v = values[choose([t for t, _ in res], values.index(cur[name]) if cur[name] in values else 0,
                                   SPLIT_MARGIN)]
            cur[name] = best_v
            rows.append((name, res))
            log("split %d/%d %s %s -> %d" % (field, p2w, name, " ".join("%d:%.3f" % (v, t) for t, v in res), best_v))
        chosen[key] = (cur, rows)
    for tr, lde in bufs.values():
        tr.free()
        lde.free()
    return chosen


def write_header(plans, splits, out_path=TUNED_H):
    from ziros_amd import host as H

    names = H.lde_plans()
    lines = [
        "// zkg_tuned.h -- static tuning layer (see zkg_tuning.hip), written by ziros_amd/tune.py --measure on",
        "// %s." % _device_name(),
        "//",
        "// Transform plans: milliseconds per call (median of 3) for each plan, per shape",
        "// (field 1 BabyBear / 0 Goldilocks, log2 height, width, log2 blowup; blowup 0 = plain NTT).",
        "// Plans: %s." % ", ".join("%d %s" % (i, n) for i, n in enumerate(names)),
        "//",
    ]
    for (f, lg, w, lb), times, best in plans:
        lines.append("//   %d/%d/%d/%d  %s  -> %d" % (f, lg, w, lb, " ".join("%7.3f" % t for t in times), best))
    lines.append("//")
    lines.append("// Split levels: fused-commit milliseconds summed over heights %s, each value tried in turn"
                 % ", ".join("2^%d x %d" % (lg, w) for lg, w, _ in SPLIT_COMMITS))
    lines.append("// with the values chosen before it.")
    for (f, p2w), (cur, rows) in splits.items():
        for name, res in rows:
            lines.append("//   %d/%d %s: %s" % (f, p2w, name, " ".join("%d=%.3f" % (v, t) for t, v in res)))
    lines.append("#pragma once")
    ent = ", ".join("{%d, %d, %d, %d, %d}" % (f, lg, (w - 1).bit_length(), lb, best)
                    for (f, lg, w, lb), _, best in plans)
    lines.append("#define ZKG_TUNED_LDE_TABLE {%s}" % ent)
    sp = ", ".join("{%d, %d, %d, %d, %d, %d}" % (f, p2w, c["merkle_lanes_log2"],
                                                  c["merkle_chunks_log2"] if f == GL else -1,
                                                  c["hash_chunks_log2"] if f == GL else -1, c["fused_levels"])
                   for (f, p2w), (c, _) in splits.items())
    lines.append("#define ZKG_TUNED_SPLIT_TABLE {%s}" % sp)
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _device_name():
    try:
        r = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=30)
        for line in r.stdout.splitlines():
            if "Marketing Name" in line and "MI" in line:
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "gfx950"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--measure", action="store_true", required=True)
    ap.add_argument("--out", default=TUNED_H, help="where the header is written")
    ap.add_argument("--skip-splits", action="store_true")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import ziros_amd as Z

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ctx = Z.Context(0)
    for field, width in ((Z.BABYBEAR, 16), (Z.GOLDILOCKS, 16), (Z.GOLDILOCKS, 8)):
        ctx.load_constants(field, width, 42)
    plans = measure_plans(ctx, log)
    splits = {} if args.skip_splits else measure_splits(ctx, log)
    ctx.close()
    write_header(plans, splits, args.out)
    log("wrote %s" % args.out)


if __name__ == "__main__":
    main()
