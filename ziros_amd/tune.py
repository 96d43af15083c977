"""Autotuner for the LDE tile plans of libzkg.

The in-LDS transforms split a 2^L-point tile into radix rounds of at most
2^ZKG_PLAN_RMAX32 (BabyBear) / 2^ZKG_PLAN_RMAX64 (Goldilocks) points per lane.
The radix fixes the tile shape: columns per tile = 256 * radix / 2^L.  It also
fixes the kernels' registers and LDS, and with them the blocks per CU.
ZKG_PLAN_RMAX64_SMALL applies to Goldilocks tiles of at most
2^ZKG_PLAN_SMALL_L64 points.  ZKG_TILE_STORE_UNROLL sets how far the tile
store loops unroll, and ZKG_LDE_MID_WAVES32/64 the waves per SIMD (blocks per
CU) that LDE pass B is compiled for, BabyBear and Goldilocks.  The tuner builds
every candidate and times the coset LDE of each on shapes that cover the
single, batched and served workloads (heights 2^13..2^22, widths 8..256,
blowup 2 and 4).  It writes the candidate with the least score (the sum over
shapes of its time relative to the best candidate on that shape) into
csrc/zkg_tuned.h, which the build picks up:

    python -m ziros_amd.tune --build     # compile the candidates (no GPU needed)
    python -m ziros_amd.tune --measure   # time them on the GPU, write zkg_tuned.h
    python -m ziros_amd.build            # rebuild with the tuned plan
"""
import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNE_DIR = os.path.join(HERE, "_tune")
TUNED_H = os.path.join(HERE, "csrc", "zkg_tuned.h")

CANDIDATES = {
    "r6_r4_u8": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_TILE_STORE_UNROLL": 8},
    "r6_r5_u8": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 5, "ZKG_TILE_STORE_UNROLL": 8},
    "r5_r4_u8": {"ZKG_PLAN_RMAX32": 5, "ZKG_PLAN_RMAX64": 4, "ZKG_TILE_STORE_UNROLL": 8},
    "r6_r4_u4": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_TILE_STORE_UNROLL": 4},
    "r6_r4_u16": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_TILE_STORE_UNROLL": 16},
    "r5_r5_u8": {"ZKG_PLAN_RMAX32": 5, "ZKG_PLAN_RMAX64": 5, "ZKG_TILE_STORE_UNROLL": 8},
    # pass B held to 3 waves per SIMD (168 registers) instead of the compiler's 2
    "r6_r4_u4_o3": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_TILE_STORE_UNROLL": 4, "ZKG_LDE_MID_WAVES32": 3},
    "r6_r4_u4_o33": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_TILE_STORE_UNROLL": 4, "ZKG_LDE_MID_WAVES32": 3,
                     "ZKG_LDE_MID_WAVES64": 3},
    # Goldilocks tiles of up to 2^8 points with 2^3-point rounds: narrow small matrices fill the block
    "r6_r4s3_u8": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_PLAN_RMAX64_SMALL": 3, "ZKG_PLAN_SMALL_L64": 8,
                   "ZKG_TILE_STORE_UNROLL": 8},
    "r6_r4s3_u4": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_PLAN_RMAX64_SMALL": 3, "ZKG_PLAN_SMALL_L64": 8,
                   "ZKG_TILE_STORE_UNROLL": 4},
    # two-round plans with the full radix first (BabyBear 2^10 tiles as 2^6 x 2^4: 16-column tiles)
    "r6h_r4s3_u8": {"ZKG_PLAN_RMAX32": 6, "ZKG_PLAN_RMAX64": 4, "ZKG_PLAN_RMAX64_SMALL": 3, "ZKG_PLAN_SMALL_L64": 8,
                    "ZKG_TILE_STORE_UNROLL": 8, "ZKG_PLAN_FIRST_HEAVY": 1},
    # BabyBear 2^10-point tiles in three rounds (2^4, 2^3, 2^3): fewer live registers per lane
    "r4_r4s3_u8": {"ZKG_PLAN_RMAX32": 4, "ZKG_PLAN_RMAX64": 4, "ZKG_PLAN_RMAX64_SMALL": 3, "ZKG_PLAN_SMALL_L64": 8,
                   "ZKG_TILE_STORE_UNROLL": 8},
}

# (field, log2 height, width, log_blowup); field 1 = BabyBear, 0 = Goldilocks
SHAPES = [
    (1, 20, 64, 1), (1, 22, 16, 1), (1, 16, 128, 2), (1, 13, 256, 1), (1, 18, 48, 1),
    (0, 20, 64, 1), (0, 18, 64, 2), (0, 16, 32, 1), (0, 21, 16, 1), (0, 14, 96, 2), (0, 14, 8, 1),
]


def _defines(cfg):
    # ZKG_TUNING_CANDIDATE: the installed zkg_tuned.h must not override the candidate
    return ["-DZKG_TUNING_CANDIDATE=1"] + ["-D%s=%d" % kv for kv in sorted(cfg.items())]


def build_candidates():
    from .build import build_variant

    for name, cfg in CANDIDATES.items():
        lib = build_variant(os.path.join(TUNE_DIR, name), _defines(cfg))
        print("built", name, lib, flush=True)


def time_shapes(reps=3):
    """Runs in a child process whose ZKG_LIBRARY names one candidate."""
    import ziros_amd as Z
    from ziros_amd._lib import check

    ctx = Z.Context(0)
    out = []
    for field, lg, w, lb in SHAPES:
        n = 1 << lg
        es = 8 if field == Z.GOLDILOCKS else 4
        tr = ctx.random(field, n * w, 1000 + lg)
        lde = ctx.alloc((n << lb) * w * es)
        shift = Z.GENERATOR[field]
        check(ctx.lib.zkg_coset_lde_batch(ctx.h, field, tr.h, n, w, lb, shift, lde.h))
        ctx.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            check(ctx.lib.zkg_coset_lde_batch(ctx.h, field, tr.h, n, w, lb, shift, lde.h))
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        out.append(sorted(ts)[len(ts) // 2] * 1e3)
        tr.free()
        lde.free()
        print("shape", field, lg, w, lb, "%.3f ms" % out[-1], file=sys.stderr, flush=True)
    ctx.close()
    print(json.dumps(out), flush=True)


def measure(out_path=TUNED_H):
    results = {}
    for name in CANDIDATES:
        lib = os.path.join(TUNE_DIR, name, "libzkg.so")
        if not os.path.exists(lib):
            raise SystemExit("candidate %s is not built: run python -m ziros_amd.tune --build" % name)
        env = dict(os.environ, ZKG_LIBRARY=lib)
        print("timing", name, flush=True)
        r = subprocess.run([sys.executable, "-m", "ziros_amd.tune", "--time"], env=env, capture_output=True,
                           text=True, timeout=300)
        if r.returncode != 0:
            raise SystemExit("timing %s failed:\n%s" % (name, r.stderr[-2000:]))
        results[name] = json.loads(r.stdout.strip().splitlines()[-1])
        print(name, " ".join("%.3f" % t for t in results[name]), flush=True)
    best = min(results, key=lambda k: score(results, k))
    write_header(best, results, out_path)
    print("best", best, "->", out_path, flush=True)


def score(results, name):
    """Sum over shapes of the candidate's time relative to the best time on
    that shape: every shape weighs the same, whatever its size."""
    best = [min(ts[i] for ts in results.values()) for i in range(len(SHAPES))]
    return sum(t / b for t, b in zip(results[name], best))


def write_header(best, results, out_path):
    lines = [
        "// zkg_tuned.h -- LDE tile plan chosen by ziros_amd/tune.py (python -m ziros_amd.tune --measure)",
        "// on %s.  Coset LDE milliseconds per shape (field: 1 BabyBear, 0 Goldilocks;" % _device_name(),
        "// log2 height, width, log_blowup), median of 3.  Score = sum over shapes of the time relative to",
        "// the best candidate on that shape; the least score wins.",
        "//",
        "// %-12s %s  score" % ("candidate", " ".join("%d/%d/%d/%d" % s for s in SHAPES)),
    ]
    for name in sorted(results, key=lambda k: score(results, k)):
        lines.append("// %-12s %s  %.3f" % (name, " ".join("%.3f" % t for t in results[name]), score(results, name)))
    lines.append("#pragma once")
    for k, v in sorted(CANDIDATES[best].items()):
        lines.append("#define %s %d" % (k, v))
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
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--build", action="store_true")
    g.add_argument("--measure", action="store_true")
    g.add_argument("--time", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--out", default=TUNED_H, help="where --measure writes the header")
    args = ap.parse_args()
    if args.build:
        build_candidates()
    elif args.measure:
        measure(args.out)
    else:
        time_shapes()


if __name__ == "__main__":
    main()
