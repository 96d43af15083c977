"""Builds libzkg.so (HIP for gfx950) in-tree.

Run as ``python -m ziros_amd.build`` or through ``__graft_entry__.build()``.
Every translation unit is compiled with hipcc for gfx950 only; the packed
FP32 instructions are disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libzkg.so")
PROBE_LIB = os.path.join(HERE, "libzkg_probe.so")
# (source, object name, extra defines): the column transforms are compiled once
# per tile plan (zkg_plans.h), each into its own namespace.
LDE_PLANS = 4
UNITS = [("zkg_ntt.hip", "zkg_ntt_p%d.o" % k, ["-DZKG_LDE_VARIANT=%d" % k]) for k in range(LDE_PLANS)] + [
    (src, src.replace(".hip", ".o"), [])
    for src in ("zkg_hash.hip", "zkg_deep.hip", "zkg_constraints.hip", "zkg_runtime.hip", "zkg_batch.hip",
                "zkg_driver.hip", "zkg_noalloc.hip", "zkg_poly.hip", "zkg_tuning.hip")]
PROBE_UNITS = [("zkg_probe.hip", "zkg_probe.o", [])]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


# FP64 code that relies on exact products and explicit fma must not be
# contracted or reassociated by the compiler.
EXTRA = {"zkg_probe.hip": ["-ffp-contract=off"], "zkg_hash.hip": ["-ffp-contract=off"],
         "zkg_deep.hip": ["-ffp-contract=off"]}


JIT_PRELUDE = os.path.join(CSRC, "_gen", "zkg_jit_prelude.inc")
JIT_HEADERS = ["zkg_field.h", "zkg_ce_device.h"]


def write_jit_prelude():
    """The source that constraint programs compiled at run time start with:
    the field arithmetic and evaluator helpers, as one C++ raw string."""
    parts = ["typedef unsigned int uint32_t;", "typedef unsigned long long uint64_t;",
             "typedef int int32_t;", "typedef long long int64_t;"]
    for h in JIT_HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            parts += [l for l in f.read().splitlines() if not l.startswith(("#include", "#pragma once"))]
    text = "\n".join(parts)
    assert ")ZKGJIT" not in text
    os.makedirs(os.path.dirname(JIT_PRELUDE), exist_ok=True)
    body = 'R"ZKGJIT(' + text + '\n)ZKGJIT"\n'
    if not os.path.exists(JIT_PRELUDE) or open(JIT_PRELUDE).read() != body:
        with open(JIT_PRELUDE, "w") as f:
            f.write(body)


def _compile(unit, obj_dir=None):
    src, obj_name, defines = unit
    obj = os.path.join(obj_dir or os.path.join(CSRC, "_obj"), obj_name)
    cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), *defines, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s %s:\n%s" % (src, " ".join(defines), "\n".join(msgs)))
    return obj, [m for m in msgs if m.strip()]


def _link(objs, out):
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)


def build(verbose=False):
    """Compile every translation unit and link libzkg.so and libzkg_probe.so."""
    os.makedirs(os.path.join(CSRC, "_obj"), exist_ok=True)
    write_jit_prelude()
    # longest units first so the pool stays busy
    with ThreadPoolExecutor(max_workers=8) as ex:
        results = list(ex.map(_compile, UNITS + PROBE_UNITS))
    if verbose:
        for _, msgs in results:
            for m in msgs:
                print(m)
    objs = [o for o, _ in results]
    _link(objs[: len(UNITS)], LIB)
    _link(objs[len(UNITS):], PROBE_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
