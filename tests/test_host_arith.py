"""The device's exact-FP64 BabyBear arithmetic, compiled for the host.

tests/host/p2_fp64_check.cpp includes the device headers (Bf in
ziros_amd/csrc/zkg_field.h, p2_permute_bf in zkg_poseidon2.h) and checks:
- the multiplication by +-2^-k mod p of the internal rounds (every shift and
  sign the diagonal of zkf-metal/src/poseidon2/babybear.rs:14-58 uses) against
  exact integer arithmetic, over edge values and random integers below 2^41;
- the FP64 product against __int128 arithmetic;
- the whole width-16 permutation against the oracle's integer Poseidon2 over
  random, near-p and 0/p-1 states, and over a second absorption on the
  unreduced capacity.
"""
import os
import subprocess

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_fp64_babybear_arithmetic_and_permutation(tmp_path):
    O.lib()  # builds oracle/liboracle.so when missing
    odir = os.path.join(ROOT, "oracle")
    exe = str(tmp_path / "p2_fp64_check")
    cmd = [HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(ROOT, "tests", "host", "p2_fp64_check.cpp"),
           "-L", odir, "-loracle", "-Wl,-rpath," + odir, "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "inv_pow2 mismatches 0; mul mismatches 0; permutation mismatches 0" in r.stdout
