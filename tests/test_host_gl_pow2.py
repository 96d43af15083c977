"""The Goldilocks power-of-two product (Gl::mul_pow2 in
ziros_amd/csrc/zkg_field.h), compiled for the host by tests/host/gl_pow2_check.cpp:
x 2^k for every k < 96 against exact 128-bit arithmetic over edge and random
canonical inputs, and the order-4, -8 and -16 roots of p3's Goldilocks two-adic
generator (0x185629dcda58878c squared down) equal to 2^48, 2^120 and 2^156, the
exponents the transforms' register sub-DFTs use in place of table twiddles."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_goldilocks_pow2_products_and_roots(tmp_path):
    exe = str(tmp_path / "gl_pow2_check")
    b = subprocess.run([HIPCC, "-O2", "-std=c++17", os.path.join(ROOT, "tests", "host", "gl_pow2_check.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "roots mismatches 0;" in r.stdout and "mismatches 0\n" in r.stdout
