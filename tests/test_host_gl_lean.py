"""The case analysis of the device's lean Goldilocks products (gl_asm in
ziros_amd/csrc/zkg_field.h), modelled instruction by instruction in
tests/host/gl_lean_steps.cpp and checked against exact 128-bit arithmetic:
which carry or borrow is corrected, with which constant, and that no
correction wraps.  The device instructions themselves are checked by the GPU
parity tests (every Goldilocks hash, tree and opening against the oracle)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_lean_goldilocks_product_steps(tmp_path):
    exe = str(tmp_path / "gl_lean_steps")
    b = subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "host", "gl_lean_steps.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0; wraps 0;" in r.stdout
