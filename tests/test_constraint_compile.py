"""Host-side checks of compiled constraint programs (no device): the HIP
source generated for a program compiles for gfx950 with hiprtc, and carries
the program's constants as immediates."""
import numpy as np

import oracle as O

import ziros_amd as Z

P = O.GL_P
I = O.instr


def test_generated_source_compiles():
    bc = [I(O.OP_LOAD, 0), I(O.OP_DUP), I(O.OP_MUL), I(O.OP_CONST, 0), I(O.OP_SUB), I(O.OP_EMIT, 1),
          I(O.OP_LOAD, 17), I(O.OP_LOAD, 3), I(O.OP_ADD), I(O.OP_EMIT, 0)]
    consts = [P - 1]
    src = Z.constraint_source(20, bc, consts, 3)
    assert "0x%xull" % (P - 1) in src and "zkg_ce_jit" in src
    Z.compile_check_constraints(20, bc, consts, 3)


def test_invalid_program_is_not_compiled():
    try:
        Z.compile_check_constraints(4, [I(O.OP_ADD)], [], 1)
    except Z.ZkgError:
        return
    raise AssertionError("underflowing program compiled")
