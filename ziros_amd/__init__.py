"""ziros_amd -- MI355X-native STARK trace commitment (coset LDE, Poseidon2 row
hashing, Merkle trees, FRI folding, the DEEP quotient of the PCS opening,
constraint evaluation)
behind the libzkg C ABI (include/zkg.h).

The compute path is hand-written HIP for gfx950 in ``csrc/``; this package
only binds it.  There is no CPU fallback: if ``libzkg.so`` is missing, every
entry point raises.
"""
from ._lib import ZkgError, load  # noqa: F401
from .host import (  # noqa: F401
    BABYBEAR,
    GENERATOR,
    GOLDILOCKS,
    BufferBridge,
    ConstraintCompiler,
    ConstraintProgram,
    Context,
    DeviceBuffer,
    FallbackSignal,
    FriCommitment,
    HipConstraintEval,
    HipDeep,
    HipDispatchDriver,
    HipDft,
    HipFri,
    HipMerkleTreeMmcs,
    HipPolyOps,
    HipPoseidon2,
    MerkleTree,
    Server,
    commit_batch,
    commit_trace,
    compile_check_constraints,
    constraint_source,
    make_job,
    translate_constraints,
)
