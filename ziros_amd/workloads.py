"""Synthetic workload draws of the batched and served commit streams.

Both draws are seeded with numpy's default generator so that the bench and the
parity tests commit exactly the same shapes.  Trace contents come from the
library's own element stream (zkg_fill_random, seed + k for trace k).

* Batched: 48 BabyBear traces, heights 2^U{12..20}, widths U{8, 16, ..., 256},
  coset LDE at blowup 2, Poseidon2-16 rows and Merkle tree.
* Served: 60% BabyBear / 40% Goldilocks requests, heights 2^U{14..22},
  widths U{16..128}, blowup 2 or 4, with unit-rate exponential gaps for a
  Poisson arrival trace.
"""
import numpy as np

SEED_BATCH = 0x5EED0003
SEED_SERVE = 0x5EED0005
GOLDILOCKS, BABYBEAR = 0, 1


def batch_shapes(n=48, seed=SEED_BATCH):
    """[(log2 height, width)] of the batched workload (BabyBear, blowup 2)."""
    rng = np.random.default_rng(seed)
    logs = rng.integers(12, 21, n)
    widths = rng.integers(1, 33, n) * 8
    return [(int(lg), int(wd)) for lg, wd in zip(logs, widths)]


def serve_requests(n, seed=SEED_SERVE):
    """([(field, log2 height, width, log_blowup)], unit exponential gaps)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        field = BABYBEAR if rng.random() < 0.6 else GOLDILOCKS
        lg = int(rng.integers(14, 23))
        w = int(rng.integers(16, 129))
        lb = int(rng.integers(1, 3))
        reqs.append((field, lg, w, lb))
    return reqs, rng.exponential(1.0, n)
