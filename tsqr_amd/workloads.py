"""Shapes of the batched workload: many tall-skinny matrices of uneven size.

batch_shapes(count, m_lo, m_hi, widths, seed, multiple) draws m log-uniform
in [m_lo, m_hi] (rounded down to a multiple of `multiple`, the reference's
divisibility constraint n | m for every width used) and n uniformly from
`widths`, from a seeded numpy generator.  The default arguments are the
512-matrix batch: m in [4096, 262144], n in {8, 16, 32, 64}, m a multiple
of 64.
"""
from __future__ import annotations

import numpy as np


def batch_shapes(count: int = 512, m_lo: int = 4096, m_hi: int = 262144, widths=(8, 16, 32, 64), seed: int = 0,
                 multiple: int = 64):
    rng = np.random.Generator(np.random.PCG64(seed))
    logm = rng.uniform(np.log(m_lo), np.log(m_hi), size=count)
    ms = np.maximum((np.exp(logm).astype(np.int64) // multiple) * multiple, max(m_lo // multiple, 1) * multiple)
    ns = rng.choice(np.asarray(widths), size=count)
    return [(int(m), int(n)) for m, n in zip(ms, ns)]
