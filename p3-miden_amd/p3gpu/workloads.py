"""Synthetic workload shapes shared by the benchmark and the tests.

Shapes are drawn from numpy's PCG64 with a fixed seed so a workload is the
same everywhere; values come from the library's generators (host SmallRng
stream, or the device counter generator for inputs too large to ship).
"""
import numpy as np

# single-trace proof: 2^22 x 64, log_blowup 2, FRI arity 4 down to degree 2^4, 100 queries
C2_LOG_ROWS, C2_WIDTH = 22, 64
C2_PARAMS = (2, 2, 4, 0, 0, 100, 0)

# single-trace proof on the CPU-sized case: 2^14 x 32, log_blowup 1, arity 2, degree 2^3, 64 queries
C1_LOG_ROWS, C1_WIDTH = 14, 32
C1_PARAMS = (1, 1, 3, 0, 0, 64, 0)

# wide bit-structured Keccak trace: 2^16 x 2600, log_blowup 1, arity 2 down to degree 2^3, 100 queries
C4_LOG_ROWS, C4_WIDTH = 16, 2600
C4_PARAMS = (1, 1, 3, 0, 0, 100, 0)
C4_BIT_FRACTION = 0.9


def c3_shapes(seed: int = 2025, k: int = 12, log_min: int = 12, log_max: int = 22, w_min: int = 8, w_max: int = 256):
    """12 traces, log height uniform in [12, 22], width uniform in [8, 256], sorted ascending by height."""
    rng = np.random.default_rng(seed)
    logs = rng.integers(log_min, log_max + 1, size=k)
    ws = rng.integers(w_min, w_max + 1, size=k)
    return sorted(((1 << int(l), int(w)) for l, w in zip(logs, ws)), key=lambda s: s[0])


def c3_params(log_blowup: int = 1):
    """Opening parameters used for the batched proof: arity 2 (matching C1), final degree 2^3, 100 queries."""
    return (log_blowup, 1, 3, 0, 0, 100, 0)


def c4_trace(seed: int = 2025, log_rows: int = C4_LOG_ROWS, width: int = C4_WIDTH,
             bit_fraction: float = C4_BIT_FRACTION) -> np.ndarray:
    """Columns are 0/1 bits with probability bit_fraction, else 16-bit limbs."""
    rng = np.random.default_rng(seed)
    is_bit = rng.random(width) < bit_fraction
    n = 1 << log_rows
    out = np.empty((n, width), np.uint64)
    nb = int(is_bit.sum())
    out[:, is_bit] = rng.integers(0, 2, size=(n, nb), dtype=np.uint64)
    out[:, ~is_bit] = rng.integers(0, 1 << 16, size=(n, width - nb), dtype=np.uint64)
    return out


def c5_request(rng: np.random.Generator):
    """One served prove request: 1-8 traces, log height in [14, 22], width in [16, 128], log_blowup in [1, 3]."""
    k = int(rng.integers(1, 9))
    logs = rng.integers(14, 23, size=k)
    ws = rng.integers(16, 129, size=k)
    lb = int(rng.integers(1, 4))
    shapes = sorted(((1 << int(l), int(w)) for l, w in zip(logs, ws)), key=lambda s: s[0])
    return shapes, lb
