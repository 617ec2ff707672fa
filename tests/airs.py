"""The example AIRs and trace generators of the constraint-evaluation tests
(p3gpu.example_airs), under the short name the tests use."""
from p3gpu.example_airs import *  # noqa: F401,F403
from p3gpu.example_airs import P, W, PERIODIC  # noqa: F401

import numpy as _np

import oracle as _O

# The cases of the reference's TinyAir tests (p3-miden-lifted-stark/tests/tiny_air.rs:
# single_trace, two_traces_same_height, two_traces_different_heights,
# three_traces_ascending_heights, single_periodic_column, periodic_column_period_4,
# multiple_periodic_columns, periodic_columns_multi_trace_same_height,
# periodic_columns_multi_trace_different_heights, periodic_columns_three_traces):
# (periods, trace heights); instance idx starts at idx + 2 (tiny_air.rs:152-156).
TINY_CASES = [([], [8]), ([], [8, 8]), ([], [4, 8]), ([], [4, 8, 16]), ([2], [8]), ([4], [8]),
              ([2, 4], [8]), ([2], [8, 8]), ([2, 4], [4, 8]), ([2, 4], [4, 8, 16])]


def tiny_periodic_table(periods, n, log_d, shift):
    """PeriodicLde::build (periodic.rs:49-81) for TinyAir's columns on the quotient domain of
    a trace of height n on the coset `shift`: the columns repeated to the longest period
    P, then their coset LDE on shift^(n/P) over P * 2^log_d rows, natural order."""
    cols = tiny_periodic_columns(periods)
    Pm = max(periods)
    per = _np.array([[c[i % len(c)] for c in cols] for i in range(Pm)], dtype=_np.uint64)
    return _O.coset_lde_batch(per, log_d, pow(shift, n // Pm, P), bit_reverse=False)
