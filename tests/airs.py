"""The example AIRs and trace generators of the constraint-evaluation tests
(p3gpu.example_airs), under the short name the tests use."""
from p3gpu.example_airs import *  # noqa: F401,F403
from p3gpu.example_airs import P, W, PERIODIC  # noqa: F401
