"""The non-finite error-estimate rule on the oracle (the GPU twin is
tests/test_gpu_nan_rule.py): a NaN error estimate is a rejection."""
from __future__ import annotations

from nan_rule_cases import global_norm_fails, percell_isolates


def test_global_norm_nan_is_rejected_until_maxiters(oracle):
    global_norm_fails(oracle)


def test_percell_nan_stays_in_its_cell(oracle):
    percell_isolates(oracle)
