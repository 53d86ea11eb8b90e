"""GPU: the non-finite error-estimate rule (rx_traced_integrator.jl:484), as
on the oracle (tests/test_nan_rule.py), and the per-cell results of the finite
cells against the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from nan_rule_cases import global_norm_fails, percell_isolates

pytestmark = pytest.mark.gpu


def test_global_norm_nan_is_rejected_until_maxiters(hip):
    print("nan rule: global norm", flush=True)
    global_norm_fails(hip)


def test_percell_nan_stays_in_its_cell(hip, oracle):
    print("nan rule: per cell", flush=True)
    ug = percell_isolates(hip)
    uo = percell_isolates(oracle)
    # chemistry parity: relative to max(|u|, atol, 1e-4 x the cell's largest species)
    floor = np.maximum(1e-6, 1e-4 * np.max(np.abs(uo), axis=0, keepdims=True))
    rel = np.abs(ug - uo) / np.maximum(np.abs(uo), floor)
    assert float(rel.max()) <= 1e-8
