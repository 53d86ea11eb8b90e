"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle): the inputs are regenerated from
their seeds and must match the stored ones; the oracle must reproduce the
stored outputs bit for bit; the HIP library must match them to tolerance."""
from __future__ import annotations

import os

import numpy as np
import pytest

from golden.make_golden import NAMES, compute

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(oracle, name):
    ref = load(name)
    got = compute(oracle, name)
    assert set(ref) == set(got)
    for k in ref:
        assert np.array_equal(ref[k], got[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_golden(hip, name):
    ref = load(name)
    got = compute(hip, name)
    tol = 1e-12 if name.startswith("transport") else 1e-8
    for k in ref:
        if k == "counts":
            assert np.array_equal(ref[k], got[k])
            continue
        a, b = got[k], ref[k]
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) if k != "eest" else \
            abs(float(a) - float(b)) / max(abs(float(b)), 1.0)
        assert rel <= tol, (k, rel)
