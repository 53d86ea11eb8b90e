"""The mechanism compiler's emitted code, built for the CPU, against the
oracle on identical inputs: with the natural elimination order and exact
division it reproduces the reference's dense pivot-free elimination
(rx_traced_integrator.jl:110-135) bit for bit; with the Markowitz order it
agrees to round-off; its adjoint agrees with the oracle's dense VJP."""
from __future__ import annotations

import numpy as np
import pytest

from cpu_chem.hostchem import HostChem
from reseact_esm_amd import problems as P


def cases():
    c1 = P.config1(nlon=6, nlat=4, nlev=3)
    c4 = P.config4(nlon=5, nlat=4, nlev=3)
    c3 = P.config3(nlon=4, nlat=3, nlev=2)
    c1.emis = P.lognormal_emissions(c1.grid, c1.ns, 9)
    c4.jfac = np.random.default_rng(2).uniform(0, 1, c4.nc)
    return {"c1": c1, "c4": c4, "c3": c3}


CASES = cases()


def relmax(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("name", ["c1", "c4", "c3"])
def test_natural_order_is_bitwise_reference_elimination(oracle, name):
    p = CASES[name]
    ctx = oracle.context().load(p)
    hc = HostChem(p.mech, ordering=0, exact_div=True)
    for dt in (0.5, 40.0):
        eo = ctx.ros23_step(0.0, dt, 1e-6, 1e-3)
        un, e = hc.attempt(p, p.u0, dt, 1e-6, 1e-3)
        assert np.array_equal(un, ctx.get_candidate())
        assert abs(e - eo) <= 1e-15 * eo


# step sizes on the scale the controller accepts for each mechanism (the
# 50-species one is stiff up to 1e4 1/s: its accepted steps start near 1e-3 s)
DTS = {"c1": (0.5, 40.0), "c4": (0.5, 40.0), "c3": (1e-3, 0.05)}


@pytest.mark.parametrize("name", ["c1", "c4", "c3"])
def test_markowitz_order_agrees(oracle, name):
    p = CASES[name]
    ctx = oracle.context().load(p)
    hc = HostChem(p.mech, ordering=1, exact_div=False)
    for dt in DTS[name]:
        eo = ctx.ros23_step(0.0, dt, 1e-6, 1e-3)
        un, e = hc.attempt(p, p.u0, dt, 1e-6, 1e-3)
        # a different elimination order rounds differently, by up to cond(W)*eps
        assert relmax(un, ctx.get_candidate()) <= 1e-9
        assert abs(e - eo) <= 1e-9 * eo


@pytest.mark.parametrize("name", ["c1", "c4"])
def test_generated_vjp_matches_oracle(oracle, name):
    p = CASES[name]
    ctx = oracle.context().load(p)
    lam = np.random.default_rng(5).standard_normal(p.ns * p.nc)
    ctx.set_lambda(lam)
    ctx.zero_grad()
    ctx.ros23_vjp(0.0, 20.0)
    lo, go = ctx.get_lambda(), ctx.get_grad_emis()
    hc = HostChem(p.mech, ordering=1)
    lh, gh = hc.vjp(p, p.u0, 20.0, lam)
    assert np.linalg.norm(lh - lo) <= 1e-12 * np.linalg.norm(lo)
    if p.emis is not None:
        assert np.linalg.norm(gh - go) <= 1e-12 * np.linalg.norm(go)


def test_sparse_plan_is_cheaper_than_dense():
    p = CASES["c3"]
    nat = HostChem(p.mech, ordering=0).info
    mk = HostChem(p.mech, ordering=1).info
    assert mk["nnz_LU"] < nat["nnz_LU"] <= p.ns * p.ns
    assert mk["flops_attempt"] < nat["flops_attempt"]
    assert mk["nnz_J"] == nat["nnz_J"]


_SANITIZED_RUN = r"""
import sys, numpy as np
sys.path[:0] = sys.argv[1:3]
from cpu_chem.hostchem import HostChem
from oracle_lib import oracle_library
from test_mechgen_cpu import CASES
for name in ("c1", "c4"):
    p = CASES[name]
    ctx = oracle_library().context().load(p)
    hc = HostChem(p.mech, ordering=0, exact_div=True, sanitize=True)
    for dt in (5.0, 0.5, 5.0, 40.0):
        eo = ctx.ros23_step(0.0, dt, 1e-6, 1e-3)
        un, e, bad = hc.attempt_wave(p, p.u0, dt, 1e-6, 1e-3)
        assert bad == 0, (name, dt, bad)
        assert np.array_equal(un, ctx.get_candidate()), (name, dt)
        assert abs(e - eo) <= 1e-15 * eo
print("clean")
"""


def test_exact_build_wave_slab_clean_under_sanitizers():
    """The exact build's attempt with the device's [3][NS][64] slab layout,
    built for the CPU with AddressSanitizer and UndefinedBehaviorSanitizer,
    launched repeatedly with changing steps: no out-of-bounds or undefined
    operation, no write into another lane's slab slots, and every launch
    bit-for-bit the oracle's natural-order elimination."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                          text=True, check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", _SANITIZED_RUN, here, os.path.dirname(here)],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "clean" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
