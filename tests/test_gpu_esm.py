"""GPU: the SuperFast mechanism ingested from its ESM reaction system
(tests/golden/superfast_g2_nodep.json) through rs_set_mechanism_stoich, on the
HIP library.

* Per-cell adaptive chemistry at tight tolerances on 128 cells (perturbed
  copies of each test's initial state plus the unperturbed one in cell 0):
  cell 0 meets the system's own assertions for all 6 tests (49 checks), and
  every cell agrees with the oracle's run of the same cell.
* The global-norm adaptive chemistry at the workloads' tolerances takes the
  oracle's accepted/rejected sequence and agrees to the chemistry bar.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from reseact_esm_amd import esm
from reseact_esm_amd import problems as P
from reseact_esm_amd.host import pictrl_ros23

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "superfast_g2_nodep.json")
NCELL = 128


@pytest.fixture(scope="module")
def sf():
    return esm.load_reaction_system(FIXTURE)


def _cells(m, u0, seed):
    rng = np.random.default_rng(seed)
    u = np.repeat(u0[:, None], NCELL, 1) * rng.lognormal(0.0, 0.2, (m.ns, NCELL))
    u[:, 0] = u0
    return np.ascontiguousarray(u.reshape(-1))


def _run(lib, m, u, times, rtol=1e-9, atol=1e-24):
    ctx = lib.context()
    ctx.set_grid(P.Grid(NCELL, 1, 1), m.ns, 0)
    ctx.set_mechanism_stoich(m)
    ctx.set_state(u)
    out, t, dt, att = {}, 0.0, np.full(NCELL, 1e-3), 0
    for tt in times:
        if tt > t:
            st = ctx.chem_percell(t, tt, 1e-3, dt, pictrl_ros23(), atol, rtol, True, maxiters=10**7)
            assert st.n_failed == 0
            att += st.attempts_total
            t = tt
        out[tt] = ctx.get_state().reshape(m.ns, NCELL)
    ctx.close()
    return out, att


def test_superfast_percell_assertions_and_oracle(sf, hip, oracle):
    worst = 0.0
    for ti, t in enumerate(sf.tests):
        m = esm.compile_system(sf, t.get("parameter_overrides"))
        u = _cells(m, esm.initial_state(sf, t.get("initial_conditions")), 100 + ti)
        times = sorted({float(a["time"]) for a in t["assertions"]})
        g, ag = _run(hip, m, u, times)
        o, ao = _run(oracle, m, u, times)
        print(f"{t['id']}: attempts gpu {ag} oracle {ao}", flush=True)
        # same step sequences up to the odd accept/reject flip at round-off
        assert abs(ag - ao) <= max(2, 0.002 * ao)
        for tt in times:
            rel = np.abs(g[tt] - o[tt]) / np.maximum(np.abs(o[tt]), 1e-12)
            worst = max(worst, float(rel.max()))
        for a in [x for x in esm.assertions(sf) if x.test == t["id"]]:
            got = float(g[a.time][m.species.index(a.variable), 0])
            assert esm.check(a.expected, got, a.rel, a.abs), (a, got)
    print(f"worst relative difference to the oracle: {worst:.2e}", flush=True)
    # far inside the assertions' 1e-4 band; the integration tolerance is 1e-9
    assert worst <= 1e-6


def test_superfast_global_norm_same_sequence(sf, hip, oracle):
    m = esm.compile_system(sf)
    u = _cells(m, esm.initial_state(sf), 7)
    res = []
    for lib in (hip, oracle):
        ctx = lib.context()
        ctx.set_grid(P.Grid(NCELL, 1, 1), m.ns, 0)
        ctx.set_mechanism_stoich(m)
        ctx.set_state(u)
        dt, k = ctx.adaptive_chem(0.0, 360.0, 1e-3, pictrl_ros23(), 1e-6, 1e-3, True)
        res.append((dt, k.naccept, k.nreject, ctx.get_state()))
        ctx.close()
    (dg, ag, rg, ug), (do, ao, ro, uo) = res
    assert (ag, rg) == (ao, ro) and ag > 5
    assert abs(dg - do) <= 1e-8 * do
    floor = np.maximum(np.abs(uo), 1e-6)
    assert float(np.max(np.abs(ug - uo) / floor)) <= 1e-8


def test_superfast_with_deposition_sinks(sf, hip, oracle):
    """SuperFast with first-order deposition sinks (compile_system(sinks=...),
    the rates of prototypes/g2_depsink/g2_run.jl:32-39) on 128 cells over one
    hour: per-cell adaptive chemistry agrees with the oracle."""
    dep = {"O3": 0.05, "NO2": 0.03, "HNO3": 0.1, "H2O2": 0.07, "CH2O": 0.04}
    m = esm.compile_system(sf, sinks=dep)
    u = _cells(m, esm.initial_state(sf), 21)
    g, ag = _run(hip, m, u, [3600.0])
    o, ao = _run(oracle, m, u, [3600.0])
    print(f"with sinks: attempts gpu {ag} oracle {ao}", flush=True)
    assert abs(ag - ao) <= max(2, 0.002 * ao)
    rel = np.abs(g[3600.0] - o[3600.0]) / np.maximum(np.abs(o[3600.0]), 1e-12)
    assert float(rel.max()) <= 1e-6
