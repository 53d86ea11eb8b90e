"""Mechanism ingestion from an ESM reaction system: the SuperFast mechanism
(12 active species, 3 constant reservoirs, 27 reactions with fractional and
multiple products, second-order self-reactions and catalysts;
prototypes/g2_depsink/g2_nodep.esm, fixture tests/golden/superfast_g2_nodep.json)
checked against the system's own 49 assertions (6 tests; expected values at
times, with the tests' tolerances).  CPU only.

* The ingestion (expression evaluation, constant folding, rate law) meets the
  assertions with an independent stiff integrator (scipy Radau at rtol 1e-12).
* The oracle, given the mechanism through rs_set_mechanism_stoich, meets them
  with its per-cell adaptive Rosenbrock23 at tight tolerances.
* The generated chemistry code (CPU build) reproduces the oracle's attempt:
  bit-for-bit in natural order with true division, to 1e-9 in Markowitz order.
"""
from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from reseact_esm_amd import esm
from reseact_esm_amd import problems as P
from reseact_esm_amd.host import pictrl_ros23

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "superfast_g2_nodep.json")


@pytest.fixture(scope="module")
def sf():
    return esm.load_reaction_system(FIXTURE)


def test_fixture_shape(sf):
    assert len(sf.species) == 15 and len(sf.active) == 12 and len(sf.reactions) == 27
    assert len(esm.assertions(sf)) == 49
    m = esm.compile_system(sf)
    assert m.ns == 12 and m.nr == 27
    # the constant reservoirs are folded into the rate constants, not integrated
    assert not {"O2", "CH4", "H2O"} & set(m.species)
    # fractional products (R21: ISOP + O3 -> 0.05 CO + 0.06 HO2 + 0.87 CH2O + 1.86 CH3O2)
    r21 = m.ids.index("R21")
    coef = {m.species[m.prod_species[q]]: m.prod_coef[q]
            for q in range(m.prod_off[r21], m.prod_off[r21 + 1])}
    assert coef == {"CO": 0.05, "HO2": 0.06, "CH2O": 0.87, "CH3O2": 1.86}
    # second-order self-reaction 2 HO2 -> H2O2 + O2: both reactant slots HO2
    r4 = m.ids.index("R4")
    assert list(m.reactants[r4]) == [m.species.index("HO2")] * 2
    # catalyst R19: ISOP + OH -> ISOP consumes OH only
    N = m.net()
    r19 = m.ids.index("R19")
    assert N[m.species.index("ISOP"), r19] == 0.0 and N[m.species.index("OH"), r19] == -1.0


def _scipy_run(m, u0, times, t1):
    N = m.net()

    def f(_t, u):
        rate = m.k.copy()
        for r in range(m.nr):
            for a in m.reactants[r]:
                if a >= 0:
                    rate[r] *= u[a]
        return N @ rate

    sol = solve_ivp(f, (0.0, t1), u0, method="Radau", rtol=1e-12, atol=1e-30, dense_output=True)
    return {t: (u0 if t == 0 else sol.sol(t)) for t in times}


def _check_all(sf, run):
    """run(mechanism, u0, sorted times, t_end) -> {time: state}."""
    bad = []
    for t in sf.tests:
        m = esm.compile_system(sf, t.get("parameter_overrides"))
        u0 = esm.initial_state(sf, t.get("initial_conditions"))
        times = sorted({float(a["time"]) for a in t["assertions"]})
        states = run(m, u0, times, float(t["time_span"]["end"]))
        for a in [x for x in esm.assertions(sf) if x.test == t["id"]]:
            got = float(states[a.time][m.species.index(a.variable)])
            if not esm.check(a.expected, got, a.rel, a.abs):
                bad.append((a.test, a.variable, a.time, got, a.expected))
    return bad


def test_ingestion_meets_assertions(sf):
    assert _check_all(sf, _scipy_run) == []


def _library_run(lib, rtol=1e-9, atol=1e-24):
    def run(m, u0, times, t1):
        ctx = lib.context()
        ctx.set_grid(P.Grid(1, 1, 1), m.ns, 0)
        ctx.set_mechanism_stoich(m)
        ctx.set_state(u0)
        out, t, dt = {}, 0.0, np.array([1e-3])
        for tt in times:
            if tt > t:
                st = ctx.chem_percell(t, tt, 1e-3, dt, pictrl_ros23(), atol, rtol, True,
                                      maxiters=10**7)
                assert st.n_failed == 0
                t = tt
            out[tt] = ctx.get_state()
        ctx.close()
        return out
    return run


def test_oracle_meets_assertions(sf, oracle):
    assert _check_all(sf, _library_run(oracle)) == []


@pytest.mark.parametrize("ordering,exact", [(0, True), (1, False)])
def test_generated_code_matches_oracle(sf, oracle, ordering, exact):
    from cpu_chem.hostchem import HostChem
    m = esm.compile_system(sf)
    hc = HostChem(m, ordering=ordering, exact_div=exact)
    ncell = 16
    rng = np.random.default_rng(3)
    u0 = esm.initial_state(sf)
    u = np.repeat(u0[:, None], ncell, 1) * rng.lognormal(0.0, 0.3, (m.ns, ncell))
    u = u.reshape(-1)
    grid = P.Grid(ncell, 1, 1)
    box = SimpleNamespace(grid=grid, jfac=None, emis=None, ns=m.ns)
    ctx = oracle.context()
    ctx.set_grid(grid, m.ns, 0)
    ctx.set_mechanism_stoich(m)
    for dt in (1e-3, 0.5, 20.0):
        ctx.set_state(u)
        eo = ctx.ros23_step(0.0, dt, 1e-6, 1e-3)
        uo = ctx.get_candidate()
        ug, eg = hc.attempt(box, u, dt, 1e-6, 1e-3)
        if exact:
            assert np.array_equal(ug, uo), dt
        else:
            assert float(np.max(np.abs(ug - uo) / np.maximum(np.abs(uo), 1e-30))) <= 1e-9, dt
        assert abs(eg - eo) <= 1e-9 * max(eo, 1.0)
    ctx.close()


def test_unsupported_reactions_raise(sf):
    import copy
    bad = copy.deepcopy(sf)
    bad.reactions[0].substrates = [("O3", 3.0)]
    with pytest.raises(ValueError):
        esm.compile_system(bad)
    bad = copy.deepcopy(sf)
    bad.reactions[0].rate = {"op": "*", "args": ["OH", 2.0]}   # an active species inside a rate
    with pytest.raises(KeyError):
        esm.compile_system(bad)


# the exaggerated deposition rates of prototypes/g2_depsink/g2_run.jl:32-39
# (k_wet on the soluble species is the caller's choice: the sink model that
# assigns it is not in the reference tree)
DEP = {"O3": 0.05, "NO2": 0.03, "HNO3": 0.08 + 0.02, "H2O2": 0.05 + 0.02, "CH2O": 0.04}


def test_deposition_sinks(sf, oracle):
    """First-order sinks from compile_system(sinks=...): alone they decay as
    exp(-k t); on SuperFast over 1 s (g2_run.jl) every species with a sink
    ends lower than without, and the NOy total NO + NO2 + HNO3 loses what
    the NO2 and HNO3 sinks remove."""
    m0 = esm.compile_system(sf)
    md = esm.compile_system(sf, sinks=DEP)
    assert md.nr == m0.nr + len(DEP) and md.ids[-1] == "sink_CH2O"
    with pytest.raises(KeyError):
        esm.compile_system(sf, sinks={"N2O5": 1.0})
    # sinks alone: zero every chemical rate constant
    import dataclasses
    only = dataclasses.replace(md, k=np.concatenate([np.zeros(m0.nr), md.k[m0.nr:]]))
    u0 = esm.initial_state(sf)
    run = _library_run(oracle, rtol=1e-10, atol=1e-30)
    got = run(only, u0, [1.0], 1.0)[1.0]
    for sp, kd in DEP.items():
        i = md.species.index(sp)
        assert abs(got[i] - u0[i] * np.exp(-kd)) <= 1e-8 * u0[i], sp
    a = run(m0, u0, [1.0], 1.0)[1.0]
    b = run(md, u0, [1.0], 1.0)[1.0]
    for sp in DEP:
        i = md.species.index(sp)
        assert b[i] < a[i], sp
    noy = lambda u: sum(u[md.species.index(x)] for x in ("NO", "NO2", "HNO3"))
    assert noy(b) < noy(a) <= noy(u0) * (1 + 1e-9)
