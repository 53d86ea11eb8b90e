"""The host-side operator API (reseact_esm_amd.operators), mirroring
tools/reactant_handoff/op_split.jl and rx_traced_integrator.jl.  CPU only: a
recording stand-in context for the driver logic, the oracle for numbers.

* Stop grid and forcing refresh (op_split.jl:122-156): with the GEOS-FP
  cadences (I3 at 10800k, A3 at 5400+10800k, A1 at 1800+3600k) a 9 h run
  refreshes at the same 12 boundaries for every macro_dt (300, 250, 900,
  10800), whether one call covers the window or the caller drives one macro
  step at a time; each boundary fires once, a boundary at t0 never.
* refresh(t) runs between the interval that ends at t and the next one, and
  the macro-boundary clamp acts after each interval: both bit-identical to
  driving rs_macro_step by hand.
* Macro-step bisection (op_split.jl:219-242): a window that runs out of inner
  iterations is redone as two halves, equal bit for bit to running the halves
  directly; recursion stops at min_macro_dt and the failure propagates.
* The single-attempt and adaptive wrappers agree with the context calls.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from reseact_esm_amd import operators as O
from reseact_esm_amd import problems as P
from reseact_esm_amd.host import (RS_E_ARG, RS_E_MAXITERS, ReseactError, SplitOpts, pictrl_ros23,
                                  pictrl_ssprk43)

H = 3600.0


def geosfp_tstops(t0, t1):
    """Forcing cadence boundaries of the three GEOS-FP collections."""
    out = set()
    for off, per in ((0.0, 3 * H), (1.5 * H, 3 * H), (0.5 * H, H)):
        k = 0
        while off + k * per <= t1:
            if off + k * per >= t0:
                out.add(off + k * per)
            k += 1
    return sorted(out)


class Recorder:
    """Stands in for a Context: records macro intervals and clamps; fails the
    intervals `fail(t0, t1)` names with RS_E_MAXITERS."""

    def __init__(self, fail=None):
        self.u = None
        self.log = []
        self.fail = fail

    def set_state(self, u):
        self.u = np.array(u, dtype=np.float64)

    def get_state(self):
        return self.u.copy()

    def clamp_state(self):
        self.log.append(("clamp",))

    def macro_step(self, t0, t1, dtT, dtC, opts):
        if self.fail is not None and self.fail(t0, t1):
            raise ReseactError(RS_E_MAXITERS, f"maxiters in ({t0}, {t1})")
        self.log.append(("step", t0, t1))
        self.u = self.u + (t1 - t0)
        c = SimpleNamespace(naccept=1, nreject=0)
        return dtT, dtC, c, c


@pytest.mark.parametrize("macro_dt", [300.0, 250.0, 900.0, 10800.0])
def test_refresh_fires_at_every_boundary(macro_dt):
    t0, t1 = 0.0, 9 * H
    fst = geosfp_tstops(t0, t1)
    expect = [t for t in fst if t0 < t <= t1]
    assert len(expect) == 12
    # one call over the whole window
    fired = []
    ctx = Recorder()
    r = O.lie_trotter_solve(ctx, np.zeros(1), (t0, t1), macro_dt, refresh=fired.append,
                            forcing_tstops=fst)
    assert fired == expect and r["retcode"] == "Success"
    # the caller drives one macro step at a time (tools/scaling_study.jl:127)
    fired2, t = [], t0
    while t < t1 - 1e-9:
        tn = min(t + macro_dt, t1)
        O.lie_trotter_solve(ctx, np.zeros(1), (t, tn), tn - t, refresh=fired2.append,
                            forcing_tstops=fst)
        t = tn
    assert fired2 == expect


def test_stop_grid_contents():
    stops, fst = O.forcing_stops(0.0, 3600.0, 600.0, [0.0, 1800.0, 2000.0000004, 3600.0, 3600.5])
    assert fst == {1800.0, 2000.0, 3600.0}
    assert stops == [600.0, 1200.0, 1800.0, 2000.0, 2400.0, 3000.0, 3600.0]
    # a lattice that does not divide the window ends on the window end
    assert O.forcing_stops(0.0, 1000.0, 300.0)[0] == [300.0, 600.0, 900.0, 1000.0]
    # the lattice stops 1e-9 short of tf, so no sliver interval
    assert O.forcing_stops(0.0, 600.0 + 5e-10, 300.0)[0] == [300.0, 600.0 + 5e-10]
    with pytest.raises(ValueError):
        O.forcing_stops(0.0, 1.0, 0.0)


def test_intervals_clamps_and_refresh_order():
    ctx = Recorder()
    log = ctx.log
    O.lie_trotter_solve(ctx, np.zeros(1), (0.0, 1000.0), 400.0, forcing_tstops=[500.0],
                        refresh=lambda t: log.append(("refresh", t)))
    assert log == [("step", 0.0, 400.0), ("clamp",), ("step", 400.0, 500.0), ("clamp",),
                   ("refresh", 500.0), ("step", 500.0, 800.0), ("clamp",),
                   ("step", 800.0, 1000.0), ("clamp",)]
    log.clear()
    O.lie_trotter_solve(ctx, np.zeros(1), (0.0, 1000.0), 400.0, clamp_nonneg=False)
    assert ("clamp",) not in log and len(log) == 3


@pytest.fixture(scope="module")
def small(oracle):
    """6x4x2 grid, 10 species, 4 tracers, 20 first-order reactions: the
    transport's undershoots reach the chemistry (the clamp acts at macro
    boundaries only), which a linear mechanism takes without blowing up."""
    p = P.config1(nlon=6, nlat=4, nlev=2)
    rng = np.random.default_rng(5)
    a = rng.integers(0, p.ns, 20)
    b = (a + 1 + rng.integers(0, p.ns - 1, 20)) % p.ns
    none = -np.ones(20, np.int64)
    p.mech = P.Mechanism(ns=p.ns, reactants=np.stack([a, none], 1).astype(np.int32),
                         products=np.stack([b, none], 1).astype(np.int32),
                         k=10 ** rng.uniform(-4, 1, 20), photo=np.zeros(20, np.int32))
    ctx = oracle.context().load(p)
    yield p, ctx
    ctx.close()


def _opts(p, flags=0, maxiters=20000):
    return SplitOpts(pictrl_ssprk43(), pictrl_ros23(), p.atol_T, p.rtol_T, p.atol_C, p.rtol_C,
                     flags, 0, maxiters)


def test_refresh_and_clamp_match_manual_steps(small):
    p, ctx = small
    emis = P.lognormal_emissions(p.grid, p.ns, 9)

    def set_emis(t):
        ctx.set_emissions(emis * (1.0 + t / 1800.0))

    set_emis(0.0)
    r = O.lie_trotter_solve(ctx, p.u0, (0.0, 1800.0), 600.0, refresh=set_emis,
                            forcing_tstops=[600.0, 1200.0, 1800.0])
    set_emis(0.0)
    ctx.set_state(p.u0)
    dtT, dtC = 15.0, 0.5
    for a, b in ((0.0, 600.0), (600.0, 1200.0), (1200.0, 1800.0)):
        dtT, dtC, _, _ = ctx.macro_step(a, b, dtT, dtC, _opts(p))
        ctx.clamp_state()
        set_emis(b)
    assert r["retcode"] == "Success" and r["nmacro"] == 3
    assert np.array_equal(r["u"], ctx.get_state())
    ctx.set_emissions(None)


def test_bisection_equals_halves(small):
    p, ctx = small
    # 600 s needs 93 chemistry attempts, each 300 s half at most 90
    r = O.lie_trotter_solve(ctx, p.u0, (0.0, 600.0), 600.0, maxiters=90)
    assert r["retcode"] == "MaxIters" and r["nmacro"] == 0
    calls = []
    rb = O.lie_trotter_solve_bisect(ctx, p.u0, (0.0, 600.0), 600.0, maxiters=90,
                                    on_bisect=lambda a, b: calls.append((a, b)))
    assert rb["retcode"] == "Success" and calls == [(0.0, 600.0)] and rb["nmacro"] == 2
    r1 = O.lie_trotter_solve(ctx, p.u0, (0.0, 300.0), 300.0, maxiters=90)
    r2 = O.lie_trotter_solve(ctx, r1["u"], (300.0, 600.0), 300.0, maxiters=90)
    assert np.array_equal(rb["u"], r2["u"])
    for k in ("naT", "nrT", "naC", "nrC"):
        assert rb[k] == r1[k] + r2[k]


def test_bisection_stops_at_min_macro_dt():
    # every interval longer than 100 s that contains t = 250 fails
    ctx = Recorder(fail=lambda a, b: b - a > 100.0 + 1e-9 and a < 250.0 < b)
    calls = []
    r = O.lie_trotter_solve_bisect(ctx, np.zeros(1), (0.0, 600.0), 600.0, min_macro_dt=9.0,
                                   on_bisect=lambda a, b: calls.append((a, b)))
    assert r["retcode"] == "Success"
    assert calls == [(0.0, 600.0), (0.0, 300.0), (150.0, 300.0)]
    steps = [e[1:] for e in ctx.log if e[0] == "step"]
    assert steps == [(0.0, 150.0), (150.0, 225.0), (225.0, 300.0), (300.0, 600.0)]
    assert r["nmacro"] == 4 and r["u"][0] == 600.0
    # with a floor of 200 s the 150 s halves may not split: the failure propagates
    ctx = Recorder(fail=lambda a, b: b - a > 100.0 + 1e-9 and a < 250.0 < b)
    calls.clear()
    r = O.lie_trotter_solve_bisect(ctx, np.zeros(1), (0.0, 600.0), 600.0, min_macro_dt=200.0,
                                   on_bisect=lambda a, b: calls.append((a, b)))
    assert r["retcode"] == "MaxIters" and calls == [(0.0, 600.0), (0.0, 300.0)]


def test_other_errors_raise():
    class BadArg(Recorder):
        def macro_step(self, *a):
            raise ReseactError(RS_E_ARG, "bad argument")

    with pytest.raises(ReseactError) as e:
        O.lie_trotter_solve(BadArg(), np.zeros(1), (0.0, 600.0), 600.0)
    assert e.value.code == RS_E_ARG


def test_step_wrappers(small):
    p, ctx = small
    u1, e1 = O.ros23_step(ctx, p.u0, 0.0, 30.0, 1e-6, 1e-3)
    ctx.set_state(p.u0)
    e2 = ctx.ros23_step(0.0, 30.0, 1e-6, 1e-3)
    assert e1 == e2 and np.array_equal(u1, ctx.get_candidate())
    u1, e1 = O.ssprk43_step(ctx, p.u0, 0.0, 30.0, 1e-6, 1e-3)
    ctx.set_state(p.u0)
    e2 = ctx.ssprk43_step(0.0, 30.0, 1e-6, 1e-3)
    assert e1 == e2 and np.array_equal(u1, ctx.get_candidate())
    u, t, dt, na, nr = O.adaptive_solve(O.ros23_step, ctx, p.u0, 0.0, 600.0, 0.5, pictrl_ros23())
    assert abs(t - 600.0) < 1e-9 and na > 0
    with pytest.raises(ValueError):
        O.adaptive_solve(len, ctx, p.u0, 0.0, 1.0, 0.5, pictrl_ros23())
