"""The out-of-domain rejection of trial states (RS_POSITIVE_DOMAIN) on the
oracle, pinned by closed forms.  CPU only.

The reference wants negative Rosenbrock23 trial sub-steps rejected and retried
shorter, values untouched (HELPERS.md:396-447): an isoutofdomain rule with
failfactor 2 (a rejected trial of size h leaves h/2 as the next step size and
the controller's memory unchanged).

* The one-step stability function of ROS23 on u' = -k u,
  R(z) = 1 + (z/w)(2 - 1/w + z/(2w)), w = 1 - d z, d = 1/(2+sqrt 2)
  (from the method's two stages), matches the oracle's attempt; R < 0 for
  k h > ~2.4, so large steps of A -> B make A negative.
* With a tolerance so loose that the controller saturates (every accepted
  step grows tenfold), the accepted/rejected sequence follows from the sign
  of R alone: the oracle's counts and end state match that sawtooth exactly,
  with and without the rule; no accepted state goes negative with it, and
  A + B = 1 is kept (the clamp breaks it).
* Per cell: every cell follows the same sawtooth.
* The transport solve checks only the transported block; in the split
  options the rule reaches the chemistry only.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import (RS_CLAMP_NONNEG, RS_E_MAXITERS, RS_POSITIVE_DOMAIN,
                                  ReseactError, pictrl_ros23, pictrl_ssprk43, split_opts)

D = 1.0 / (2.0 + math.sqrt(2.0))


def R(z: float) -> float:
    w = 1.0 - D * z
    return 1.0 + (z / w) * (2.0 - 1.0 / w + z / (2.0 * w))


def decay_box(oracle, k=1.0, u0=(1.0, 0.0), ncell=1):
    m = P.Mechanism(ns=2, reactants=np.array([[0, -1]], np.int32),
                    products=np.array([[1, -1]], np.int32), k=np.array([k]),
                    photo=np.zeros(1, np.int32))
    g = P.Grid(ncell, 1, 1)
    u = np.zeros((2, ncell))
    u[0], u[1] = u0[0], u0[1]
    p = P.Problem(name="decay", grid=g, ns=2, nt=0, wind=P.solid_body_wind(g, 1.0, 0.0),
                  mech=m, u0=np.ascontiguousarray(u.reshape(-1)))
    return p, oracle.context().load(p)


def sawtooth(k, dt0, t1, posdom):
    """Saturated controller (dt x10 per accepted step), R's sign decides the
    domain rejections; returns (naccept, nreject, A(t1)/A0, min A over accepted)."""
    t, dt, a, acc, rej, amin = 0.0, dt0, 1.0, 0, 0, 1.0
    while t < t1 - 1e-9:
        h = min(dt, t1 - t)
        r = R(-k * h)
        if posdom and r < 0.0:
            dt, rej = h / 2.0, rej + 1
            continue
        a *= r
        amin = min(amin, a)
        t, dt, acc = t + h, 10.0 * h, acc + 1
    return acc, rej, a, amin


# a tolerance that scales every error estimate of states up to 1e3 below 1e-5,
# where the PI rule saturates at qmax
LOOSE = dict(atol=1e9, rtol=1e-3)


def test_stability_function_closed_form(oracle):
    p, ctx = decay_box(oracle, k=1.0)
    for h in (0.01, 0.7, 2.0, 2.5, 3.0, 10.0, 1e3):
        ctx.set_state(p.u0)
        ctx.ros23_step(0.0, h, 1e-6, 1e-3)
        a = ctx.get_candidate()[0]
        assert abs(a - R(-h)) <= 2e-15, (h, a, R(-h))
    assert R(-2.0) > 0.0 > R(-2.5)


@pytest.mark.parametrize("k,dt0,t1", [(1.0, 1.0, 100.0), (0.3, 0.5, 600.0), (5.0, 0.05, 60.0)])
def test_sawtooth_counts_and_state(oracle, k, dt0, t1):
    p, ctx = decay_box(oracle, k=k)
    # without the rule: the same saturated sequence goes negative
    acc0, rej0, a0, amin0 = sawtooth(k, dt0, t1, False)
    assert amin0 < 0.0 and rej0 == 0
    dt, c = ctx.adaptive_chem(0.0, t1, dt0, pictrl_ros23(), LOOSE["atol"], LOOSE["rtol"], 0)
    assert (c.naccept, c.nreject) == (acc0, 0)
    u = ctx.get_state()
    assert abs(u[0] - a0) <= 1e-9 * abs(a0) + 1e-300
    # the clamp zeroes A and breaks A + B = 1
    ctx.set_state(p.u0)
    ctx.adaptive_chem(0.0, t1, dt0, pictrl_ros23(), LOOSE["atol"], LOOSE["rtol"], RS_CLAMP_NONNEG)
    u = ctx.get_state()
    assert u[0] >= 0.0 and u[0] + u[1] > 1.0 + 1e-3
    # with the rule: halvings where R < 0, every accepted A >= 0, A + B = 1
    acc, rej, a, amin = sawtooth(k, dt0, t1, True)
    assert rej > 0 and amin >= 0.0
    for flags in (RS_POSITIVE_DOMAIN, RS_POSITIVE_DOMAIN | RS_CLAMP_NONNEG):
        ctx.set_state(p.u0)
        dt, c = ctx.adaptive_chem(0.0, t1, dt0, pictrl_ros23(), LOOSE["atol"], LOOSE["rtol"],
                                  flags)
        assert (c.naccept, c.nreject) == (acc, rej), flags
        u = ctx.get_state()
        # R near its root loses digits to cancellation: 1e-9 over the product
        assert u[0] >= 0.0 and abs(u[0] - a) <= 1e-9 * a + 1e-300
        assert abs(u[0] + u[1] - 1.0) <= 1e-14
    ctx.close()


def test_sawtooth_per_cell(oracle):
    k, dt0, t1, ncell = 1.0, 1.0, 100.0, 5
    scale = np.array([1.0, 1e-3, 7.0, 0.5, 1e3])
    p, ctx = decay_box(oracle, k=k, u0=(scale, np.zeros(ncell)), ncell=ncell)
    acc, rej, a, _ = sawtooth(k, dt0, t1, True)
    st = ctx.chem_percell(0.0, t1, dt0, None, pictrl_ros23(), LOOSE["atol"], LOOSE["rtol"],
                          RS_POSITIVE_DOMAIN)
    assert (st.accepts_total, st.rejects_total) == (ncell * acc, ncell * rej)
    assert st.attempts_max == st.attempts_min == acc + rej and st.n_failed == 0
    u = ctx.get_state().reshape(2, ncell)
    assert np.all(u[0] >= 0.0)
    np.testing.assert_allclose(u[0], scale * a, rtol=1e-9)
    np.testing.assert_allclose(u.sum(0), scale, rtol=1e-14)
    ctx.close()


def test_rejection_leaves_controller_memory(oracle):
    """A domain rejection sets dt = h/2 and nothing else: a run whose first
    trial is rejected equals a run started at h/2 (same counts less one
    rejection, same end state and dt)."""
    k, t1 = 1.0, 100.0
    p, ctx = decay_box(oracle, k=k)
    res = []
    for dt0 in (4.0, 2.0):
        ctx.set_state(p.u0)
        dt, c = ctx.adaptive_chem(0.0, t1, dt0, pictrl_ros23(), 1e-8, 1e-6, RS_POSITIVE_DOMAIN)
        res.append((dt, c.naccept, c.nreject, ctx.get_state()))
    (d4, a4, r4, u4), (d2, a2, r2, u2) = res
    assert R(-4.0) < 0.0 <= R(-2.0)
    assert (a4, r4 - 1, d4) == (a2, r2, d2) and np.array_equal(u4, u2)
    ctx.close()


def test_transport_checks_own_block(oracle):
    """The transport solve looks at the transported species only: a negative
    chemistry-only species does not stop it."""
    g = P.Grid(12, 8, 4)
    u = np.empty((2, g.nc))
    u[0] = 5.0 + P.blob_state(g, 1, 3)
    u[1] = -1.0
    p = P.Problem(name="own", grid=g, ns=2, nt=1, wind=P.solid_body_wind(g, 600.0, 0.8),
                  mech=None, u0=np.ascontiguousarray(u.reshape(-1)))
    ctx = oracle.context().load(p)
    out = []
    for flags in (0, RS_POSITIVE_DOMAIN):
        ctx.set_state(p.u0)
        dt, c = ctx.adaptive_transport(0.0, 120.0, 15.0, pictrl_ssprk43(), 1e-6, 1e-3, flags)
        out.append((dt, c.naccept, c.nreject, ctx.get_state()))
    assert out[0][:3] == out[1][:3] and np.array_equal(out[0][3], out[1][3])
    assert np.all(out[1][3][:g.nc] > 0.0) and np.all(out[1][3][g.nc:] == -1.0)
    ctx.close()


def test_transport_rule_can_exhaust_maxiters(oracle):
    """Why the split options keep the rule off the transport: the linear
    flux-form scheme undershoots by an amount proportional to the step at a
    sharp front, so halving never clears it and the solve runs out of
    iterations; the chemistry-only rule in rs_macro_step leaves the transport
    as it was."""
    g = P.Grid(12, 8, 4)
    u = np.zeros((2, 4, 8, 12))
    u[:, :, 2:5, 3:6] = 1.0
    m = P.Mechanism(ns=2, reactants=np.array([[0, -1]], np.int32),
                    products=np.array([[1, -1]], np.int32), k=np.array([1e-3]),
                    photo=np.zeros(1, np.int32))
    p = P.Problem(name="front", grid=g, ns=2, nt=2, wind=P.solid_body_wind(g, 600.0, 0.8),
                  mech=m, u0=u.reshape(-1), dt_macro=600.0, nmacro=1)
    ctx = oracle.context().load(p)
    with pytest.raises(ReseactError) as e:
        ctx.adaptive_transport(0.0, 600.0, 15.0, pictrl_ssprk43(), 1e-6, 1e-3,
                               RS_POSITIVE_DOMAIN, maxiters=200)
    assert e.value.code == RS_E_MAXITERS
    # in a split step (with the clamp, so the chemistry starts non-negative)
    # the transport takes the same steps with and without the rule
    res = []
    for posdom in (False, True):
        ctx.set_state(p.u0)
        dtT, dtC, cT, cC = ctx.macro_step(0.0, 600.0, 15.0, 0.5, split_opts(p, positive_domain=posdom))
        res.append((dtT, cT.naccept, cT.nreject, ctx.get_state()))
    (a0, b0, c0, s0), (a1, b1, c1, s1) = res
    assert (a0, b0, c0) == (a1, b1, c1) and s1.min() >= 0.0
    ctx.close()
