"""Pins of the oracle's PI step-size controller (PICtrl / pictrl_ros23 /
adaptive_solve, tools/reactant_handoff/rx_traced_integrator.jl:86-99,
:480-505) and of ROS23's embedded error estimate (ros23_step, :318-347).

These decide which steps are accepted, i.e. "the same accepted step sequence"
every GPU parity test compares, so each check is chosen to fail on a plausible
transcription mistake.  `test_planted_mistakes_are_caught` builds oracle
variants with such a mistake planted in a copy of its source (on the CPU) and
shows that a check fails for each:
  * beta1 and beta2 swapped in the PI update;
  * the sign of the c32 term of k3's right-hand side flipped;
  * the qmin clamp of a rejection dropped;
  * qsteady_max = 1 instead of 6/5 for ROS23 (a wrong controller constant).
CPU only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import EXTRA, ORACLE_DIR, oracle_library
from reseact_esm_amd import problems as P
from reseact_esm_amd.host import Library, PICtrl_make, pictrl_ros23, pictrl_ssprk43

# The reference's constants, rx_traced_integrator.jl:86-95 (qmin, qmax, gamma,
# qsteady_min, qoldinit) and :98 (pictrl_ros23: beta1, beta2, qsteady_max).
BETA1, BETA2, QMIN, QMAX, GAMMA = 7 / 20, 1 / 5, 1 / 5, 10.0, 9 / 10
QSMIN, QSMAX, QOLDINIT = 1.0, 6 / 5, 1e-4

_D = C.POINTER(C.c_double)


def _box(lib, mech, u0):
    g = P.Grid(len(u0) // mech.ns, 1, 1)
    p = P.Problem(name="box", grid=g, ns=mech.ns, nt=0, wind=P.solid_body_wind(g, 1.0, 0.0),
                  mech=mech, u0=np.ascontiguousarray(u0, np.float64))
    return lib.context().load(p)


def attempt_log(ctx) -> np.ndarray:
    """Every attempt of the last adaptive solve: rows (t, dtc, EEst, outcome)."""
    n = C.c_longlong(0)
    ctx._chk(ctx.lib.rs_oracle_attempt_log(ctx.h, None, 0, C.byref(n)))
    out = np.empty((n.value, 4))
    ctx._chk(ctx.lib.rs_oracle_attempt_log(ctx.h, out.ctypes.data_as(_D), n.value, C.byref(n)))
    return out


def decay_mech(k=1.0):
    """u' = -k u (A -> nothing)."""
    return P.Mechanism(ns=1, reactants=np.array([[0, -1]], np.int32),
                       products=np.array([[-1, -1]], np.int32), k=np.array([k]),
                       photo=np.zeros(1, np.int32))


def linear_mech():
    """Five species, first-order reactions only (u' = K u)."""
    from test_oracle_chem import first_order_mech
    return first_order_mech(5, 1), np.linspace(0.5, 2.0, 5)


def nonlinear_mech():
    """A + B -> C, C -> A, 2B -> A: a smooth nonlinear cell."""
    m = P.Mechanism(ns=3, reactants=np.array([[0, 1], [2, -1], [1, 1]], np.int32),
                    products=np.array([[2, -1], [0, -1], [0, -1]], np.int32),
                    k=np.array([2.0, 0.5, 0.3]), photo=np.zeros(3, np.int32))
    return m, np.array([1.2, 0.8, 0.3])


# ---- checks (each returns (ok, detail)) --------------------------------------

def check_eest_order(lib):
    """(a) The embedded estimate ut = h/6 (k1 - 2 k2 + k3) estimates the local
    error of the 2nd-order solution unew:
      * it is O(h^3): EEst(h)/EEst(h/2) -> 8 (+-0.8) as h -> 0, on a linear
        system and on a nonlinear cell (atol small against u);
      * it is that error to leading order: on u' = -k u with atol = 0,
        |ut| = EEst rtol |u| and |ut| / |unew - u e^(-kh)| -> 1 (within 1%
        at k h = 0.05 and 0.025).  A sign slip in the c32 term keeps the
        order but not the constant (the ratio goes to ~2.7)."""
    out = {}
    for name, (m, u0), hs in (("linear", linear_mech(), (0.05, 0.025)),
                              ("nonlinear", nonlinear_mech(), (0.0125, 0.00625, 0.003125))):
        ctx = _box(lib, m, u0)
        ratios = []
        for h in hs:
            ctx.set_state(u0)
            e1 = ctx.ros23_step(0.0, h, 1e-12, 1e-3)
            ctx.set_state(u0)
            e2 = ctx.ros23_step(0.0, h / 2, 1e-12, 1e-3)
            ratios.append(e1 / e2)
        out[name] = ratios
    ok = all(abs(r[-1] - 8.0) <= 0.8 for r in out.values())
    ctx = _box(lib, decay_mech(1.0), np.array([1.0]))
    est = []
    for h in (0.05, 0.025):
        ctx.set_state(np.array([1.0]))
        e = ctx.ros23_step(0.0, h, 0.0, 1e-3)
        err = ctx.get_candidate()[0] - np.exp(-h)
        est.append(e * 1e-3 / abs(err))
    out["estimate/true error"] = est
    ok = ok and all(abs(r - 1.0) <= 0.01 for r in est)
    return ok, out


def check_steady_state(lib, ctrl):
    """(b) On u' = -k u with atol = 0 the error estimate is a function of k*dt
    alone (it scales with u), so the controller must settle to a constant
    accepted step (the qsteady snap) whose EEst lies in the band where the
    snapped q = EEst^(beta1-beta2)/gamma is in [qsteady_min, qsteady_max] and
    the step is accepted: [0.9^(20/3), 1] ~ [0.495, 1].  The accepted sequence
    must also be invariant under scaling u by 2^20."""
    seqs = []
    for u0 in (1.0, 2.0 ** 20):
        ctx = _box(lib, decay_mech(), np.array([u0]))
        try:
            ctx.adaptive_chem(0.0, 60.0, 1e-3, ctrl, 0.0, 1e-3, False, maxiters=2000)
        except Exception as e:   # noqa: BLE001 -- a controller that never settles fails here
            return False, f"solve failed: {e}"
        seqs.append(attempt_log(ctx))
    L = seqs[0]
    if not np.array_equal(L[:, 1], seqs[1][:, 1]):
        return False, "accepted sequence depends on the scale of u"
    acc = L[L[:, 3] == 1]
    tail = acc[-21:-1]            # the last step is cut to end at t = 60
    lo = 0.9 ** (20 / 3)
    # constant to rounding: at the band's lower edge q = 1 up to the last bit
    same_dt = np.all(np.abs(tail[:, 1] / tail[0, 1] - 1.0) <= 1e-12)
    in_band = np.all((tail[:, 2] >= lo * (1 - 1e-9)) & (tail[:, 2] <= 1.0))
    return bool(same_dt and in_band), dict(dt=tail[:, 1].tolist()[-3:], eest=tail[:, 2].tolist()[-3:])


def check_reject_qmin_clamp(lib, ctrl):
    """(c) A rejection with EEst >= 4.5^(20/7) (q11/gamma >= 1/qmin) shrinks
    dt by exactly 1/qmin = 5; a rejection below that shrinks it by q11/gamma,
    so between two unsaturated rejections log(shrink) changes as beta1 *
    log(EEst)."""
    sat, unsat = [], []
    for dt0 in (40.0, 25.0, 12.0, 5.0, 2.5, 2.0, 1.5, 1.2):
        ctx = _box(lib, decay_mech(), np.array([1.0]))
        ctx.adaptive_chem(0.0, 60.0, dt0, ctrl, 1e-6, 1e-3, False)
        L = attempt_log(ctx)
        e0 = L[0, 2]
        if L[0, 3] != 0:
            continue
        shrink = L[0, 1] / L[1, 1]
        (sat if e0 >= 4.5 ** (20 / 7) else unsat).append((e0, shrink))
    if not sat or len(unsat) < 2:
        return False, ("no cases", sat, unsat)
    ok_sat = all(s == 5.0 for _, s in sat)
    (ea, sa), (eb, sb) = unsat[0], unsat[-1]
    slope = (np.log(sa) - np.log(sb)) / (np.log(ea) - np.log(eb))
    ok_unsat = abs(slope - BETA1) < 1e-9 and all(s < 5.0 for _, s in unsat)
    return bool(ok_sat and ok_unsat), dict(sat=sat, unsat=unsat, slope=slope)


def check_transitions(lib, ctrl, op="chem"):
    """(d) Every logged attempt's successor step follows the reference's PI
    update with the reference's constants (typed here from
    rx_traced_integrator.jl:86-99, :480-505, independently of the oracle;
    pictrl_ros23 for the chemistry, pictrl_ssprk43 for the transport):
    accept -> dtc / (1 if q in [qsteady_min, qsteady_max] else q),
    reject -> dtc / min(1/qmin, q11/gamma), with q = clamp(q11/qold^beta2/gamma,
    1/qmax, 1/qmin), q11 = max(EEst, 1e-35)^beta1, qold = max(EEst, qoldinit)
    after an accept.  On a run of the coupled workload's chemistry that takes
    every branch: unclamped accept, snap, the qmin-clamped accept of the first
    step, unclamped and clamped rejections.  The transport run (SSPRK43 on
    the same workload's grid) takes accepts and both kinds of rejection."""
    tend = 600.0
    if op == "chem":
        ctx = lib.context().load(P.config1(nlon=2, nlat=1, nlev=1))
        dt_end, _ = ctx.adaptive_chem(0.0, tend, 50.0, ctrl, 1e-6, 1e-3, False)
        BETA1, BETA2, QSMAX = 7 / 20, 1 / 5, 6 / 5             # :98
        want = {"acc", "snap", "acc_qmin", "rej", "rej_qmin"}
    else:
        ctx = lib.context().load(P.config1(nlon=12, nlat=4, nlev=4))
        tend = 7200.0
        dt_end, _ = ctx.adaptive_transport(0.0, tend, tend, ctrl, 1e-6, 1e-3, False)
        BETA1, BETA2, QSMAX = 7 / 30, 2 / 15, 1.0              # :99
        want = {"acc", "rej", "rej_qmin"}
    L = attempt_log(ctx)
    qold, seen, bad = QOLDINIT, set(), []
    for i, (t, dtc, e, outcome) in enumerate(L):
        E = 1e10 if np.isnan(e) else e
        q11 = max(E, 1e-35) ** BETA1
        q = max(1 / QMAX, min(1 / QMIN, q11 / qold ** BETA2 / GAMMA))
        if outcome == 1:
            snap = QSMIN <= q <= QSMAX
            nxt = dtc / (1.0 if snap else q)
            seen.add("snap" if snap else ("acc_qmin" if q == 1 / QMIN else "acc"))
            qold = max(E, QOLDINIT)
            tn = t + dtc
        else:
            clamp = q11 / GAMMA >= 1 / QMIN
            nxt = dtc / min(1 / QMIN, q11 / GAMMA)
            seen.add("rej_qmin" if clamp else "rej")
            tn = t
        if i + 1 < len(L):
            got = L[i + 1, 1]
            if got < tend - tn - 1e-9 and abs(got - nxt) > 1e-12 * nxt:   # not cut at tend
                bad.append((i, got, nxt))
        elif abs(dt_end - nxt) > 1e-12 * nxt:
            bad.append((i, dt_end, nxt))
    return (not bad and want <= seen), dict(bad=bad[:3], seen=sorted(seen))


# ---- tests on the oracle ------------------------------------------------------

def test_eest_is_third_order(oracle):
    ok, d = check_eest_order(oracle)
    assert ok, d


def test_controller_settles_in_the_qsteady_band(oracle):
    ok, d = check_steady_state(oracle, pictrl_ros23())
    assert ok, d


def test_reject_shrinks_by_qmin_clamp_and_beta1(oracle):
    ok, d = check_reject_qmin_clamp(oracle, pictrl_ros23())
    assert ok, d


def test_controller_transitions_follow_reference_update(oracle):
    ok, d = check_transitions(oracle, pictrl_ros23())
    assert ok, d


def test_transport_controller_transitions_follow_reference_update(oracle):
    ok, d = check_transitions(oracle, pictrl_ssprk43(), op="transport")
    assert ok, d


def test_pictrl_constants():
    c = pictrl_ros23()
    assert (c.beta1, c.beta2, c.qmin, c.qmax, c.gamma, c.qsteady_min, c.qsteady_max,
            c.qoldinit) == (BETA1, BETA2, QMIN, QMAX, GAMMA, QSMIN, QSMAX, QOLDINIT)
    c = pictrl_ssprk43()
    assert (c.beta1, c.beta2, c.qsteady_max) == (7 / 30, 2 / 15, 1.0)


# ---- the checks catch planted mistakes ----------------------------------------

PLANTS = {
    # beta1 and beta2 swapped in the PI update
    "swap_beta": [("std::pow(std::max(EEst, 1.0e-35), c.beta1)",
                   "std::pow(std::max(EEst, 1.0e-35), c.beta2)"),
                  ("std::pow(s.qold, c.beta2)", "std::pow(s.qold, c.beta1)")],
    # sign slip in the c32 term of k3's right-hand side
    "c32_sign": [("a = ROS23_C32 * a;", "a = -ROS23_C32 * a;")],
    # the rejection's qmin clamp dropped
    "no_qmin_clamp": [("double qr = std::min(invqmin, q11 / c.gamma);",
                       "double qr = q11 / c.gamma;")],
}


def _planted_library(tmp_path, name):
    src = open(os.path.join(ORACLE_DIR, "reseact_cpu.cpp")).read()
    for old, new in PLANTS[name]:
        assert src.count(old) == 1, (name, old)
        src = src.replace(old, new)
    d = tmp_path / name
    d.mkdir()
    cpp = d / "reseact_cpu.cpp"
    cpp.write_text(src.replace('#include "../include/reseact_hip.h"',
                               '#include "%s"' % os.path.join(ORACLE_DIR, "..", "include",
                                                              "reseact_hip.h")))
    so = d / "libplanted.so"
    return cpp, so


def test_planted_mistakes_are_caught(tmp_path):
    builds = {}
    for name in PLANTS:
        cpp, so = _planted_library(tmp_path, name)
        builds[name] = (so, subprocess.Popen(
            ["g++", "-O1", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fPIC",
             "-std=c++17", "-shared", "-I", ORACLE_DIR, "-o", str(so), str(cpp)]))
    libs = {}
    for name, (so, proc) in builds.items():
        assert proc.wait() == 0, name
        libs[name] = Library(str(so), EXTRA)
    ref = pictrl_ros23()
    # swapped betas: the controller never settles into the band
    assert not check_steady_state(libs["swap_beta"], ref)[0]
    assert not check_transitions(libs["swap_beta"], ref)[0]
    # c32 sign: the estimate falls to O(h^2)
    assert not check_eest_order(libs["c32_sign"])[0]
    # no qmin clamp on rejections
    assert not check_reject_qmin_clamp(libs["no_qmin_clamp"], ref)[0]
    assert not check_transitions(libs["no_qmin_clamp"], ref)[0]
    # qsteady_max = 1 for ROS23 (a wrong constant in the controller settings)
    wrong = PICtrl_make(BETA1, BETA2, 1.0)
    assert not check_transitions(oracle_library(), wrong)[0]
    # the ROS23 constants used for SSPRK43 (or the reverse)
    assert not check_transitions(oracle_library(), ref, op="transport")[0]
    assert not check_transitions(oracle_library(), pictrl_ssprk43())[0]
    # and the unplanted oracle passes them all
    assert check_eest_order(oracle_library())[0]


@pytest.mark.parametrize("name", sorted(PLANTS))
def test_plants_apply_to_the_current_source(name):
    src = open(os.path.join(ORACLE_DIR, "reseact_cpu.cpp")).read()
    for old, _ in PLANTS[name]:
        assert src.count(old) == 1
