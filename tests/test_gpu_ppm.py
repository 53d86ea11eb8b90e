"""GPU: the monotone PPM transport (CWC mixing-ratio form) on the HIP
library against the oracle.  Both evaluate the same statements in the same
order without contraction, so single attempts agree bit for bit; the error
norm is the same sum in another order (1e-12), so adaptive solves take the
same accepted/rejected sequence with step sizes equal to the last bits and
states to 1e-12 of the largest entry.

* The reference's CWC gate on the 7x7x7 CONUS patch: q stays exactly 1 to
  t = 0.1 on the GPU too, and its falsifier moves q off 1 exactly as on the
  oracle.
* Regional open walls with random mixing ratios and random inflow halos, and
  a global grid (periodic longitude, poles, hybrid layers, continuity Mz),
  each also with columns taller than a workgroup's level block: single
  attempts bitwise equal to the oracle, adaptive solves (with the clamp) with
  the oracle's accepted/rejected counts.
"""
from __future__ import annotations

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import RS_CLAMP_NONNEG, pictrl_ssprk43

pytestmark = pytest.mark.gpu


def regional_random(nlon=20, nlat=12, nlev=9, nq=3, seed=4):
    p = P.transport3d_regional(nlon, nlat, nlev, nq=nq)
    rng = np.random.default_rng(seed)
    nc = p.nc
    p.u0[nc:] = rng.lognormal(0.0, 0.7, nq * nc)
    pp = p.ppm
    pp.qbc_w = rng.lognormal(0.0, 0.7, pp.qbc_w.shape)
    pp.qbc_e = rng.lognormal(0.0, 0.7, pp.qbc_e.shape)
    pp.qbc_s = rng.lognormal(0.0, 0.7, pp.qbc_s.shape)
    pp.qbc_n = rng.lognormal(0.0, 0.7, pp.qbc_n.shape)
    return p


CASES = {
    "gate-7x7x7": lambda: P.transport3d_regional(),
    "falsifier-7x7x7": lambda: P.transport3d_regional(qbc_scale=1.000001),
    "regional-20x12x9": regional_random,
    "global-48x24x10": lambda: P.ppm_global(48, 24, 10, nq=3, seed=9),
    # columns taller than one workgroup's level block: the walk restarts at
    # levels 18 and 36 (53 levels: blocks of 18, 18, 17), or 15 (30 levels)
    "global-40x12x53": lambda: P.ppm_global(40, 12, 53, nq=3, seed=11),
    "regional-20x12x30": lambda: regional_random(20, 12, 30, seed=12),
    # with the PBL mixing term (reseact.esm:1563-1690): a 1-4 km boundary
    # layer over many of 40 levels; the regional patch relaxes in 0.05 s
    "global-pbl-40x12x40": lambda: P.with_pbl_mixing(P.ppm_global(40, 12, 40, nq=3, seed=13),
                                                     pblh_mean=2500.0, pblh_amp=1500.0),
    "regional-pbl-20x12x30": lambda: P.with_pbl_mixing(regional_random(20, 12, 30, seed=14),
                                                       pblh_mean=800.0, pblh_amp=400.0, tau=0.05),
}


@pytest.mark.parametrize("name", list(CASES))
def test_ppm_step_bitwise(hip, oracle, name):
    print(f"ppm step {name}", flush=True)
    p = CASES[name]()
    cg = hip.context().load(p)
    co = oracle.context().load(p)
    for dt in ((30.0, 300.0) if name.startswith("global") else (0.01, 0.05)):
        eg = cg.ssprk43_step(0.0, dt, 1e-6, 1e-3)
        eo = co.ssprk43_step(0.0, dt, 1e-6, 1e-3)
        ug, uo = cg.get_candidate(), co.get_candidate()
        assert np.array_equal(ug, uo), (name, dt, float(np.max(np.abs(ug - uo))))
        assert abs(eg - eo) <= 1e-12 * max(eo, 1e-300)
    cg.close()
    co.close()


@pytest.mark.parametrize("name", list(CASES))
def test_ppm_adaptive_bitwise(hip, oracle, name):
    print(f"ppm adaptive {name}", flush=True)
    p = CASES[name]()
    # the patch's air mass (m ~ 2 Pa) against its fluxes (~0.7 Pa/s) makes
    # sub-second steps; the global grid moves at 20 m/s
    # (the 30-layer patch's thin layers stall the solve near t = 0.34 on
    # both libraries, so it runs a shorter window)
    t1, dt0 = (1800.0, 60.0) if name.startswith("global") else \
        ((0.1, 0.01) if "7x7x7" in name else ((0.05, 0.001) if name.endswith("x30") else (0.5, 0.01)))
    out = []
    for lib in (hip, oracle):
        ctx = lib.context().load(p)
        dt, c = ctx.adaptive_transport(0.0, t1, dt0, pictrl_ssprk43(), 1e-6, 1e-3, RS_CLAMP_NONNEG)
        out.append((dt, c.naccept, c.nreject, ctx.get_state()))
        ctx.close()
    (dg, ag, rg, ug), (do, ao, ro, uo) = out
    print(f"  gpu {ag}/{rg} oracle {ao}/{ro}", flush=True)
    assert (ag, rg) == (ao, ro) and ag >= 1
    assert abs(dg - do) <= 1e-10 * do
    # the error norm is the same sum in another order (per workgroup here,
    # in runs of cells on the oracle), so the accepted step sizes differ in
    # their last bits and the states by round-off after many steps; every
    # single attempt is bit-for-bit (test_ppm_step_bitwise)
    rel = float(np.max(np.abs(ug - uo)) / np.max(np.abs(uo)))
    print(f"  max entry difference relative to the largest entry: {rel:.1e}", flush=True)
    assert rel <= 1e-12
    if name == "gate-7x7x7":
        assert np.all(ug[p.nc:] == 1.0)
    if name == "falsifier-7x7x7":
        assert np.max(np.abs(ug[p.nc:] - 1.0)) > 1e-8


def _relnorm(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


VJP_CASES = {
    "regional-20x12x9": regional_random,
    "global-48x24x10": lambda: P.ppm_global(48, 24, 10, nq=3, seed=9),
    # a periodic row shorter than a stencil: the per-cell gather kernel
    "global-4x6x5": lambda: P.ppm_global(4, 6, 5, nq=2, seed=3),
    # one level (the vertical sweep folds every clamped index onto it) and
    # lines shorter than a stencil in latitude and the vertical
    "thin-12x3x1": lambda: P.ppm_global(12, 3, 1, nq=2, seed=5),
    "thin-8x2x2": lambda: P.ppm_global(8, 2, 2, nq=2, seed=5),
    "regional-9x2x2": lambda: regional_random(9, 2, 2, nq=2, seed=6),
    # several level blocks per column in the recompute stages
    "global-40x12x53": lambda: P.ppm_global(40, 12, 53, nq=3, seed=11),
    # the PBL mixing term, on the sweeps and on the per-cell gather
    "global-pbl-40x12x40": lambda: P.with_pbl_mixing(P.ppm_global(40, 12, 40, nq=3, seed=13),
                                                     pblh_mean=2500.0, pblh_amp=1500.0),
    "global-pbl-4x6x12": lambda: P.with_pbl_mixing(P.ppm_global(4, 6, 12, nq=2, seed=15),
                                                   pblh_mean=6000.0, pblh_amp=3000.0),
}


@pytest.mark.parametrize("name", list(VJP_CASES))
def test_ppm_step_vjp(hip, oracle, name):
    """One SSPRK43 step VJP of the PPM transport at the context's state: the
    stage states are the forward's bits, the adjoint sums run in another order
    (line sweeps or a per-cell gather on the GPU, a scatter in the oracle)."""
    print(f"ppm step vjp {name}", flush=True)
    p = VJP_CASES[name]()
    # steps of a few tenths of the Courant limit, so J^T lam is not lost in lam
    dt = {"regional-20x12x9": 0.03, "global-48x24x10": 3000.0, "global-4x6x5": 50000.0,
          "thin-12x3x1": 20000.0, "thin-8x2x2": 20000.0, "regional-9x2x2": 0.05,
          "global-40x12x53": 2000.0, "global-pbl-40x12x40": 2000.0,
          "global-pbl-4x6x12": 50000.0}[name]
    lam = np.random.default_rng(21).standard_normal(p.ns * p.nc)
    out = []
    for lib in (hip, oracle):
        with lib.context() as c:
            c.load(p)
            c.set_lambda(lam)
            c.ssprk43_vjp(0.0, dt)
            out.append(c.get_lambda())
    lg, lo = out
    assert _relnorm(lg, lo) <= 1e-12
    assert float(np.max(np.abs(lg - lo))) <= 1e-11 * float(np.max(np.abs(lo)))


@pytest.mark.parametrize("pbl", [False, True])
def test_ppm_adjoint_run(hip, oracle, pbl):
    """Checkpointed time-loop adjoint of split steps with the PPM transport
    (with and without the PBL mixing term) and a small mechanism: same
    accepted sequences, gradient to 1e-10."""
    from reseact_esm_amd.host import split_opts
    print(f"ppm adjoint run pbl={pbl}", flush=True)
    p = P.ppm_coupled_regional()
    if pbl:
        P.with_pbl_mixing(p, pblh_mean=800.0, pblh_amp=400.0, tau=0.05)
    res = []
    prof = None
    for lib in (hip, oracle):
        with lib.context() as c:
            c.load(p)
            if lib is hip:
                c.profile(True)
            res.append(c.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p), loss_species=1))
            if lib is hip:
                prof = c.profile_read()
    (Jg, gug, _, sg), (Jo, guo, _, so) = res
    assert (sg.naccept_T, sg.nreject_T, sg.naccept_C) == (so.naccept_T, so.nreject_T, so.naccept_C)
    assert abs(Jg - Jo) <= 1e-10 * abs(Jo)
    assert _relnorm(gug, guo) <= 1e-10
    assert sg.replay_maxrel == 0.0
    # the replay keeps every transport step's stage states in the tape, so the
    # VJPs recompute none: every stage launch belongs to an attempt (three
    # stages and the last; a recomputing VJP would add four stage launches)
    n_stage = prof.get("transport_stage", (0, 0.0))[0]
    n_last = prof.get("transport_last", (0, 0.0))[0]
    assert sg.nvjp_T > 0 and n_last > 0 and n_stage == 3 * n_last, (n_stage, n_last, sg.nvjp_T)


KEEP_CASES = ["regional-20x12x9", "global-48x24x10", "global-40x12x53", "global-pbl-40x12x40"]


@pytest.mark.parametrize("name", KEEP_CASES)
def test_kept_stages_vjp_bitwise(hip, name):
    """rs_keep_transport_stages: the step's candidate is unchanged, the VJP that
    reuses the kept stages equals the recomputing VJP bit for bit, a second
    VJP of the same step reuses them again, and anything that may change the
    state (here rs_set_state) or another dt makes the VJP recompute."""
    print(f"kept stages {name}", flush=True)
    p = CASES[name]()
    dt = 30.0 if name.startswith("global") else 0.01
    rng = np.random.default_rng(7)
    lam0 = rng.normal(size=p.ns * p.nc)
    out = {}
    for keep in (False, True):
        c = hip.context().load(p)
        c.keep_transport_stages(keep)
        c.ssprk43_step(0.0, dt, 1e-6, 1e-3)
        cand = c.get_candidate()
        c.set_lambda(lam0)
        c.ssprk43_vjp(0.0, dt)
        l1 = c.get_lambda()
        c.set_lambda(lam0)
        c.ssprk43_vjp(0.0, dt)                    # the same step again
        l2 = c.get_lambda()
        c.set_lambda(lam0)
        c.ssprk43_vjp(0.0, 0.5 * dt)              # another dt: recomputed
        l3 = c.get_lambda()
        c.ssprk43_step(0.0, dt, 1e-6, 1e-3)
        c.set_state(p.u0 * 1.001)                 # state changed: recomputed
        c.set_lambda(lam0)
        c.ssprk43_vjp(0.0, dt)
        l4 = c.get_lambda()
        out[keep] = (cand, l1, l2, l3, l4)
        c.close()
    for a, b in zip(out[False], out[True]):
        assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    assert np.array_equal(out[True][1], out[True][2])
    assert not np.array_equal(out[True][1], out[True][3])
    assert not np.array_equal(out[True][1], out[True][4])
