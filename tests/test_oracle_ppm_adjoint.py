"""Pins of the oracle's derivatives of the monotone PPM transport (the
tangent J(u) v and the adjoint J(u)^T lam of the tendency, and of one SSPRK43
step at fixed dt).  CPU only.

The tendency is piecewise smooth: every slope and parabola limiter switches on
a product of neighbour differences (CW84 eq. 1.8, 1.10).  The derivatives are
those of the branch the forward takes, so they are pinned where the branch is
locally constant, on seeded random states:
* the dot-product identity <J^T lam, v> = <lam, J v> between the two
  independently derived forms (forward-mode tangent, reverse-mode face
  gradients), for the tendency and for the step;
* central finite differences of the tendency and of the step map;
* closed forms: with the flux of one face only and every limiter on its flat
  branch the face value is the upwind cell's mean, so the derivative is a
  single entry.
The reference's caveat (DIFFERENTIABILITY_PLAN.md:709-736) is reproduced on
its own case, the CWC gate (uniform mixing ratios): every guard product is 0
at the base point and quadratic in a perturbation, so the forward and the
backward one-sided quotients agree with each other and both differ from the
active-branch derivative; a jittered base point removes the disagreement.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from reseact_esm_amd import problems as P

_D = C.POINTER(C.c_double)


def _dp(a):
    return a.ctypes.data_as(_D)


def rhs(ctx, u):
    u = np.ascontiguousarray(u, np.float64)
    f = np.empty_like(u)
    ctx._chk(ctx.lib.rs_oracle_transport_rhs(ctx.h, _dp(u), _dp(f)))
    return f


def rhs_jvp(ctx, v):
    v = np.ascontiguousarray(v, np.float64)
    o = np.empty_like(v)
    ctx._chk(ctx.lib.rs_oracle_transport_rhs_jvp(ctx.h, _dp(v), _dp(o)))
    return o


def rhs_vjp(ctx, lam):
    lam = np.ascontiguousarray(lam, np.float64)
    o = np.empty_like(lam)
    ctx._chk(ctx.lib.rs_oracle_transport_rhs_adj(ctx.h, _dp(lam), _dp(o)))
    return o


def step_jvp(ctx, dt, v):
    v = np.ascontiguousarray(v, np.float64)
    o = np.empty_like(v)
    ctx._chk(ctx.lib.rs_oracle_ssprk43_jvp(ctx.h, dt, _dp(v), _dp(o)))
    return o


def step_map(ctx, u, dt):
    ctx.set_state(u)
    ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
    return ctx.get_candidate()


def step_vjp(ctx, u, dt, lam):
    ctx.set_state(u)
    ctx.set_lambda(lam)
    ctx.ssprk43_vjp(0.0, dt)
    return ctx.get_lambda()


def regional(seed=3, nq=2):
    """The CONUS patch with open walls, random mixing ratios and inflow halos."""
    p = P.transport3d_regional(nlon=9, nlat=7, nlev=8, nq=nq)
    rng = np.random.default_rng(seed)
    nc = p.nc
    p.u0 = p.u0.copy()
    p.u0[nc:] = rng.lognormal(0.0, 0.5, nq * nc)
    for name in ("qbc_w", "qbc_e", "qbc_s", "qbc_n"):
        a = getattr(p.ppm, name)
        setattr(p.ppm, name, rng.lognormal(0.0, 0.5, a.shape))
    return p, 0.02


def global_grid(seed=4):
    """Global grid: periodic longitude, polar faces of zero length."""
    p = P.ppm_global(nlon=16, nlat=10, nlev=6, nq=2, seed=seed)
    return p, 20000.0


def regional_pbl():
    p, dt = regional()
    return P.with_pbl_mixing(p, pblh_mean=800.0, pblh_amp=400.0, tau=0.5), dt


def global_pbl():
    p, dt = global_grid()
    # a deep boundary layer (3-9 km), so that it spans several of the 6 layers
    return P.with_pbl_mixing(p, pblh_mean=6000.0, pblh_amp=3000.0, tau=3600.0), dt


# the PBL mixing term (reseact.esm:1563-1690) on both grids
CASES = {"regional": regional, "global": global_grid, "regional-pbl": regional_pbl,
         "global-pbl": global_pbl}


@pytest.mark.parametrize("case", list(CASES))
def test_rhs_dot_identity(oracle, case):
    p, _ = CASES[case]()
    ctx = oracle.context().load(p)
    rng = np.random.default_rng(7)
    for _ in range(3):
        lam = rng.standard_normal(p.ns * p.nc)
        v = rng.standard_normal(p.ns * p.nc)
        a = float(np.dot(rhs_vjp(ctx, lam), v))
        b = float(np.dot(lam, rhs_jvp(ctx, v)))
        assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))
    ctx.close()


@pytest.mark.parametrize("case", list(CASES))
def test_rhs_tangent_matches_central_fd(oracle, case):
    p, _ = CASES[case]()
    ctx = oracle.context().load(p)
    rng = np.random.default_rng(8)
    u = p.u0
    for _ in range(3):
        v = rng.standard_normal(u.size) * np.abs(u)
        jv = rhs_jvp(ctx, v)
        e = 1e-7
        fd = (rhs(ctx, u + e * v) - rhs(ctx, u - e * v)) / (2 * e)
        assert np.linalg.norm(fd - jv) <= 1e-7 * np.linalg.norm(jv)
    ctx.close()


@pytest.mark.parametrize("case", list(CASES))
def test_step_adjoint(oracle, case):
    """One SSPRK43 step at fixed dt: the VJP against the step's tangent (dot
    product) and against central differences of the step map."""
    p, dt = CASES[case]()
    ctx = oracle.context().load(p)
    rng = np.random.default_rng(9)
    u = p.u0.copy()
    lam = rng.standard_normal(u.size)
    g = step_vjp(ctx, u, dt, lam)
    ctx.set_state(u)
    for _ in range(3):
        v = rng.standard_normal(u.size) * np.abs(u)
        jv = step_jvp(ctx, dt, v)
        a, b = float(np.dot(g, v)), float(np.dot(lam, jv))
        assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))
        e = 1e-7
        fd = (step_map(ctx, u + e * v, dt) - step_map(ctx, u - e * v, dt)) / (2 * e)
        ctx.set_state(u)
        assert abs(float(np.dot(lam, fd)) - a) <= 1e-7 * abs(a)
    ctx.close()


def test_flat_branch_closed_form(oracle):
    """A single non-zero longitude face flux M > 0 on a field whose every
    limiter takes the flat branch (a square step sampled so that each upwind
    stencil has a zero neighbour difference): the face value is the upwind
    cell's mean, so d f / d q is one entry per adjacent cell, -+M gx / m."""
    g = P.Grid(8, 1, 1)
    ce = np.ones(2)
    dS = np.ones(1)
    dphi = np.ones(1)
    Mx = np.zeros((1, 1, 9))
    Mx[0, 0, 4] = 0.3                                  # face between cells 3 and 4
    My = np.zeros((1, 2, 8))
    Mz = np.zeros((2, 1, 8))
    dp = np.ones((1, 1, 8))
    ppm = P.PPMSetup(True, 1.0, ce, dS, dphi, Mx, My, Mz, dp)
    q = np.array([1.0, 1.0, 2.0, 2.0, 3.0, 3.0, 1.0, 1.0])
    m = np.full(8, 2.0)
    p = P.Problem(name="line", grid=g, ns=2, nt=2, wind=P.solid_body_wind(g, 1.0, 0.0), mech=None,
                  u0=np.concatenate([m, q]), ppm=ppm)
    ctx = oracle.context().load(p)
    gx = 1.0 / (1.0 * 1.0)
    J = np.zeros((8, 8))
    for a in range(8):
        v = np.zeros(16)
        v[8 + a] = 1.0
        J[:, a] = rhs_jvp(ctx, v)[8:]
    want = np.zeros((8, 8))
    # f_3 = (q3 divM3 - M gx qface) / m, divM3 = M gx; f_4 = (q4 divM4 + M gx qface) / m
    want[3, 3] = (0.3 * gx - 0.3 * gx) / 2.0
    want[4, 4] = -0.3 * gx / 2.0
    want[4, 3] = 0.3 * gx / 2.0
    assert np.allclose(J, want, rtol=0, atol=1e-15)
    lam = np.zeros(16)
    lam[8 + 4] = 1.0
    assert np.allclose(rhs_vjp(ctx, lam)[8:], want[4], rtol=0, atol=1e-15)
    ctx.close()


def test_kink_at_uniform_field(oracle):
    """The reference's caveat on its own case: at the CWC gate's uniform
    mixing ratios the active-branch derivative (first-order upwind faces)
    differs from finite differences while the forward and backward quotients
    agree with each other; from a jittered base point they all agree."""
    p = P.transport3d_regional(nlon=9, nlat=7, nlev=8, nq=1)
    ctx = oracle.context().load(p)
    nc = p.nc
    rng = np.random.default_rng(10)
    v = np.zeros(p.ns * nc)
    v[nc:] = rng.standard_normal(nc)
    u0 = p.u0.copy()
    jv = rhs_jvp(ctx, v)[nc:]
    e = 1e-6
    f0 = rhs(ctx, u0)[nc:]
    fwd = (rhs(ctx, u0 + e * v)[nc:] - f0) / e
    bwd = (f0 - rhs(ctx, u0 - e * v)[nc:]) / e
    assert np.linalg.norm(fwd - bwd) <= 1e-6 * np.linalg.norm(fwd)     # one-sided agree
    assert np.linalg.norm(fwd - jv) >= 0.1 * np.linalg.norm(fwd)       # ... but not with AD
    # jittered base point: the branch is locally constant again
    uj = u0.copy()
    uj[nc:] += 1e-3 * rng.standard_normal(nc)
    ctx.set_state(uj)
    jv = rhs_jvp(ctx, v)[nc:]
    e = 1e-8
    fd = (rhs(ctx, uj + e * v)[nc:] - rhs(ctx, uj - e * v)[nc:]) / (2 * e)
    assert np.linalg.norm(fd - jv) <= 1e-6 * np.linalg.norm(jv)
    ctx.close()


@pytest.mark.parametrize("pbl", [False, True])
def test_time_loop_adjoint_against_central_fd(oracle, pbl):
    """The discrete adjoint of split steps with the PPM transport (checkpoint,
    fixed-sequence replay, reverse sweep) against central differences of the
    frozen-sequence loss (tools/adjoint_gradient.jl:1183-1244), at a random
    (off-kink) state; also with the PBL mixing term (relaxation time of the
    patch's own time scale)."""
    from reseact_esm_amd.host import split_opts
    p = P.ppm_coupled_regional()
    if pbl:
        P.with_pbl_mixing(p, pblh_mean=800.0, pblh_amp=400.0, tau=0.05)
    ctx = oracle.context().load(p)
    J, gu, ge, st = ctx.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p), loss_species=1)
    assert st.naccept_T >= 2 and st.naccept_C >= 2 and st.replay_maxrel == 0.0
    rng = np.random.default_rng(11)
    for _ in range(3):
        v = rng.standard_normal(p.u0.size) * p.u0
        pred = float(np.dot(gu, v))
        best = np.inf
        for eps in (1e-5, 1e-6, 1e-7):
            ctx.set_state(p.u0 + eps * v)
            jp = ctx.replay_loss(1)
            ctx.set_state(p.u0 - eps * v)
            jm = ctx.replay_loss(1)
            best = min(best, abs((jp - jm) / (2 * eps) - pred) / abs(pred))
        assert best <= 1e-6, best
    ctx.close()
