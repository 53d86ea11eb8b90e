"""Pins of the oracle's time-loop discrete adjoint (checkpoint, fixed-sequence
replay, reverse sweep with the clamp mask) against central finite differences
of the frozen-sequence forward map (the reference's fdtape check,
tools/adjoint_gradient.jl:1183-1244) and the structural identity
s dJ/ds = <E, dJ/dE> for emissions scaled by s (:962-970).  CPU only."""
from __future__ import annotations

import numpy as np

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import split_opts


def small_c4(nmacro=2):
    return P.config4(nlon=8, nlat=5, nlev=4, nmacro=nmacro)


def run(oracle, p):
    ctx = oracle.context().load(p)
    J, gu, ge, st = ctx.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p))
    return ctx, J, gu, ge, st


def frozen_J(ctx, p, u0, emis):
    ctx.set_emissions(emis)
    ctx.set_state(u0)
    return ctx.replay_loss(0)


def test_replay_is_bit_identical(oracle):
    p = small_c4()
    ctx, J, gu, ge, st = run(oracle, p)
    assert st.replay_maxrel == 0.0
    assert frozen_J(ctx, p, p.u0, p.emis) == J


def test_gradient_u0_against_central_fd(oracle):
    p = small_c4()
    ctx, J, gu, ge, st = run(oracle, p)
    rng = np.random.default_rng(1004)
    for _ in range(3):
        v = rng.standard_normal(p.u0.size) * p.u0
        pred = float(np.dot(gu, v))
        best = np.inf
        for eps in (1e-4, 1e-5, 1e-6, 1e-7):
            jp = frozen_J(ctx, p, p.u0 + eps * v, p.emis)
            jm = frozen_J(ctx, p, p.u0 - eps * v, p.emis)
            best = min(best, abs((jp - jm) / (2 * eps) - pred) / abs(pred))
        assert best <= 1e-6, best


def test_gradient_emissions_against_central_fd(oracle):
    p = small_c4()
    ctx, J, gu, ge, st = run(oracle, p)
    rng = np.random.default_rng(7)
    for _ in range(3):
        v = rng.standard_normal(p.emis.size) * p.emis
        pred = float(np.dot(ge, v))
        best = np.inf
        for eps in (1e-3, 1e-4, 1e-5, 1e-6):
            jp = frozen_J(ctx, p, p.u0, p.emis + eps * v)
            jm = frozen_J(ctx, p, p.u0, p.emis - eps * v)
            best = min(best, abs((jp - jm) / (2 * eps) - pred) / abs(pred))
        assert best <= 1e-6, best


def test_structural_identity_emission_scale(oracle):
    # J(s) with E -> s E: s dJ/ds at s = 1 equals <E, dJ/dE>
    p = small_c4()
    ctx, J, gu, ge, st = run(oracle, p)
    pred = float(np.dot(p.emis, ge))
    eps = 1e-5
    jp = frozen_J(ctx, p, p.u0, p.emis * (1 + eps))
    jm = frozen_J(ctx, p, p.u0, p.emis * (1 - eps))
    assert abs((jp - jm) / (2 * eps) - pred) <= 1e-6 * abs(pred)


def test_loss_gradient_seed_without_steps(oracle):
    # with zero-length macro steps the gradient is dJ/du = 2 u on surface cells of species 0
    p = small_c4(nmacro=1)
    ctx = oracle.context().load(p)
    J, gu, ge, st = ctx.adjoint_run(np.array([0.0, 0.0]), p.dt0_T, p.dt0_C, split_opts(p))
    ncol = p.grid.ncol
    expect = np.zeros_like(p.u0)
    expect[:ncol] = 2 * p.u0[:ncol]
    assert np.array_equal(gu, expect)
    assert abs(J - float(np.sum(p.u0[:ncol] ** 2))) <= 1e-14 * J
    assert not ge.any()
