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


# ---- the reference's linear objective (adjoint_gradient.jl:408-426) -----------

def test_objective_weights_select_the_reference_scopes():
    g = P.Grid(6, 4, 3)
    w, n = P.objective_weights(g, 5, 2, "all")
    assert n == g.nc and np.isclose(w.sum(), 1.0)
    assert np.all(w[2 * g.nc:3 * g.nc] == 1.0 / g.nc) and w[:2 * g.nc].sum() == 0.0
    w, n = P.objective_weights(g, 5, 0, "surf")
    assert n == g.ncol and np.all(w[:g.ncol] == 1.0 / g.ncol) and w[g.ncol:].sum() == 0.0
    w, n = P.objective_weights(g, 5, 1, "lev3")            # the top of 3 levels, 1-based
    lo = g.nc + 2 * g.ncol
    assert n == g.ncol and np.all(w[lo:lo + g.ncol] == 1.0 / g.ncol)
    for bad in ("lev0", "lev4", "column"):
        try:
            P.objective_weights(g, 5, 0, bad)
        except ValueError:
            pass
        else:
            raise AssertionError(bad)


def test_linear_objective_gradient_against_central_fd(oracle):
    """J = mean of species 3 over level 2: the seed is w, J = <w, u_end>, and
    the gradient matches central differences of the frozen-sequence map."""
    p = small_c4()
    w, _ = P.objective_weights(p.grid, p.ns, 3, "lev2")
    ctx = oracle.context().load(p)
    ctx.set_objective_weights(w)
    J, gu, ge, st = ctx.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p), loss_kind=1)
    ctx.set_state(p.u0)
    assert ctx.replay_loss(loss_kind=1) == J
    rng = np.random.default_rng(408)
    for v, grad, apply in ((rng.standard_normal(p.u0.size) * p.u0, gu, "u0"),
                           (rng.standard_normal(p.emis.size) * p.emis, ge, "emis")):
        pred = float(np.dot(grad, v))
        best = np.inf
        for eps in (1e-4, 1e-5, 1e-6):
            out = []
            for sgn in (1, -1):
                if apply == "u0":
                    ctx.set_emissions(p.emis)
                    ctx.set_state(p.u0 + sgn * eps * v)
                else:
                    ctx.set_emissions(p.emis + sgn * eps * v)
                    ctx.set_state(p.u0)
                out.append(ctx.replay_loss(loss_kind=1))
            best = min(best, abs((out[0] - out[1]) / (2 * eps) - pred) / abs(pred))
        assert best <= 1e-6, (apply, best)


# ---- non-finite detection in the backward sweep (adjoint_gradient.jl:763-779) --

def poisoned_weights(p, species=2):
    """Linear objective weights with one NaN, at a surface cell whose state is
    positive at the end of the run (so the clamp mask keeps it)."""
    w, _ = P.objective_weights(p.grid, p.ns, species, "surf")
    w[species * p.grid.nc + 3] = np.nan
    return w


def test_nonfinite_lambda_is_reported_at_its_first_vjp(oracle):
    from reseact_esm_amd.host import RS_E_NONFINITE, ReseactError
    p = small_c4()
    ctx = oracle.context().load(p)
    ctx.set_objective_weights(P.objective_weights(p.grid, p.ns, 2, "surf")[0])
    J, gu, ge, st = ctx.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p), loss_kind=1)
    assert st.bad_op == -1 and np.isfinite(gu).all() and np.isfinite(ge).all()
    ctx.set_state(p.u0)
    ctx.set_objective_weights(poisoned_weights(p))
    try:
        ctx.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p), loss_kind=1)
    except ReseactError as e:
        assert e.code == RS_E_NONFINITE
        s = e.stats
        msg = str(e)
    else:
        raise AssertionError("a NaN in lambda was not reported")
    # the last chemistry step of the last macro step is the first VJP of the sweep
    assert (s.bad_op, s.bad_macro, s.bad_step) == (1, p.nmacro - 1, s.bad_nsteps - 1)
    assert s.bad_lam_in == 1 and s.bad_lam_out > 1 and s.bad_grad >= 1
    assert f"chemistry step {s.bad_nsteps} of {s.bad_nsteps} in macro step {p.nmacro}" in msg
    # t and dt are that step's: it ends the run
    assert abs(s.bad_t + s.bad_dt - p.tstops()[-1]) <= 1e-9 * p.tstops()[-1]



def test_adjoint_on_a_given_sequence_is_the_adjoint(oracle):
    """The oracle's adjoint with the forward forced onto a recorded accepted
    sequence equals the free-running adjoint that recorded it (bit for bit):
    the check GPU parity tests use when the GPU's global error norm, summed in
    another order, takes another sequence."""
    from oracle_lib import adjoint_on_sequence
    p = small_c4()
    ctx, J, gu, ge, st = run(oracle, p)
    seqT = [ctx.recorded_steps(m, 0) for m in range(p.nmacro)]
    seqC = [ctx.recorded_steps(m, 1) for m in range(p.nmacro)]
    assert sum(len(s) for s in seqC) == st.naccept_C
    ctx.set_state(p.u0)
    J2, gu2, ge2, st2 = adjoint_on_sequence(ctx, p, split_opts(p), seqT, seqC)
    assert J2 == J and np.array_equal(gu2, gu) and np.array_equal(ge2, ge)
    assert st2.replay_maxrel == 0.0 and st2.nreject_C == 0
