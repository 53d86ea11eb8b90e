"""GPU, full 144x91x47 grid of the coupled adjoint workload (20 species,
40 reactions): size-independent checks of the time-loop adjoint where the
oracle would take minutes.

* The replay of the recorded step sequence is bit-identical to the forward,
  and the frozen-sequence loss at the forward's inputs is the forward's loss.
* Directional derivatives <dJ/du0, v> and <dJ/dE, v> agree with central
  differences of the frozen-sequence loss (the reference's fdtape check,
  tools/adjoint_gradient.jl:1183-1244) on 32 random directions (16 in u0, 16
  in E, seeded 1004), best over a step sweep, to 1e-6 relative.
* The structural identity s dJ/ds = <E, dJ/dE> for emissions scaled by s
  (:962-970).
Three split steps of 900 s, with the clamp max(u, 0) on.  The clamp makes the
step map piecewise smooth: over more steps some directions move entries across
a kink within the difference step, where a central difference is not the
derivative the adjoint takes (the reference lists the same confounder of its
fd check, adjoint_gradient.jl:1168-1181).
"""
from __future__ import annotations

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import split_opts

pytestmark = pytest.mark.gpu

NDIR = 16


@pytest.fixture(scope="module")
def run(hip):
    p = P.config4(nmacro=3)
    ctx = hip.context().load(p)
    J, gu, ge, st = ctx.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p))
    yield p, ctx, J, gu, ge, st
    ctx.close()


def frozen_J(ctx, u0, emis):
    ctx.set_emissions(emis)
    ctx.set_state(u0)
    return ctx.replay_loss(0)


def test_fullsize_replay_bit_identical(run):
    p, ctx, J, gu, ge, st = run
    print(f"full-size adjoint: J={J:.6e} T {st.naccept_T} C {st.naccept_C}", flush=True)
    assert st.replay_maxrel == 0.0
    assert frozen_J(ctx, p.u0, p.emis) == J
    assert np.isfinite(gu).all() and np.isfinite(ge).all()


def _fd_best(ctx, p, pred, perturb, steps):
    best = np.inf
    for eps in steps:
        jp = frozen_J(ctx, *perturb(eps))
        jm = frozen_J(ctx, *perturb(-eps))
        best = min(best, abs((jp - jm) / (2 * eps) - pred) / abs(pred))
        if best <= 1e-7:
            break
    return best


def test_fullsize_gradient_u0_fd(run):
    p, ctx, J, gu, ge, st = run
    rng = np.random.default_rng(1004)
    worst = 0.0
    for d in range(NDIR):
        v = rng.standard_normal(p.u0.size) * p.u0
        pred = float(np.dot(gu, v))
        best = _fd_best(ctx, p, pred, lambda e: (p.u0 + e * v, p.emis), (1e-5, 1e-6, 1e-4, 1e-7))
        print(f"u0 direction {d}: rel {best:.2e}", flush=True)
        worst = max(worst, best)
    assert worst <= 1e-6, worst


def test_fullsize_gradient_emissions_fd(run):
    p, ctx, J, gu, ge, st = run
    rng = np.random.default_rng(1005)
    worst = 0.0
    for d in range(NDIR):
        v = rng.standard_normal(p.emis.size) * p.emis
        pred = float(np.dot(ge, v))
        best = _fd_best(ctx, p, pred, lambda e: (p.u0, p.emis + e * v), (1e-4, 1e-5, 1e-3, 1e-6))
        print(f"E direction {d}: rel {best:.2e}", flush=True)
        worst = max(worst, best)
    assert worst <= 1e-6, worst


def test_fullsize_emission_scale_identity(run):
    p, ctx, J, gu, ge, st = run
    pred = float(np.dot(p.emis, ge))
    eps = 1e-5
    jp = frozen_J(ctx, p.u0, p.emis * (1 + eps))
    jm = frozen_J(ctx, p.u0, p.emis * (1 - eps))
    assert abs((jp - jm) / (2 * eps) - pred) <= 1e-6 * abs(pred)
