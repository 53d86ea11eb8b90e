"""GPU per-cell chemistry adjoint (rs_chem_percell_adj through rs_chem_percell_vjp
and rs_adjoint_run with per-cell chemistry) against the oracle, same seeded
inputs on both libraries, through the C ABI.  The oracle's per-cell adjoint is
pinned on the CPU (test_oracle_percell_adjoint.py)."""
from __future__ import annotations

import numpy as np
import pytest

from percell_adj import chemistry_only, compose_adjoint
from reseact_esm_amd import problems as P
from reseact_esm_amd.host import RS_E_NONFINITE, ReseactError, pictrl_ros23, split_opts

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def weights(p, seed):
    return np.random.default_rng(seed).uniform(0.5, 1.5, p.ns * p.nc)


@pytest.mark.parametrize("clamp", [True, False])
def test_percell_vjp_matches_oracle(hip, oracle, clamp):
    p = chemistry_only(P.config4(nlon=12, nlat=7, nlev=5))
    w = weights(p, 21)
    ts = p.tstops()
    out = []
    for lib in (hip, oracle):
        with lib.context() as c:
            c.load(p)
            c.set_lambda(w)
            c.zero_grad()
            st = c.chem_percell_vjp(ts[0], ts[1], p.dt0_C, pictrl_ros23(), p.atol_C, p.rtol_C,
                                    clamp)
            out.append((c.get_lambda(), c.get_grad_emis(), st, c.get_state()))
    (lg, gg, sg, ug), (lo, go, so, uo) = out
    print(f"per-cell VJP {p.grid}: attempts {sg.attempts_total} (oracle {so.attempts_total}), "
          f"lambda rel {relnorm(lg, lo):.2e}, grad E rel {relnorm(gg, go):.2e}", flush=True)
    assert (sg.attempts_total, sg.accepts_total, sg.attempts_max) == \
           (so.attempts_total, so.accepts_total, so.attempts_max)
    assert np.array_equal(ug, p.u0)
    assert relnorm(lg, lo) <= 1e-8 and relnorm(gg, go) <= 1e-8


def test_percell_time_loop_adjoint_matches_oracle_and_composition(hip, oracle):
    p = P.config4(nlon=12, nlat=7, nlev=5, nmacro=2)
    w = weights(p, 22)
    opts = split_opts(p, percell=True)
    res = []
    for lib in (hip, oracle):
        with lib.context() as c:
            c.load(p)
            c.set_objective_weights(w)
            J, gu, ge, st = c.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, opts, loss_kind=1)
            seqT = [c.recorded_steps(m, 0) for m in range(p.nmacro)]
            res.append((J, gu, ge, st, seqT))
            if lib is hip:
                c.set_state(p.u0)
                Jr = c.replay_loss(loss_kind=1)
                Jc, guc, gec = compose_adjoint(c, p, opts, seqT, w)
    (Jg, gug, geg, sg, qg), (Jo, guo, geo, so, qo) = res
    print(f"per-cell adjoint {p.grid} x {p.nmacro}: J {Jg:.12e} (oracle {Jo:.12e}), grad u0 rel "
          f"{relnorm(gug, guo):.2e}, grad E rel {relnorm(geg, geo):.2e}, VJPs {sg.nvjp_C} "
          f"(oracle {so.nvjp_C})", flush=True)
    # the same transport steps: t and dt to 1e-12 of the split step (the
    # global norm is summed in another order; the last step of a macro step,
    # t1 - t, is short, so its own relative difference is larger)
    for m, (a, b) in enumerate(zip(qg, qo)):
        assert a.shape == b.shape
        assert np.max(np.abs(a - b)) <= 1e-12 * p.dt_macro, m
    assert sg.nvjp_C == so.nvjp_C and sg.naccept_C == so.naccept_C
    assert sg.replay_maxrel == 0.0 and sg.bad_op == -1
    assert abs(Jg - Jo) <= 1e-10 * abs(Jo)
    assert relnorm(gug, guo) <= 1e-8 and relnorm(geg, geo) <= 1e-8
    # the GPU's time loop is its own operator VJPs composed
    assert abs(Jr - Jg) <= 1e-13 * abs(Jg) and abs(Jc - Jg) <= 1e-13 * abs(Jg)
    assert relnorm(gug, guc) <= 1e-12 and relnorm(geg, gec) <= 1e-12


def test_percell_nonfinite_report_matches_oracle(hip, oracle):
    p = P.config4(nlon=8, nlat=5, nlev=4, nmacro=2)
    w, _ = P.objective_weights(p.grid, p.ns, 2, "surf")
    w[2 * p.grid.nc + 3] = np.nan
    rep = []
    for lib in (hip, oracle):
        with lib.context() as c:
            c.load(p)
            c.set_objective_weights(w)
            with pytest.raises(ReseactError) as e:
                c.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p, percell=True),
                              loss_kind=1)
            assert e.value.code == RS_E_NONFINITE
            print(f"{lib.backend}: {e.value}", flush=True)
            rep.append(e.value.stats)
    g, o = rep
    assert (g.bad_op, g.bad_macro, g.bad_step, g.bad_nsteps, g.bad_t, g.bad_dt) == \
           (o.bad_op, o.bad_macro, o.bad_step, o.bad_nsteps, o.bad_t, o.bad_dt) == \
           (1, 1, 0, 1, p.tstops()[1], p.dt_macro)
    assert g.bad_lam_in == o.bad_lam_in == 1 and g.bad_grad >= 1 and g.bad_lam_out >= 1


def test_percell_adjoint_mid_grid_against_oracle(hip, oracle):
    """36x24x12, 20 species, 3 split steps of 900 s with per-cell chemistry:
    gradients against the oracle's (each cell takes the same steps on both)."""
    p = P.config4(nlon=36, nlat=24, nlev=12, nmacro=3)
    opts = split_opts(p, percell=True)
    res = []
    for lib in (hip, oracle):
        with lib.context() as c:
            c.load(p)
            res.append(c.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, opts))
    (Jg, gug, geg, sg), (Jo, guo, geo, so) = res
    print(f"per-cell adjoint {p.grid} x {p.nmacro}: J rel {abs(Jg - Jo) / Jo:.2e}, grad u0 "
          f"{relnorm(gug, guo):.2e}, grad E {relnorm(geg, geo):.2e}, per-cell VJPs {sg.nvjp_C} "
          f"(oracle {so.nvjp_C}), backward {sg.seconds_backward:.3f} s", flush=True)
    assert sg.nvjp_C == so.nvjp_C
    assert abs(Jg - Jo) <= 1e-10 * abs(Jo)
    assert relnorm(gug, guo) <= 1e-7 and relnorm(geg, geo) <= 1e-7
