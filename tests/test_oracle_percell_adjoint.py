"""Pins of the oracle's per-cell chemistry adjoint (percell_vjp: the
reference's backward sweep, tools/adjoint_gradient.jl:731-834, over each cell's
own accepted sequence, rx_traced_integrator.jl:455-507 with NC = 1).  CPU only.

* On one cell the per-cell solve is the global-norm solve (the norm over one
  cell is the cell's norm), so the per-cell VJP and the per-cell time-loop
  adjoint must be bit for bit the global-norm ones, which the frozen-sequence
  finite-difference tests of test_oracle_adjoint.py pin.
* Cells are independent: the VJP over a grid is, cell by cell, the VJP of each
  cell run alone.
* The time-loop adjoint with per-cell chemistry is the composition of the
  operator VJPs (transport steps replayed from their record, the per-cell
  chemistry VJP per macro step), bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from percell_adj import chemistry_only, compose_adjoint, one_cell
from reseact_esm_amd import problems as P
from reseact_esm_amd.host import RS_E_NONFINITE, ReseactError, pictrl_ros23, split_opts


def small(nmacro=1):
    return P.config4(nlon=8, nlat=5, nlev=4, nmacro=nmacro)


def weights(p, seed=3):
    return np.random.default_rng(seed).uniform(0.5, 1.5, p.ns * p.nc)


@pytest.mark.parametrize("cell", [0, 3, 57, 131])
def test_one_cell_percell_adjoint_is_the_global_adjoint(oracle, cell):
    """One macro step: the per-cell mode starts every macro step's chemistry
    at dt0 in each cell, the global mode carries its proposed dt into the next
    macro step, so over several steps the two take other step sizes."""
    q = one_cell(small(nmacro=1), cell)
    w = weights(q)
    out = []
    for percell in (False, True):
        with oracle.context() as c:
            c.load(q)
            c.set_objective_weights(w)
            out.append(c.adjoint_run(q.tstops(), q.dt0_T, q.dt0_C, split_opts(q, percell=percell),
                                     loss_kind=1))
    (Jg, gug, geg, sg), (Jp, gup, gep, sp) = out
    assert Jg == Jp and np.array_equal(gug, gup) and np.array_equal(geg, gep)
    assert (sg.naccept_C, sg.nreject_C) == (sp.naccept_C, sp.nreject_C)
    assert sp.nvjp_C == sg.nvjp_C and sp.replay_maxrel == 0.0
    # the operator-level VJP over the first macro step agrees with the
    # global-norm sweep of that step
    ts = q.tstops()
    with oracle.context() as c:
        c.load(q)
        c.adaptive_chem(ts[0], ts[1], q.dt0_C, pictrl_ros23(), q.atol_C, q.rtol_C, True)
        steps = c.step_log(1)
        states = [q.u0.copy()]
        c.set_state(q.u0)
        for t, dt in steps:
            c.ros23_step(t, dt, q.atol_C, q.rtol_C)
            c.accept_candidate(True)
            states.append(c.get_state())
        lam = w.copy()
        c.zero_grad()
        for j in range(len(steps) - 1, -1, -1):
            lam = lam * (states[j + 1] > 0.0)
            c.set_state(states[j])
            c.set_lambda(lam)
            c.ros23_vjp(steps[j][0], steps[j][1])
            lam = c.get_lambda()
        ge_ref = c.get_grad_emis()
        c.set_state(q.u0)
        c.set_lambda(w)
        c.zero_grad()
        st = c.chem_percell_vjp(ts[0], ts[1], q.dt0_C, pictrl_ros23(), q.atol_C, q.rtol_C, True)
        assert st.accepts_total == len(steps)
        assert np.array_equal(c.get_lambda(), lam) and np.array_equal(c.get_grad_emis(), ge_ref)
        assert np.array_equal(c.get_state(), q.u0)   # the state is left as it was


def test_percell_vjp_cells_are_independent(oracle):
    p = chemistry_only(small())
    w = weights(p, seed=5)
    ts = p.tstops()
    with oracle.context() as c:
        c.load(p)
        c.set_lambda(w)
        c.zero_grad()
        st = c.chem_percell_vjp(ts[0], ts[1], p.dt0_C, pictrl_ros23(), p.atol_C, p.rtol_C, True)
        lam = c.get_lambda().reshape(p.ns, p.nc)
        ge = c.get_grad_emis().reshape(p.ns, p.grid.ncol)
    assert st.n_failed == 0 and st.attempts_max > st.attempts_min
    for cell in (0, 7, p.grid.ncol - 1, p.grid.ncol, 101, p.nc - 1):
        q = one_cell(p, cell)
        with oracle.context() as c:
            c.load(q)
            c.set_lambda(w.reshape(p.ns, p.nc)[:, cell])
            c.zero_grad()
            c.chem_percell_vjp(ts[0], ts[1], q.dt0_C, pictrl_ros23(), q.atol_C, q.rtol_C, True)
            assert np.array_equal(c.get_lambda(), lam[:, cell]), cell
            if cell < p.grid.ncol:
                assert np.array_equal(c.get_grad_emis(), ge[:, cell]), cell


def test_percell_time_loop_adjoint_is_the_composition(oracle):
    p = small(nmacro=2)
    w = weights(p, seed=9)
    opts = split_opts(p, percell=True)
    with oracle.context() as c:
        c.load(p)
        c.set_objective_weights(w)
        J, gu, ge, st = c.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, opts, loss_kind=1)
        seqT = [c.recorded_steps(m, 0) for m in range(p.nmacro)]
        assert all(len(c.recorded_steps(m, 1)) == 0 for m in range(p.nmacro))
        assert st.replay_maxrel == 0.0 and st.bad_op == -1
        c.set_state(p.u0)
        assert c.replay_loss(loss_kind=1) == J
        Jc, guc, gec = compose_adjoint(c, p, opts, seqT, w)
    # J summed in another order; the gradients bit for bit
    assert abs(J - Jc) <= 1e-13 * abs(J)
    assert np.array_equal(gu, guc) and np.array_equal(ge, gec)
    assert st.nvjp_C > 0 and np.isfinite(gu).all() and np.abs(ge).max() > 0


def test_percell_nonfinite_is_reported_as_the_macro_steps_chemistry(oracle):
    p = small(nmacro=2)
    w, _ = P.objective_weights(p.grid, p.ns, 2, "surf")
    w[2 * p.grid.nc + 3] = np.nan
    with oracle.context() as c:
        c.load(p)
        c.set_objective_weights(w)
        with pytest.raises(ReseactError) as e:
            c.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p, percell=True), loss_kind=1)
    assert e.value.code == RS_E_NONFINITE
    s = e.value.stats
    ts = p.tstops()
    assert (s.bad_op, s.bad_macro, s.bad_step, s.bad_nsteps) == (1, p.nmacro - 1, 0, 1)
    assert (s.bad_t, s.bad_dt) == (ts[-2], ts[-1] - ts[-2])
    assert s.bad_lam_in == 1 and s.bad_lam_out >= 1 and s.bad_grad >= 1
