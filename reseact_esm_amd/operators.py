"""The reference's operator API for the split step, on top of the C ABI.

Same names, argument meaning and error behaviour as the Julia plug-ins of
tools/reactant_handoff/rx_traced_integrator.jl and op_split.jl, so a caller of
the reference finds the calls it already makes:

  ssprk43_step(ctx, u, t, dt, abstol, reltol) -> (unew, EEst)      (:362-400)
  ros23_step(ctx, u, t, dt, abstol, reltol) -> (unew, EEst)        (:279-349, jac=:sym)
  adaptive_solve(step, ctx, u0, t0, tend, dt0, ctrl; maxiters, clamp_nonneg)
      -> (u, t, dt, nacc, nrej)                                    (:455-507)
  lie_trotter_solve(ctx, u0, tspan; macro_dt, reltols, abstols, clamp_nonneg)
      -> dict(u, retcode, naT, nrT, naC, nrC, nmacro)              (op_split.jl:100-164)
  ssprk43_step_vjp(ctx, u, t, dt, lam) -> (lam_in, None)           (:618-623)
  ros23_step_vjp(ctx, u, t, dt, lam) -> (lam_in, grad_emis)        (:610-617)

`ctx` is a loaded host.Context (grid, wind, mechanism, emissions already set);
the RHS callables `f`/`g` of the reference are replaced by that context.
Differences from the reference are deliberate and few: the Jacobian is the
exact analytic mass-action Jacobian (the reference's jac=:sym); maxiters
raises ReseactError(RS_E_MAXITERS) like host_adaptive! (adjoint_gradient.jl:580).
"""
from __future__ import annotations

import numpy as np

from .host import Context, PICtrl, SplitOpts, pictrl_ros23, pictrl_ssprk43  # noqa: F401


def ssprk43_step(ctx: Context, u, t, dt, abstol, reltol):
    ctx.set_state(u)
    e = ctx.ssprk43_step(t, dt, abstol, reltol)
    return ctx.get_candidate(), e


def ros23_step(ctx: Context, u, t, dt, abstol, reltol):
    ctx.set_state(u)
    e = ctx.ros23_step(t, dt, abstol, reltol)
    return ctx.get_candidate(), e


def adaptive_solve(step, ctx: Context, u0, t0, tend, dt0, ctrl: PICtrl, abstol=1e-6,
                   reltol=1e-3, maxiters=20000, clamp_nonneg=False):
    """`step` is ssprk43_step or ros23_step (the plug point of :455)."""
    ctx.set_state(u0)
    if step is ssprk43_step:
        dt, c = ctx.adaptive_transport(t0, tend, dt0, ctrl, abstol, reltol, clamp_nonneg, maxiters)
    elif step is ros23_step:
        dt, c = ctx.adaptive_chem(t0, tend, dt0, ctrl, abstol, reltol, clamp_nonneg, maxiters)
    else:
        raise ValueError("step must be ssprk43_step or ros23_step")
    return ctx.get_state(), c.t_end, dt, c.naccept, c.nreject


def lie_trotter_solve(ctx: Context, u0, tspan, macro_dt, reltols=(1e-3, 1e-3),
                      abstols=(1e-6, 1e-6), clamp_nonneg=True, dt0=(15.0, 0.5),
                      chem_percell=False, maxiters=20000):
    """Macro-step lattice t0 + k*macro_dt then tf (op_split.jl:142-156)."""
    t0, tf = float(tspan[0]), float(tspan[1])
    grid = list(np.arange(t0 + macro_dt, tf - 1e-9, macro_dt))
    stops = sorted(set(grid + [tf]))
    opts = SplitOpts(pictrl_ssprk43(), pictrl_ros23(), abstols[0], reltols[0], abstols[1],
                     reltols[1], 1 if clamp_nonneg else 0, 1 if chem_percell else 0, maxiters)
    ctx.set_state(u0)
    dtT, dtC = dt0
    tcur = t0
    naT = nrT = naC = nrC = nmacro = 0
    for tnext in stops:
        if tnext <= tcur + 1e-9:
            continue
        dtT, dtC, cT, cC = ctx.macro_step(tcur, tnext, dtT, dtC, opts)
        naT += cT.naccept; nrT += cT.nreject; naC += cC.naccept; nrC += cC.nreject
        tcur = tnext
        nmacro += 1
    return dict(u=ctx.get_state(), retcode="Success", naT=naT, nrT=nrT, naC=naC, nrC=nrC,
                nmacro=nmacro)


def ssprk43_step_vjp(ctx: Context, u, t, dt, lam):
    ctx.set_state(u)
    ctx.set_lambda(lam)
    ctx.ssprk43_vjp(t, dt)
    return ctx.get_lambda(), None


def ros23_step_vjp(ctx: Context, u, t, dt, lam):
    ctx.set_state(u)
    ctx.set_lambda(lam)
    ctx.zero_grad()
    ctx.ros23_vjp(t, dt)
    return ctx.get_lambda(), ctx.get_grad_emis()
