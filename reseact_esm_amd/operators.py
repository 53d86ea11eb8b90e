"""The reference's operator API for the split step, on top of the C ABI.

Same names, argument meaning and error behaviour as the Julia plug-ins of
tools/reactant_handoff/rx_traced_integrator.jl and op_split.jl, so a caller of
the reference finds the calls it already makes:

  ssprk43_step(ctx, u, t, dt, abstol, reltol) -> (unew, EEst)      (:362-400)
  ros23_step(ctx, u, t, dt, abstol, reltol) -> (unew, EEst)        (:279-349, jac=:sym)
  adaptive_solve(step, ctx, u0, t0, tend, dt0, ctrl; maxiters, clamp_nonneg)
      -> (u, t, dt, nacc, nrej)                                    (:455-507)
  lie_trotter_solve(ctx, u0, tspan; macro_dt, reltols, abstols, refresh,
                    forcing_tstops, clamp_nonneg)
      -> dict(u, retcode, naT, nrT, naC, nrC, nmacro)              (op_split.jl:100-164)
  lie_trotter_solve_bisect(ctx, u0, tspan; macro_dt, min_macro_dt, on_bisect, ...)
      -> the same dict                                             (op_split.jl:219-242)
  simulate(ctx, u0, tspan, macro_dt; sink, out_every, flush_every, statedump, ...)
      -> dict(u, t, ok, retcode, nstep, nT, nC, nbisect)           (run_reseact.jl:339-406)
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

from .host import (RS_E_MAXITERS, RS_E_NONFINITE, RS_POSITIVE_DOMAIN, Context, PICtrl,  # noqa: F401
                   ReseactError, SplitOpts, pictrl_ros23, pictrl_ssprk43)


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


def forcing_stops(t0: float, tf: float, macro_dt: float, forcing_tstops=()):
    """Stop grid of one Lie-Trotter window (op_split.jl:122-144): the
    macro-step lattice t0 + k*macro_dt short of tf, the forcing cadence
    boundaries in the HALF-OPEN interval t0 < x <= tf (rounded to 6 digits),
    and tf itself, sorted and unique.  Returns (stops, fstops).

    The right end is included on purpose: a caller that drives one macro step
    at a time would otherwise drop every boundary that lands on a step edge
    (excluded as the end of one step and again as the start of the next), and
    the forcing would never refresh.  The strict left end keeps a boundary at
    t0 from firing twice."""
    t0, tf, macro_dt = float(t0), float(tf), float(macro_dt)
    if not macro_dt > 0.0:
        raise ValueError("macro_dt must be positive")
    fstops = {round(float(t), 6) for t in forcing_tstops if t0 + 1e-6 < float(t) <= tf + 1e-6}
    grid, k = [], 1
    while t0 + k * macro_dt <= tf - 1e-9:
        grid.append(t0 + k * macro_dt)
        k += 1
    return sorted(set(grid) | fstops | {tf}), fstops


# Inner-solve errors that end a window with a failed retcode instead of
# raising (the inner integrators' own failure retcodes in the reference).
_FAILURES = {RS_E_MAXITERS: "MaxIters", RS_E_NONFINITE: "Unstable"}


def lie_trotter_solve(ctx: Context, u0, tspan, macro_dt, reltols=(1e-3, 1e-3),
                      abstols=(1e-6, 1e-6), refresh=None, forcing_tstops=(),
                      clamp_nonneg=True, dt0=(15.0, 0.5), chem_percell=False, maxiters=20000,
                      positive_domain=False):
    """One Lie-Trotter window over `tspan` (op_split.jl:100-164).

    Steps through `forcing_stops`: transport then chemistry over each stop
    interval (rs_macro_step), then max(u, 0) on the state when
    `clamp_nonneg` (:151), then `refresh(t)` when the interval ended on a
    forcing boundary (:154), so the next interval sees fresh forcing.  The
    inner solvers start from `dt0` and carry their step sizes across the
    window.  `positive_domain` rejects negative chemistry trial states
    (RS_POSITIVE_DOMAIN), which the reference wants and cannot yet pass
    through its splitting package (HELPERS.md:396-447).

    An inner solve that runs out of iterations or goes non-finite ends the
    window: the result carries retcode "MaxIters" / "Unstable" and the state
    reached.  Returns dict(u, retcode, naT, nrT, naC, nrC, nmacro)."""
    t0, tf = float(tspan[0]), float(tspan[1])
    stops, fstops = forcing_stops(t0, tf, macro_dt, forcing_tstops)
    flags = RS_POSITIVE_DOMAIN if positive_domain else 0
    opts = SplitOpts(pictrl_ssprk43(), pictrl_ros23(), abstols[0], reltols[0], abstols[1],
                     reltols[1], flags, 1 if chem_percell else 0, maxiters)
    ctx.set_state(u0)
    dtT, dtC = dt0
    tcur = t0
    retcode = "Success"
    naT = nrT = naC = nrC = nmacro = 0
    for tnext in stops:
        if tnext <= tcur + 1e-9:
            continue
        try:
            dtT, dtC, cT, cC = ctx.macro_step(tcur, tnext, dtT, dtC, opts)
        except ReseactError as e:
            if e.code not in _FAILURES:
                raise
            retcode = _FAILURES[e.code]
            break
        naT += cT.naccept; nrT += cT.nreject; naC += cC.naccept; nrC += cC.nreject
        if clamp_nonneg:
            ctx.clamp_state()
        tcur = tnext
        nmacro += 1
        if refresh is not None and round(tcur, 6) in fstops:
            refresh(tcur)
    return dict(u=ctx.get_state(), retcode=retcode, naT=naT, nrT=nrT, naC=naC, nrC=nrC,
                nmacro=nmacro)


def _ok(r) -> bool:
    return r["retcode"] in ("Success", "Default")


def lie_trotter_solve_bisect(ctx: Context, u0, tspan, macro_dt, min_macro_dt=9.0,
                             on_bisect=None, **kw):
    """`lie_trotter_solve`, with a failed window retried as two half windows,
    recursively, until a window would be no longer than `min_macro_dt`; then
    the failure propagates (op_split.jl:219-242).  Each half starts its inner
    solvers afresh; `on_bisect(t0, tf)` is called before each split.  Counts
    are summed over the halves that succeeded (the failed attempt's work is
    not in them).  Call it one macro step at a time: a failed multi-step
    window is redone whole."""
    t0, tf = float(tspan[0]), float(tspan[1])
    r = lie_trotter_solve(ctx, u0, (t0, tf), macro_dt, **kw)
    if _ok(r) or (tf - t0) <= min_macro_dt + 1e-9:
        return r
    tm = 0.5 * (t0 + tf)
    if on_bisect is not None:
        on_bisect(t0, tf)
    r1 = lie_trotter_solve_bisect(ctx, u0, (t0, tm), min(float(macro_dt), tm - t0),
                                  min_macro_dt, on_bisect, **kw)
    if not _ok(r1):
        return r1
    r2 = lie_trotter_solve_bisect(ctx, r1["u"], (tm, tf), min(float(macro_dt), tf - tm),
                                  min_macro_dt, on_bisect, **kw)
    if not _ok(r2):
        return r2
    out = dict(u=r2["u"], retcode=r2["retcode"])
    for k in ("naT", "nrT", "naC", "nrC", "nmacro"):
        out[k] = r1[k] + r2[k]
    return out


def simulate(ctx: Context, u0, tspan, macro_dt, min_macro_dt=9.0, sink=None, out_every=1,
             flush_every=1, statedump: str | None = None, names=None, on_bisect=None,
             report=None, **kw):
    """The run drivers' macro-step loop (run_reseact.jl:339-406): one
    `lie_trotter_solve_bisect` call per macro step (so a bisection halves the
    failed step itself), the state written to `sink` at t0 and after every
    `out_every` steps, the sink flushed every `out_every * flush_every` steps
    and at the end.  A step that still fails stops the run without advancing
    t; with `statedump` the state it started from (restartable) and the state
    where it gave up are written there (run_reseact.jl:378-391).  `report(t,
    u)` runs after each step.  `kw` goes to lie_trotter_solve (tolerances,
    refresh, forcing_tstops, clamp_nonneg, ...).

    Returns dict(u, t, ok, retcode, nstep, nT, nC, nbisect)."""
    from .output import write_statedump

    t0, tend = float(tspan[0]), float(tspan[1])
    u = np.array(u0, np.float64, copy=True)
    nbisect = 0

    def bis(a, b):
        nonlocal nbisect
        nbisect += 1
        if on_bisect is not None:
            on_bisect(a, b)

    if sink is not None:
        sink.write(t0, u)
    t, nstep, nT, nC, ok, retcode = t0, 0, 0, 0, True, "Success"
    while t < tend - 1e-9:
        nstep += 1
        tnext = min(t + macro_dt, tend)
        res = lie_trotter_solve_bisect(ctx, u, (t, tnext), tnext - t, min_macro_dt=min_macro_dt,
                                       on_bisect=bis, **kw)
        if not _ok(res):
            ok, retcode = False, res["retcode"]
            if statedump:
                write_statedump(statedump, ctx.ns, ctx.nc, t, u, res["u"], names)
            break
        u = res["u"]
        nT += res["naT"]
        nC += res["naC"]
        t = tnext
        if report is not None:
            report(t, u)
        if sink is not None and nstep % out_every == 0:
            sink.write(t, u)
            if nstep % (out_every * flush_every) == 0:
                sink.flush()
    if sink is not None:
        sink.flush()
    return dict(u=u, t=t, ok=ok, retcode=retcode, nstep=nstep, nT=nT, nC=nC, nbisect=nbisect)


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
