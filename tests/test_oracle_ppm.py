"""Pins of the oracle's monotone PPM transport (CWC mixing-ratio form),
against the reference's own test and what the mathematics fixes.  CPU only.

* The reference's CWC gate (prototypes/transport_3d/transport_3d.esm:269-300,
  freestream_q_stays_one_regional_inflow): on the 7x7x7 CONUS patch with
  open walls, inflow halos at 1 and a divergent air-mass flux, q stays
  EXACTLY 1 (dev = 0) through an adaptive solve to t = 0.1 while m changes;
  its falsifier (every halo x 1.000001) moves q off 1.
* Face values, read back from the tendency with a single non-zero face flux:
  - a step is transported by its upwind value (no overshoot), linear data by
    the exact edge value, and the CW84 worked case (0,1,2,3,2,1) gives 8/3;
  - cell averages of a cubic give the exact edge value, on uniform cells
    (lat) and on non-uniform layer widths (the weighted eq. 1.6/1.7 form);
  - the weighted form with equal widths equals the uniform form.
* Conservation on the global grid: sum of A*dm/dt and of A*d(m q)/dt vanish
  to round-off (periodic longitude, zero-length polar faces, no flux through
  surface and lid); the air mass is conserved through a solve.
* The column pressure fixer: the air-mass column tendency is dPS/dt for any
  fluxes; with Mz from the corrected continuity, m is stationary.
* Monotonicity: a square wave carried by a uniform zonal flow makes no new
  extrema beyond 1e-10 of the jump (the 3rd-order upwind scheme of the
  synthetic workloads overshoots by percents on the same profile).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import ReseactError, pictrl_ssprk43

_D = C.POINTER(C.c_double)


def rhs(ctx, u):
    u = np.ascontiguousarray(u, np.float64)
    f = np.empty_like(u)
    ctx._chk(ctx.lib.rs_oracle_transport_rhs(ctx.h, u.ctypes.data_as(_D), f.ctypes.data_as(_D)))
    return f


# ------------------------------------------------------------------ CWC gate
def test_cwc_gate_and_falsifier(oracle):
    p = P.transport3d_regional()
    ctx = oracle.context().load(p)
    nc = p.nc
    assert np.array_equal(rhs(ctx, p.u0)[nc:], np.zeros(nc))        # dev'(0) = 0
    dt, c = ctx.adaptive_transport(0.0, 0.1, 0.01, pictrl_ssprk43(), 1e-6, 1e-3, 0)
    u = ctx.get_state()
    assert c.naccept >= 1 and abs(c.t_end - 0.1) < 1e-12
    assert np.all(u[nc:] == 1.0)                                      # q EXACTLY 1
    assert np.max(np.abs(u[:nc] - p.u0[:nc])) > 1e-3                  # m moved
    ctx.close()
    pf = P.transport3d_regional(qbc_scale=1.000001)
    ctx = oracle.context().load(pf)
    ctx.adaptive_transport(0.0, 0.1, 0.01, pictrl_ssprk43(), 1e-6, 1e-3, 0)
    dev = ctx.get_state()[nc:] - 1.0
    assert np.max(np.abs(dev)) > 1e-8                                 # the falsifier fails the gate
    ctx.close()


# ------------------------------------------------------------- face values
def _line_problem(n, axis, dp=None):
    """A line of n cells along `axis` ("lat" or "lev") with unit metrics, m = 1,
    one mixing ratio, and a single unit face flux at face 3 (set later)."""
    if axis == "lat":
        g = P.Grid(1, n, 1)
    else:
        g = P.Grid(1, 1, n)
    ce = np.ones(g.nlat + 1)
    dS = np.ones(g.nlat)
    dphi = np.ones(g.nlat)
    Mx = np.zeros((g.nlev, g.nlat, 2))
    My = np.zeros((g.nlev, g.nlat + 1, 1))
    Mz = np.zeros((g.nlev + 1, g.nlat, 1))
    dpa = np.ones((g.nlev, g.nlat, 1)) if dp is None else np.asarray(dp, float).reshape(g.nlev, g.nlat, 1)
    ppm = P.PPMSetup(True, 1.0, ce, dS, dphi, Mx, My, Mz, dpa)
    return g, ppm


def face_value(oracle, q, axis, face, M=1.0, dp=None):
    """Face value of the PPM reconstruction of `q` at `face` (between cells
    face-1 and face) from the tendency of cell `face` (M > 0) or face-1."""
    n = len(q)
    g, ppm = _line_problem(n, axis, dp)
    if axis == "lat":
        ppm.My[0, face, 0] = M
    else:
        ppm.Mz[face, 0, 0] = M
    p = P.Problem(name="line", grid=g, ns=2, nt=2, wind=P.solid_body_wind(g, 1.0, 0.0), mech=None,
                  u0=np.concatenate([np.ones(n), q]), ppm=ppm)
    ctx = oracle.context().load(p)
    f = rhs(ctx, p.u0)[n:]
    ctx.close()
    # with m = 1 and only this face carrying flux, the cell downstream of the
    # face, r, has div(M) = -|M| and div(M q_f) = -|M| q_f, so
    # dq_r/dt = |M| (q_f - q_r)
    r = face if M > 0 else face - 1
    return f[r] / abs(M) + q[r]


@pytest.mark.parametrize("axis", ["lat", "lev"])
def test_face_values_closed_forms(oracle, axis):
    step = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    assert face_value(oracle, step, axis, 3) == 0.0               # upwind value, no overshoot
    assert face_value(oracle, step, axis, 3, M=-1.0) == 1.0
    lin = np.arange(8, dtype=float)
    assert face_value(oracle, lin, axis, 3) == 2.5                 # exact edge of linear data
    assert face_value(oracle, lin, axis, 4, M=-1.0) == 3.5
    worked = np.array([-1.0, 0.0, 1.0, 2.0, 3.0, 2.0, 1.0, 0.0])
    # upwind cell 3 (mean 2) between 1 and 3; its right neighbour is an extremum
    assert abs(face_value(oracle, worked, axis, 4) - 8.0 / 3.0) < 1e-15


def _cubic_averages(edges):
    Q = lambda x: 0.5 + 0.3 * x + 0.05 * x ** 2 + 0.002 * x ** 3     # monotone on [0, 10]
    I = lambda x: 0.5 * x + 0.15 * x ** 2 + 0.05 / 3 * x ** 3 + 0.0005 * x ** 4
    return Q, (I(edges[1:]) - I(edges[:-1])) / np.diff(edges)


def test_cubic_exact_uniform(oracle):
    edges = np.arange(11, dtype=float)
    Q, a = _cubic_averages(edges)
    for face in (3, 4, 5, 6, 7):
        for M in (1.0, -1.0):
            got = face_value(oracle, a, "lat", face, M)
            assert abs(got - Q(edges[face])) < 2e-14, (face, M, got, Q(edges[face]))


def test_cubic_exact_nonuniform_layers(oracle):
    widths = np.array([0.7, 1.3, 0.9, 1.6, 0.8, 1.1, 1.4, 0.6, 1.2, 1.0])
    edges = np.concatenate([[0.0], np.cumsum(widths)])
    Q, a = _cubic_averages(edges)
    for face in (3, 4, 5, 6, 7):
        for M in (1.0, -1.0):
            got = face_value(oracle, a, "lev", face, M, dp=widths)
            assert abs(got - Q(edges[face])) < 1e-13, (face, M, got, Q(edges[face]))
    # a non-uniform grid the uniform formula is not exact on: the weights matter
    g_unif = face_value(oracle, a, "lat", 5)
    assert abs(g_unif - Q(edges[5])) > 1e-6


def test_weighted_form_reduces_to_uniform(oracle):
    rng = np.random.default_rng(3)
    q = rng.random(10)
    for face in (3, 4, 5, 6, 7):
        for M in (1.0, -1.0):
            a = face_value(oracle, q, "lat", face, M)
            b = face_value(oracle, q, "lev", face, M, dp=np.full(10, 2.5))
            assert abs(a - b) <= 1e-15 * max(1.0, abs(a))


# ------------------------------------------------------------ conservation
def test_conservation_global(oracle):
    p = P.ppm_global(24, 12, 8, nq=3, seed=7)
    p.ppm.dPSdt = None          # the raw continuity: dm/dt = -div(M) alone
    ctx = oracle.context().load(p)
    nc, g, pp = p.nc, p.grid, p.ppm
    A = np.broadcast_to((pp.dlon * pp.dS_lat)[None, :, None], (g.nlev, g.nlat, g.nlon)).reshape(-1)
    rng = np.random.default_rng(1)
    u = p.u0.copy()
    u[:nc] *= rng.uniform(0.9, 1.1, nc)                 # m not in balance with the fluxes
    f = rhs(ctx, u)
    m, fm = u[:nc], f[:nc]
    scale = np.sum(A * np.abs(fm)) + 1e-300
    assert abs(np.sum(A * fm)) <= 1e-13 * scale
    for s in range(1, p.ns):
        q, fq = u[s * nc:(s + 1) * nc], f[s * nc:(s + 1) * nc]
        tm = A * (m * fq + q * fm)
        assert abs(np.sum(tm)) <= 1e-12 * np.sum(np.abs(tm)), s
    ctx.set_state(u)
    ctx.adaptive_transport(0.0, 3600.0, 60.0, pictrl_ssprk43(), 1e-6, 1e-3, 0)
    mt = ctx.get_state()[:nc]
    assert abs(np.sum(A * mt) - np.sum(A * m)) <= 1e-13 * np.sum(A * m)
    ctx.close()


def test_pressure_fixer(oracle):
    """The column pressure fixer (reseact.esm:5 caveat 4): the air-mass column
    tendency equals dPS/dt per column whatever the fluxes; with Mz diagnosed
    from the same corrected divergence and dPS/dt = 0 the air mass is
    stationary to round-off; the mixing-ratio tendencies do not see it."""
    p = P.ppm_global(24, 12, 8, nq=2, seed=5)
    g, pp, nc = p.grid, p.ppm, p.nc
    ctx = oracle.context().load(p)
    f = rhs(ctx, p.u0)
    fm = f[:nc].reshape(g.nlev, g.nlat, g.nlon)
    divscale = np.max(np.abs(np.diff(pp.Mx, axis=2)))
    assert np.max(np.abs(fm)) <= 1e-12 * divscale            # stationary (the fixer on)
    ctx.close()
    rng = np.random.default_rng(2)
    Mz = np.array(pp.Mz) * rng.uniform(0.5, 1.5, pp.Mz.shape)  # any Mz, zero at the ends
    dps = rng.normal(0.0, 1e-3, (g.nlat, g.nlon))
    pp2 = P.PPMSetup(**{**pp.__dict__, "Mz": Mz, "dPSdt": dps})
    p2 = P.Problem(**{**p.__dict__, "ppm": pp2})
    ctx = oracle.context().load(p2)
    f2 = rhs(ctx, p.u0)
    col = f2[:nc].reshape(g.nlev, g.nlat, g.nlon).sum(0)
    assert np.max(np.abs(col - dps)) <= 1e-12 * (np.max(np.abs(dps)) + divscale)
    ctx._chk(ctx.lib.rs_set_pressure_fixer(ctx.h, None))
    f3 = rhs(ctx, p.u0)
    assert np.array_equal(f3[nc:], f2[nc:])                   # mixing ratios: raw div(M)
    assert not np.array_equal(f3[:nc], f2[:nc])
    ctx.close()


def test_monotone_square_wave(oracle):
    g = P.Grid(48, 2, 2)
    ce, dS, dphi = np.ones(3), np.ones(2), np.ones(2)
    Mx = np.ones((2, 2, 49))
    My = np.zeros((2, 3, 48))
    Mz = np.zeros((3, 2, 48))
    dp = np.ones((2, 2, 48))
    q = np.zeros((2, 2, 48))
    q[:, :, 10:20] = 1.0
    ppm = P.PPMSetup(True, 1.0, ce, dS, dphi, Mx, My, Mz, dp)
    p = P.Problem(name="square", grid=g, ns=2, nt=2, wind=P.solid_body_wind(g, 1.0, 0.0), mech=None,
                  u0=np.concatenate([np.ones(g.nc), q.reshape(-1)]), ppm=ppm)
    ctx = oracle.context().load(p)
    ctx.adaptive_transport(0.0, 12.0, 0.1, pictrl_ssprk43(), 1e-8, 1e-6, 0)
    qt = ctx.get_state()[g.nc:]
    assert qt.max() <= 1.0 + 1e-10 and qt.min() >= -1e-10
    assert 0.5 < qt.reshape(2, 2, 48)[0, 0, 22:32].mean() < 1.0      # it moved ~12 cells
    ctx.close()
    # the synthetic workloads' upwind operator on the same profile overshoots
    w = P.solid_body_wind(g, 1.0, 0.0)
    w.cx[:] = 1.0
    pu = P.Problem(name="square-up", grid=g, ns=1, nt=1, wind=w, mech=None, u0=q.reshape(-1))
    cu = oracle.context().load(pu)
    cu.adaptive_transport(0.0, 12.0, 0.1, pictrl_ssprk43(), 1e-8, 1e-6, 0)
    qu = cu.get_state()
    assert qu.max() > 1.02 or qu.min() < -0.02
    cu.close()


def test_setup_errors(oracle):
    p = P.transport3d_regional()
    ctx = oracle.context()
    ctx.set_grid(p.grid, p.ns, p.nt)
    ctx.set_transport_scheme(1)
    bad = np.array(p.ppm.Mz)
    bad[-1, 0, 0] = 1e-3                              # flux through the lid
    with pytest.raises(ReseactError):
        ctx._chk(ctx.lib.rs_set_mass_fluxes(ctx.h, *[np.ascontiguousarray(a).ctypes.data_as(_D)
                                                     for a in (p.ppm.Mx, p.ppm.My, bad, p.ppm.dp)]))
    with pytest.raises(ReseactError):
        ctx.set_transport_scheme(7)
    ctx.set_state(p.u0)
    with pytest.raises(ReseactError):                # geometry and fluxes missing
        ctx.ssprk43_step(0.0, 0.01, 1e-6, 1e-3)
    ctx.close()
