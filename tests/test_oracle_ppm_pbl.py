"""Pins of the oracle's PBL mixing of the mixing ratios (the PPM tracer
tendency's pbl_f (pbl_mean - q) / tau, reseact.esm:1563-1690).  CPU only.

* Column conservation, the reference's own claim (reseact.esm:1563 comment):
  sum_k dp_k pbl_f_k (qbar - q_k) = 0 per column, so the PBL part of the
  tendency moves no tracer mass sum_k dp q.
* Closed form: with zero air-mass fluxes and pbl_f = 1 in the lowest layers,
  0 above, each PBL layer relaxes to the conserved PBL mean as
  qbar + (q0 - qbar) exp(-t / tau), and the layers above stay put; an
  adaptive SSPRK43 solve at tight tolerance follows it.
* The floor of the weight (max(sum dp pbl_f, 1)) on a column whose PBL holds
  less than 1 Pa: the mean is the plain weighted sum there.
* pbl_fraction against hand-computed layer overlaps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import pictrl_ssprk43

_D = C.POINTER(C.c_double)


def rhs(ctx, u):
    u = np.ascontiguousarray(u, np.float64)
    f = np.empty_like(u)
    ctx._chk(ctx.lib.rs_oracle_transport_rhs(ctx.h, u.ctypes.data_as(_D), f.ctypes.data_as(_D)))
    return f


def pbl_term(oracle, p):
    """The PBL part of the tendency: with it minus without it."""
    ctx = oracle.context().load(p)
    a = rhs(ctx, p.u0)
    ctx.set_pbl_mixing(None)
    b = rhs(ctx, p.u0)
    ctx.close()
    return a - b


def test_column_conservation(oracle):
    # 40 levels: the 1-4 km boundary layer spans several layers
    p = P.with_pbl_mixing(P.ppm_global(24, 10, 40, nq=2, seed=3), pblh_mean=2500.0, pblh_amp=1500.0)
    g, nc = p.grid, p.nc
    t = pbl_term(oracle, p)
    dp = p.ppm.dp.reshape(g.nlev, -1)
    assert np.max(np.abs(t[:nc])) == 0.0                 # the air mass has no PBL term
    for s in (1, 2):
        ts = t[s * nc:(s + 1) * nc].reshape(g.nlev, -1)
        col = np.sum(dp * ts, axis=0)
        scale = np.sum(np.abs(dp * ts), axis=0)
        assert np.all(scale > 0.0)
        # round-off of the two tendencies the term is taken from
        assert np.max(np.abs(col) / scale) <= 1e-12


def test_relaxation_closed_form(oracle):
    p = P.ppm_global(12, 6, 10, nq=1, seed=5)
    pp = p.ppm
    pp.Mx = np.zeros_like(pp.Mx)
    pp.My = np.zeros_like(pp.My)
    pp.Mz = np.zeros_like(pp.Mz)
    pp.dPSdt = None
    K = 4
    f = np.zeros_like(pp.dp)
    f[:K] = 1.0
    pp.pbl_f, pp.tau_pbl = f, 900.0
    g, nc = p.grid, p.nc
    q0 = p.u0[nc:].reshape(g.nlev, -1).copy()
    dp = pp.dp.reshape(g.nlev, -1)
    qbar = np.sum(dp[:K] * q0[:K], axis=0) / np.sum(dp[:K], axis=0)
    T = 600.0
    ctx = oracle.context().load(p)
    ctx.adaptive_transport(0.0, T, 10.0, pictrl_ssprk43(), 1e-12, 1e-9, False)
    q = ctx.get_state()[nc:].reshape(g.nlev, -1)
    ctx.close()
    want = q0.copy()
    want[:K] = qbar[None] + (q0[:K] - qbar[None]) * np.exp(-T / 900.0)
    assert np.max(np.abs(q - want) / np.abs(want)) <= 1e-7
    # above the PBL only the rounding of the stage combination (2/3 u + 1/3 u3)
    assert np.max(np.abs(q[K:] - q0[K:]) / np.abs(q0[K:])) <= 1e-13


def test_weight_floor(oracle):
    """A column whose PBL weight is below 1 Pa divides by 1, not by the weight
    (zero air-mass fluxes, so the tendency is the PBL term alone)."""
    p = P.ppm_global(8, 4, 6, nq=1, seed=7)
    pp = p.ppm
    pp.Mx = np.zeros_like(pp.Mx)
    pp.My = np.zeros_like(pp.My)
    pp.Mz = np.zeros_like(pp.Mz)
    pp.dPSdt = None
    g, nc = p.grid, p.nc
    f = np.zeros_like(pp.dp)
    f[:3] = 1.0
    f[:3, 1, 2] = 1e-8 / pp.dp[:3, 1, 2]          # that column's weight: 3e-8 Pa
    pp.pbl_f = f
    ctx = oracle.context().load(p)
    t = rhs(ctx, p.u0)
    ctx.close()
    assert np.max(np.abs(t[:nc])) == 0.0
    t = t[nc:].reshape(g.nlev, g.nlat, g.nlon)
    q = p.u0[nc:].reshape(g.nlev, g.nlat, g.nlon)
    w = pp.dp * f
    W = np.sum(w, axis=0)
    assert W[1, 2] < 1.0 and np.all(np.delete(W.reshape(-1), 1 * g.nlon + 2) > 1.0)
    mean = np.sum(w * q, axis=0) / np.maximum(W, 1.0)
    want = f * (mean[None] - q) / 900.0
    assert np.allclose(t, want, rtol=1e-13, atol=1e-300)
    # the floored column: the mean is the weighted sum itself, far below q
    assert abs(mean[1, 2] - 1e-8 * q[:3, 1, 2].sum()) <= 1e-20


def test_pbl_fraction_overlaps():
    dp = np.array([100.0, 200.0, 300.0, 400.0])[:, None, None]
    # PS = 1000 + 0 lid; layer edges (bottom, top): (1000, 900), (900, 700), (700, 400), (400, 0)
    T = 290.0
    p_top = 750.0                                    # choose pblh so that p_pbl_top = 750 Pa
    pblh = np.full((1, 1), -np.log(p_top / 1000.0) * P.RD_AIR * T / P.G_ACC)
    f = P.pbl_fraction(dp, pblh, T)[:, 0, 0]
    assert np.allclose(f, [1.0, (900.0 - 750.0) / 200.0, 0.0, 0.0], rtol=1e-12, atol=1e-15)
