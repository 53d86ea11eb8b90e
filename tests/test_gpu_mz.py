"""GPU: the vertical air-mass flux diagnosed from continuity with the column
pressure fixer (rs_diagnose_mz; reseact.esm:262-291) on the HIP library
against the oracle.  Both evaluate the same statements in the same order
without contraction, so the diagnosed Mz and the lid residual are equal bit
for bit; with the pressure fixer of the same dPSdt, a PPM attempt on the
diagnosed flux is bit-for-bit the oracle's, and the air mass follows
dm/dt = dB dPSdt on the GPU too.  Cases: a regional patch with open walls and
a periodic global grid, random per-level fluxes and layer widths, and a
global grid taller than a stage workgroup's level block.
"""
from __future__ import annotations

import numpy as np
import pytest

from test_oracle_mz import random_flux_problem

pytestmark = pytest.mark.gpu

CASES = {
    "regional-11x7x9": (False, 51, dict()),
    "global-11x7x9": (True, 52, dict()),
    "global-64x32x53": (True, 53, dict(nlon=64, nlat=32, nlev=53, nq=3)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_diagnosed_mz_bitwise(hip, oracle, name):
    periodic, seed, kw = CASES[name]
    print(f"diagnose mz {name}", flush=True)
    prob, dB, dps = random_flux_problem(periodic, seed, **kw)
    out = []
    for lib in (hip, oracle):
        ctx = lib.context().load(prob)
        Mz, resid = ctx.diagnose_mz(dB, dps)
        ctx.set_pressure_fixer(dps)
        dt = 1e-3
        eest = ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
        out.append((Mz, resid, eest, ctx.get_candidate(), ctx.get_state()))
        ctx.close()
    (mg, rg, eg, cg, ug), (mo, ro, eo, co, uo) = out
    assert np.array_equal(mg, mo), float(np.max(np.abs(mg - mo)))
    assert rg == ro
    assert np.array_equal(cg, co), float(np.max(np.abs(cg - co)))
    assert abs(eg - eo) <= 1e-12 * max(eo, 1e-300)
    g, nc = prob.grid, prob.nc
    # air mass: (m_new - m)/dt = dB dPSdt (the air-mass row does not depend on
    # the state), up to the rounding of m_new - m
    dm = (cg[:nc] - ug[:nc]).reshape(g.nlev, g.nlat, g.nlon) / 1e-3
    want = dB[:, None, None] * dps[None]
    tol = 1e-12 * np.max(np.abs(ug[:nc])) / 1e-3 + 1e-12 * np.max(np.abs(mg))
    assert np.max(np.abs(dm - want)) <= tol
