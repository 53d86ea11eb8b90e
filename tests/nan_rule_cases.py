"""Shared cases of the non-finite error-estimate rule: the reference treats a
NaN error estimate as EEst = 1e10, a rejection (rx_traced_integrator.jl:484),
so a state that stays NaN is rejected until maxiters
(adjoint_gradient.jl:580 raises), and per-cell solves keep it in its cell."""
from __future__ import annotations

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import RS_E_MAXITERS, ReseactError

NAN_CELL, NAN_SPECIES = 7, 3


def problem(nan: bool):
    p = P.config4(nlon=6, nlat=4, nlev=3)
    p.u0 = p.u0.copy()
    if nan:
        p.u0[NAN_SPECIES * p.nc + NAN_CELL] = np.nan
    return p


def global_norm_fails(lib, maxiters=25):
    """Adaptive chemistry with the global norm on a state holding one NaN:
    every attempt is rejected, the solve ends with RS_E_MAXITERS."""
    p = problem(True)
    with lib.context() as c:
        c.load(p)
        with pytest.raises(ReseactError) as e:
            c.adaptive_chem(0.0, 900.0, 0.5, None, 1e-6, 1e-3, True, maxiters)
        assert e.value.code == RS_E_MAXITERS


def percell_isolates(lib, maxiters=300):
    """Per-cell chemistry: the NaN cell alone hits maxiters; every other cell
    ends exactly where it ends without the NaN."""
    ref = None
    for nan in (False, True):
        p = problem(nan)
        with lib.context() as c:
            c.load(p)
            if nan:
                with pytest.raises(ReseactError) as e:
                    c.chem_percell(0.0, 900.0, 0.5, None, None, 1e-6, 1e-3, True, maxiters)
                assert e.value.code == RS_E_MAXITERS and "1 cells" in str(e.value)
            else:
                c.chem_percell(0.0, 900.0, 0.5, None, None, 1e-6, 1e-3, True, maxiters)
            u = c.get_state().reshape(p.ns, p.nc)
        if not nan:
            ref = u
            continue
        others = np.ones(p.nc, bool)
        others[NAN_CELL] = False
        assert np.array_equal(u[:, others], ref[:, others])
        assert np.isnan(u[NAN_SPECIES, NAN_CELL])
    return ref
