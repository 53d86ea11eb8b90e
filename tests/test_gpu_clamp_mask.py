"""The backward sweep's clamp masks: the chemistry VJP of step q multiplies
lambda by 1[u_{q+1} > 0], and the sweep hands it the sign bits of u_{q+1}
written by the VJP of step q + 1 (whose input it is) instead of u_{q+1}
itself.  These tests run the sweep with the sign bits and with every VJP
reading its own step output (RS_VJP_SIGN_BITS=0), and compare a case where
the mask zeroes whole species with the oracle, whose VJPs read each step's
output (adjoint_gradient.jl:731-784).  The two VJP kernels are compiled
separately and round a few contracted operations differently, so the two
sweeps agree to rounding (measured: 1.7e-16 of the largest gradient), not
bit for bit; a wrong mask entry would change a gradient by O(1)."""
import os

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import split_opts

pytestmark = pytest.mark.gpu

TOL_G = 1e-6


def relnorm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def adjoint(lib, p, sign_bits=True):
    key = "RS_VJP_SIGN_BITS"
    old = os.environ.get(key)
    os.environ[key] = "1" if sign_bits else "0"
    try:
        with lib.context() as c:
            c.load(p)
            return c.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p))
    finally:
        if old is None:
            os.environ.pop(key, None)
        else:
            os.environ[key] = old


def zero_species(p, s):
    """Species s set to zero everywhere with no emission of it.  The random
    mechanism never produces it (checked), so it stays exactly zero: its mask
    entry is 0 after every step and the sweep must zero its lambda."""
    assert all(v < 0 for r in p.mech.net() for q, v in r if q == s)
    u0 = p.u0.reshape(p.ns, p.nc).copy()
    u0[s] = 0.0
    p.u0 = u0.reshape(-1)
    if p.emis is not None:
        e = p.emis.reshape(p.ns, -1).copy()
        e[s] = 0.0
        p.emis = e.reshape(p.emis.shape)
    return p


def test_sign_bits_equal_reading_outputs_full_grid(hip):
    """144x91x47, 3 split steps: the same loss, gradients equal to rounding,
    and the never-produced species masked to exactly zero either way."""
    p = zero_species(P.config4(nmacro=3), 18)
    print("case: 144x91x47 x 3 split steps, species 18 zero", flush=True)
    Ja, gua, gea, sa = adjoint(hip, p, True)
    Jb, gub, geb, sb = adjoint(hip, p, False)
    assert Ja == Jb
    assert np.max(np.abs(gua - gub)) <= 1e-13 * np.max(np.abs(gub))
    assert np.max(np.abs(gea - geb)) <= 1e-13 * np.max(np.abs(geb))
    assert np.all(gua.reshape(p.ns, p.nc)[18] == 0.0)
    assert np.all(gub.reshape(p.ns, p.nc)[18] == 0.0)


def test_zero_species_masked_like_oracle(hip, oracle):
    p = zero_species(P.config4(nlon=12, nlat=7, nlev=5, nmacro=2), 18)
    print("case: 12x7x5 x 2 split steps, species 18 zero", flush=True)
    Jg, gug, geg, sg = adjoint(hip, p, True)
    with oracle.context() as co:
        co.load(p)
        Jo, guo, geo, so = co.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p))
    assert (sg.naccept_T, sg.naccept_C) == (so.naccept_T, so.naccept_C)
    assert abs(Jg - Jo) <= 1e-8 * abs(Jo)
    assert relnorm(gug, guo) <= TOL_G
    assert relnorm(geg, geo) <= TOL_G
    assert np.all(gug.reshape(p.ns, p.nc)[18] == 0.0)
    assert np.all(guo.reshape(p.ns, p.nc)[18] == 0.0)
