"""GPU: the autotuner and the graph-captured ensemble (forward + adjoint of
several differently sized regional domains) against per-domain oracle runs."""
from __future__ import annotations

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import Ensemble, split_opts

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def domains():
    return P.config5(sizes=[(16, 9, 5), (24, 7, 6), (12, 11, 4)], nmacro=2)


def test_autotune_keeps_results(hip, oracle):
    p = P.config4(nlon=40, nlat=12, nlev=9, nmacro=1)
    cg = hip.context().load(p)
    r = cg.autotune()
    assert r.transport_tile_levels in (4, 8, 16, 32)
    assert r.chem_block in (64, 128, 256)
    assert all(t > 0 for t in r.transport_ms)
    # a workgroup size whose LDS slab does not fit the device is skipped (0)
    assert r.chem_ms[0] > 0 and all(t >= 0 for t in r.chem_ms)
    co = oracle.context().load(p)
    opts = split_opts(p)
    rg = cg.macro_step(0.0, 900.0, 15.0, 0.5, opts)
    ro = co.macro_step(0.0, 900.0, 15.0, 0.5, opts)
    assert (rg[2].naccept, rg[3].naccept) == (ro[2].naccept, ro[3].naccept)
    a, b = cg.get_state(), co.get_state()
    # relative to max(|b|, atol): entries below the absolute tolerance compare at atol
    assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6)) <= 1e-8


@pytest.mark.parametrize("budget_gib", [64.0, 1e-6])
def test_ensemble_adjoint_matches_oracle(hip, oracle, budget_gib):
    probs = domains()
    members = [hip.context().load(p) for p in probs]
    ens = Ensemble(members)
    tunes = ens.autotune()
    assert len(tunes) == len(members)
    p0 = probs[0]
    opts = split_opts(p0)
    J, st, es = ens.adjoint_run(p0.tstops(), p0.dt0_T, p0.dt0_C, opts, 0, budget_gib)
    if budget_gib < 1:
        assert es.backward_groups == len(members)
    else:
        assert es.backward_groups == 1
    assert es.graph_launches_forward >= 2 and es.graph_nodes_backward > 0
    for d, p in enumerate(probs):
        co = oracle.context().load(p)
        Jo, guo, geo, so = co.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, opts)
        assert (st[d].naccept_T, st[d].naccept_C) == (so.naccept_T, so.naccept_C)
        assert abs(J[d] - Jo) <= 1e-8 * Jo
        assert relnorm(members[d].get_lambda(), guo) <= 1e-6
        assert relnorm(members[d].get_grad_emis(), geo) <= 1e-6
        assert st[d].replay_maxrel < 1e-12


def test_ensemble_adjoint_ppm(hip, oracle):
    """Members on the PPM transport (regional patches of three sizes with open
    walls, random inflow halos and the three-reaction mechanism; one with the
    PBL mixing term): the
    graph-captured forward and backward sweeps against each member's oracle
    adjoint, loss on the first mixing ratio."""
    probs = [P.ppm_coupled_regional(nlon, nlat, nlev, seed=s)
             for nlon, nlat, nlev, s in ((10, 8, 6, 5), (12, 6, 7, 6), (9, 9, 5, 7))]
    # the second member also mixes its boundary layer
    P.with_pbl_mixing(probs[1], pblh_mean=800.0, pblh_amp=400.0, tau=0.05)
    members = [hip.context().load(p) for p in probs]
    ens = Ensemble(members)
    p0 = probs[0]
    opts = split_opts(p0)
    J, st, es = ens.adjoint_run(p0.tstops(), p0.dt0_T, p0.dt0_C, opts, 1, 64.0)
    assert es.graph_launches_forward >= 2 and es.graph_nodes_backward > 0
    for d, p in enumerate(probs):
        co = oracle.context().load(p)
        Jo, guo, geo, so = co.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, opts, loss_species=1)
        assert (st[d].naccept_T, st[d].nreject_T, st[d].naccept_C) == \
            (so.naccept_T, so.nreject_T, so.naccept_C)
        assert abs(J[d] - Jo) <= 1e-10 * abs(Jo)
        assert relnorm(members[d].get_lambda(), guo) <= 1e-9
        assert st[d].replay_maxrel == 0.0
        co.close()
