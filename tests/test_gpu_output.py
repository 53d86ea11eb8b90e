"""GPU: the run loop with streaming Zarr output and the STATEDUMP restart on
the HIP library.  The records hold exactly the states the loop reported; they
agree with the oracle's run to the split-step bar; a failed run restarted
from its dump reproduces the run that never failed, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from reseact_esm_amd import operators as O
from reseact_esm_amd import output as OUT
from reseact_esm_amd import problems as P

pytestmark = pytest.mark.gpu


def linear_problem():
    p = P.config1(nlon=24, nlat=12, nlev=6)
    rng = np.random.default_rng(5)
    a = rng.integers(0, p.ns, 20)
    b = (a + 1 + rng.integers(0, p.ns - 1, 20)) % p.ns
    none = -np.ones(20, np.int64)
    p.mech = P.Mechanism(ns=p.ns, reactants=np.stack([a, none], 1).astype(np.int32),
                         products=np.stack([b, none], 1).astype(np.int32),
                         k=10 ** rng.uniform(-4, 1, 20), photo=np.zeros(20, np.int32))
    return p


def test_run_with_output_and_restart(hip, oracle, tmp_path):
    p = linear_problem()
    g = p.grid
    names = [f"S{s:02d}" for s in range(p.ns)]
    times = [0.0, 1200.0, 2400.0]
    res = {}
    for tag, lib in (("hip", hip), ("oracle", oracle)):
        print(f"simulate {tag}", flush=True)
        ctx = lib.context().load(p)
        store = str(tmp_path / f"{tag}.zarr")
        sink = OUT.ZarrSink(store, g, names, times).open()
        seen = {0.0: p.u0}
        r = O.simulate(ctx, p.u0, (0.0, 2400.0), 600.0, sink=sink, out_every=2,
                       report=lambda t, u: seen.__setitem__(t, u.copy()))
        sink.close()
        assert r["ok"] and r["nstep"] == 4
        recs = np.stack([OUT.read_zarr_array(store, nm) for nm in names], 1)
        for i, t in enumerate(times):
            assert np.array_equal(recs[i].reshape(-1), seen[t]), (tag, t)
        res[tag] = (r, recs, ctx)
    (rh, ah, ch), (ro, ao, co) = res["hip"], res["oracle"]
    assert (rh["nT"], rh["nC"]) == (ro["nT"], ro["nC"])
    assert float(np.max(np.abs(ah - ao) / np.maximum(np.abs(ao), 1e-6))) <= 1e-8

    print("statedump restart", flush=True)
    path = str(tmp_path / "fail.bin")
    bad = O.simulate(ch, p.u0, (0.0, 2400.0), 600.0, min_macro_dt=600.0, maxiters=40,
                     statedump=path)
    assert not bad["ok"] and bad["retcode"] == "MaxIters"
    d = OUT.read_statedump(path)
    r = O.simulate(ch, d.u_pre, (d.t, 2400.0), 600.0)
    assert r["ok"] and np.array_equal(r["u"], rh["u"])
    ch.close()
    co.close()
