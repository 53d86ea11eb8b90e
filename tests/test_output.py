"""Output and restart formats (reseact_esm_amd.output) and the run loop
(operators.simulate).  CPU only.

* STATEDUMP: the byte layout of run_reseact.jl:378-391 (Int64 NS, Int64 NC,
  Float64 t, then the pre-step and post-failure states CELL-major), checked
  against a hand-packed file; the `.names` companion; read-back is exact.
* A run that fails writes the dump of the step it could not take; restarting
  from the dump with more iterations reproduces, bit for bit, the run that
  never failed.
* The Zarr v3 sink: metadata per the format specification, one chunk per
  record, species-major arrays in the state's own cell order, records that
  read back exactly, NaN for records not written, the written count in the
  group attributes after each flush; the run loop writes t0 and every
  `out_every` steps.
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np
import pytest

from reseact_esm_amd import operators as O
from reseact_esm_amd import output as OUT
from reseact_esm_amd import problems as P


def test_statedump_layout(tmp_path):
    ns, nc, t = 3, 4, 5400.5
    u_pre = np.arange(ns * nc, dtype=np.float64)            # species-major
    u_post = -u_pre
    path = str(tmp_path / "dump.bin")
    OUT.write_statedump(path, ns, nc, t, u_pre, u_post, ["O3", "NO", "NO2"])
    raw = open(path, "rb").read()
    # hand-packed: header, then cell-major u[c*NS + s] = species-major u[s*NC + c]
    cm = [float(u_pre[s * nc + c]) for c in range(nc) for s in range(ns)]
    want = struct.pack("<qqd", ns, nc, t) + struct.pack(f"<{ns * nc}d", *cm) \
        + struct.pack(f"<{ns * nc}d", *[-x for x in cm])
    assert raw == want
    names = open(path + ".names").read().splitlines()
    assert names == ["# NS=3 NC=4 t=5400.5  layout: cell-major, u[(c-1)*NS+s]",
                     "1\tO3", "2\tNO", "3\tNO2"]
    d = OUT.read_statedump(path)
    assert (d.ns, d.nc, d.t, d.names) == (ns, nc, t, ["O3", "NO", "NO2"])
    assert np.array_equal(d.u_pre, u_pre) and np.array_equal(d.u_post, u_post)


def test_statedump_rejects_bad_files(tmp_path):
    path = str(tmp_path / "bad.bin")
    with open(path, "wb") as f:
        f.write(struct.pack("<qqd", 2, 3, 0.0) + b"\0" * 8)
    with pytest.raises(ValueError):
        OUT.read_statedump(path)
    with pytest.raises(ValueError):
        OUT.write_statedump(path, 2, 3, 0.0, np.zeros(5), np.zeros(6))


@pytest.fixture(scope="module")
def linear(oracle):
    """The 6x4x2 first-order-chemistry problem of the operator tests."""
    p = P.config1(nlon=6, nlat=4, nlev=2)
    rng = np.random.default_rng(5)
    a = rng.integers(0, p.ns, 20)
    b = (a + 1 + rng.integers(0, p.ns - 1, 20)) % p.ns
    none = -np.ones(20, np.int64)
    p.mech = P.Mechanism(ns=p.ns, reactants=np.stack([a, none], 1).astype(np.int32),
                         products=np.stack([b, none], 1).astype(np.int32),
                         k=10 ** rng.uniform(-4, 1, 20), photo=np.zeros(20, np.int32))
    ctx = oracle.context().load(p)
    yield p, ctx
    ctx.close()


def test_failed_run_dumps_and_restarts(linear, tmp_path):
    p, ctx = linear
    ok = O.simulate(ctx, p.u0, (0.0, 1800.0), 600.0)
    assert ok["ok"] and ok["nstep"] == 3
    # too few inner iterations for a 600 s step, and no bisection below 600 s
    path = str(tmp_path / "fail.bin")
    bad = O.simulate(ctx, p.u0, (0.0, 1800.0), 600.0, min_macro_dt=600.0, maxiters=60,
                     statedump=path)
    assert not bad["ok"] and bad["retcode"] == "MaxIters" and bad["t"] == 0.0
    d = OUT.read_statedump(path)
    assert d.t == 0.0 and np.array_equal(d.u_pre, p.u0)
    assert not np.array_equal(d.u_post, d.u_pre)
    # restart from the dump with the default iterations: the run that never failed
    r = O.simulate(ctx, d.u_pre, (d.t, 1800.0), 600.0)
    assert r["ok"] and np.array_equal(r["u"], ok["u"])


def test_zarr_sink_records(linear, tmp_path):
    p, ctx = linear
    g = p.grid
    names = [f"S{s:02d}" for s in range(p.ns)]
    store = str(tmp_path / "run.zarr")
    times = np.arange(0.0, 3600.0 + 1, 1200.0)               # out_every = 2 of 600 s steps
    sink = OUT.ZarrSink(store, g, names, times, meta={"title": "test"}).open()
    seen = {}
    r = O.simulate(ctx, p.u0, (0.0, 3600.0), 600.0, sink=sink, out_every=2,
                   report=lambda t, u: seen.__setitem__(t, u.copy()))
    sink.close()
    assert r["ok"] and r["nstep"] == 6
    grp = json.load(open(os.path.join(store, "zarr.json")))
    assert grp["zarr_format"] == 3 and grp["node_type"] == "group"
    assert grp["attributes"]["records_written"] == 4 and grp["attributes"]["title"] == "test"
    assert grp["attributes"]["species"] == names
    m = json.load(open(os.path.join(store, names[3], "zarr.json")))
    assert m["shape"] == [4, g.nlev, g.nlat, g.nlon]
    assert m["chunk_grid"] == {"name": "regular",
                               "configuration": {"chunk_shape": [1, g.nlev, g.nlat, g.nlon]}}
    assert m["dimension_names"] == ["time", "lev", "lat", "lon"]
    assert [c["name"] for c in m["codecs"]] == ["bytes", "gzip"]
    assert os.path.exists(os.path.join(store, names[3], "c", "2", "0", "0", "0"))
    assert np.array_equal(OUT.read_zarr_array(store, "time"), times)
    expect = {0.0: p.u0, 1200.0: seen[1200.0], 2400.0: seen[2400.0], 3600.0: seen[3600.0]}
    for s in (0, 3, p.ns - 1):
        a = OUT.read_zarr_array(store, names[s])
        for rec, t in enumerate(times):
            want = expect[t].reshape(p.ns, g.nlev, g.nlat, g.nlon)[s]
            assert np.array_equal(a[rec], want), (s, t)


def test_zarr_unwritten_records_are_nan(tmp_path):
    g = P.Grid(5, 3, 2)
    store = str(tmp_path / "part.zarr")
    sink = OUT.ZarrSink(store, g, ["A", "B"], [0.0, 10.0, 20.0], compress=False).open()
    u = np.arange(2 * g.nc, dtype=np.float64)
    sink.write(10.0, u)
    sink.flush()
    assert json.load(open(os.path.join(store, "zarr.json")))["attributes"]["written"] == [0, 1, 0]
    with pytest.raises(ValueError):
        sink.write(15.0, u)
    sink.close()
    a = OUT.read_zarr_array(store, "B")
    assert np.isnan(a[0]).all() and np.isnan(a[2]).all()
    assert np.array_equal(a[1].reshape(-1), u[g.nc:])
    with pytest.raises(FileExistsError):
        OUT.ZarrSink(store, g, ["A", "B"], [0.0]).open()
    with pytest.raises(ValueError):
        OUT.ZarrSink(str(tmp_path / "x"), g, ["A", "A"], [0.0])
