"""Output and restart formats of the reference's run drivers.

* STATEDUMP (run_reseact.jl:378-391): the state a failed macro step started
  from, restartable, plus the state where the integrator gave up.  Binary,
  little-endian, no padding:

      Int64 NS, Int64 NC, Float64 t, Float64 u[N] (pre-step), Float64 u[N] (post-failure)

  with both states CELL-major, u[c*NS + s], and a text companion
  `<path>.names`:

      # NS=<NS> NC=<NC> t=<t>  layout: cell-major, u[(c-1)*NS+s]
      <1-based position>\\t<species name>        (one line per species, in order)

  The library's state is species-major (u[s*NC + c]); the writer and reader
  transpose.

* Gridded streaming output (run_reseact.jl:257-271, :384-406; the sink
  protocol open / write / flush / close of HELPERS.md:236): a Zarr v3 store
  written to the format specification (no Zarr library is needed to write or
  read it).  One group, one float64 array per species named as in the state
  (species-major var_map order), shape [ntime, nlev, nlat, nlon] (longitude
  fastest: the state's own cell order), one chunk per output record so a
  record is written once and never rewritten, `time`, `lev`, `lat`, `lon`
  coordinate arrays, and gzip-compressed chunks (the reference's default is
  Blosc-zstd, which this image cannot encode; a Zarr reader decodes either
  from the codec list).  Records are compressed and written by a background
  thread, so the caller's next macro step overlaps the I/O; `flush` waits for
  them and records how many are complete in the group attributes, so a store
  cut short by a crash says which records are valid.  Unwritten records read
  as NaN (the array fill value).
"""
from __future__ import annotations

import gzip
import json
import os
import queue
import struct
import threading
from dataclasses import dataclass

import numpy as np

from .problems import Grid


# ----------------------------------------------------------------- STATEDUMP
@dataclass
class StateDump:
    ns: int
    nc: int
    t: float
    u_pre: np.ndarray      # species-major, restartable
    u_post: np.ndarray     # species-major, where the integrator gave up
    names: list[str]


def _to_cell_major(u: np.ndarray, ns: int, nc: int) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(u, np.float64).reshape(ns, nc).T).reshape(-1)


def _to_species_major(u: np.ndarray, ns: int, nc: int) -> np.ndarray:
    return np.ascontiguousarray(u.reshape(nc, ns).T).reshape(-1)


def write_statedump(path: str, ns: int, nc: int, t: float, u_pre, u_post,
                    names: list[str] | None = None) -> None:
    n = ns * nc
    u_pre = np.asarray(u_pre, np.float64)
    u_post = np.asarray(u_post, np.float64)
    if u_pre.size != n or u_post.size != n:
        raise ValueError(f"states must hold NS*NC = {n} values")
    names = names or [f"s{s}" for s in range(ns)]
    if len(names) != ns:
        raise ValueError("one name per species")
    with open(path, "wb") as f:
        f.write(struct.pack("<qqd", ns, nc, float(t)))
        f.write(_to_cell_major(u_pre, ns, nc).astype("<f8").tobytes())
        f.write(_to_cell_major(u_post, ns, nc).astype("<f8").tobytes())
    with open(path + ".names", "w") as f:
        f.write(f"# NS={ns} NC={nc} t={float(t)!r}  layout: cell-major, u[(c-1)*NS+s]\n")
        for s, nm in enumerate(names):
            f.write(f"{s + 1}\t{nm}\n")


def read_statedump(path: str) -> StateDump:
    with open(path, "rb") as f:
        head = f.read(24)
        if len(head) != 24:
            raise ValueError("truncated STATEDUMP header")
        ns, nc, t = struct.unpack("<qqd", head)
        if ns <= 0 or nc <= 0:
            raise ValueError(f"bad STATEDUMP header NS={ns} NC={nc}")
        n = ns * nc
        body = np.frombuffer(f.read(16 * n), dtype="<f8")
        if body.size != 2 * n or f.read(1):
            raise ValueError("STATEDUMP body is not 2*NS*NC doubles")
    names = [f"s{s}" for s in range(ns)]
    if os.path.exists(path + ".names"):
        with open(path + ".names") as f:
            for line in f:
                if line.startswith("#") or not line.strip():
                    continue
                b, nm = line.rstrip("\n").split("\t", 1)
                names[int(b) - 1] = nm
    return StateDump(ns, nc, t, _to_species_major(body[:n].astype(np.float64), ns, nc),
                     _to_species_major(body[n:].astype(np.float64), ns, nc), names)


# ------------------------------------------------------------------ Zarr v3
def _write_json(path: str, obj) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _array_meta(shape, chunk_shape, dims, compress: bool, attrs=None) -> dict:
    codecs = [{"name": "bytes", "configuration": {"endian": "little"}}]
    if compress:
        codecs.append({"name": "gzip", "configuration": {"level": 1}})
    return {
        "zarr_format": 3,
        "node_type": "array",
        "shape": list(shape),
        "data_type": "float64",
        "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": list(chunk_shape)}},
        "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}},
        "fill_value": "NaN",
        "codecs": codecs,
        "attributes": attrs or {},
        "dimension_names": list(dims),
    }


def _encode(a: np.ndarray, compress: bool) -> bytes:
    raw = np.ascontiguousarray(a, dtype="<f8").tobytes()
    return gzip.compress(raw, compresslevel=1, mtime=0) if compress else raw


class ZarrSink:
    """Streaming gridded output (open -> write* -> flush* -> close)."""

    def __init__(self, path: str, grid: Grid, names: list[str], output_times,
                 meta: dict | None = None, compress: bool = True):
        self.path = path
        self.grid = grid
        self.names = list(names)
        if len(set(self.names)) != len(self.names) or any("/" in n or n.startswith("__")
                                                          for n in self.names):
            raise ValueError("species names must be unique Zarr node names")
        self.times = np.asarray(output_times, np.float64)
        self.meta = dict(meta or {})
        self.compress = compress
        self.written = np.zeros(self.times.size, bool)
        self._q: queue.Queue | None = None
        self._thread = None
        self._err: BaseException | None = None

    # -- protocol -------------------------------------------------------------
    def open(self):
        g = self.grid
        os.makedirs(self.path, exist_ok=False)
        shape = (self.times.size, g.nlev, g.nlat, g.nlon)
        for nm in self.names:
            os.makedirs(os.path.join(self.path, nm))
            _write_json(os.path.join(self.path, nm, "zarr.json"),
                        _array_meta(shape, (1,) + shape[1:], ("time", "lev", "lat", "lon"),
                                    self.compress))
        coords = {"time": self.times, "lev": (np.arange(g.nlev) + 0.5) * g.dz,
                  "lat": np.degrees(g.lat), "lon": np.degrees(g.lon)}
        units = {"time": "s", "lev": "m", "lat": "degrees_north", "lon": "degrees_east"}
        for nm, v in coords.items():
            os.makedirs(os.path.join(self.path, nm))
            _write_json(os.path.join(self.path, nm, "zarr.json"),
                        _array_meta((v.size,), (v.size,), (nm,), False, {"units": units[nm]}))
            os.makedirs(os.path.join(self.path, nm, "c"))
            with open(os.path.join(self.path, nm, "c", "0"), "wb") as f:
                f.write(_encode(v, False))
        self._write_group()
        self._q = queue.Queue(maxsize=4)
        self._thread = threading.Thread(target=self._worker, daemon=True)
        self._thread.start()
        return self

    def write(self, t: float, u) -> None:
        """Queues the species-major state `u` as the record at time t."""
        if self._q is None:
            raise RuntimeError("sink is not open")
        self._raise_pending()
        rec = int(np.argmin(np.abs(self.times - t)))
        if abs(self.times[rec] - t) > 1e-6 * max(1.0, abs(t)):
            raise ValueError(f"t={t} is not an output time")
        g = self.grid
        u = np.array(u, np.float64, copy=True).reshape(len(self.names), g.nlev, g.nlat, g.nlon)
        self._q.put((rec, u))

    def flush(self) -> None:
        if self._q is not None:
            self._q.join()
        self._raise_pending()
        self._write_group()

    def close(self) -> None:
        if self._q is None:
            return
        self.flush()
        self._q.put(None)
        self._thread.join()
        self._q = None

    # -- internals ------------------------------------------------------------
    def _worker(self):
        while True:
            item = self._q.get()
            try:
                if item is None:
                    return
                rec, u = item
                for s, nm in enumerate(self.names):
                    d = os.path.join(self.path, nm, "c", str(rec), "0", "0")
                    os.makedirs(d, exist_ok=True)
                    with open(os.path.join(d, "0"), "wb") as f:
                        f.write(_encode(u[s][None], self.compress))
                self.written[rec] = True
            except BaseException as e:   # reported on the caller's next call
                self._err = e
            finally:
                self._q.task_done()

    def _raise_pending(self):
        if self._err is not None:
            e, self._err = self._err, None
            raise RuntimeError("output write failed") from e

    def _write_group(self):
        attrs = dict(self.meta)
        attrs.update({"species": self.names, "layout": "species-major, lon fastest",
                      "records_written": int(self.written.sum()),
                      "written": self.written.astype(int).tolist()})
        _write_json(os.path.join(self.path, "zarr.json"),
                    {"zarr_format": 3, "node_type": "group", "attributes": attrs})


def read_zarr_array(store: str, name: str) -> np.ndarray:
    """Reads a whole array of a Zarr v3 store (regular chunk grid, default
    chunk keys, bytes(+gzip) codecs), fill value where a chunk is absent."""
    with open(os.path.join(store, name, "zarr.json")) as f:
        m = json.load(f)
    if m.get("zarr_format") != 3 or m.get("node_type") != "array" or m["data_type"] != "float64":
        raise ValueError(f"{name}: not a float64 Zarr v3 array")
    shape = m["shape"]
    cs = m["chunk_grid"]["configuration"]["chunk_shape"]
    sep = m["chunk_key_encoding"].get("configuration", {}).get("separator", "/")
    names = [c["name"] for c in m["codecs"]]
    if names[0] != "bytes" or any(n not in ("bytes", "gzip") for n in names):
        raise ValueError(f"{name}: unsupported codecs {names}")
    fill = float(m["fill_value"]) if not isinstance(m["fill_value"], str) else float("nan")
    out = np.full(shape, fill)
    grid = [-(-n // c) for n, c in zip(shape, cs)]
    for idx in np.ndindex(*grid):
        key = os.path.join(store, name, *("c" + sep + sep.join(map(str, idx))).split("/"))
        if not os.path.exists(key):
            continue
        with open(key, "rb") as f:
            raw = f.read()
        if "gzip" in names:
            raw = gzip.decompress(raw)
        block = np.frombuffer(raw, "<f8").reshape(cs)
        sl = tuple(slice(i * c, min((i + 1) * c, n)) for i, c, n in zip(idx, cs, shape))
        out[sl] = block[tuple(slice(0, s.stop - s.start) for s in sl)]
    return out
