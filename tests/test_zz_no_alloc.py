"""rs_ros23_step_noalloc (the entry point for callers that cannot allocate
device memory) against rs_ros23_step, through the C ABI, on two calls in a
row, the second with a smaller input.  Device buffers come from the HIP
runtime the library itself links (test plumbing)."""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import pytest

from reseact_esm_amd import problems as P
from reseact_esm_amd.host import RS_E_ARG, ReseactError

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class DeviceBuffer:
    """A device allocation of float64, through the HIP runtime already loaded
    by libreseact_hip.so (same soname, so the same instance)."""

    rt = None

    def __init__(self, n: int):
        if DeviceBuffer.rt is None:
            DeviceBuffer.rt = C.CDLL("libamdhip64.so.7")
        self.n = n
        self.ptr = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        assert self.rt.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n),
                                 _H2D) == 0

    def download(self) -> np.ndarray:
        out = np.empty(self.n)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n),
                                 _D2H) == 0
        return out

    def free(self):
        if self.ptr:
            assert self.rt.hipFree(self.ptr) == 0
            self.ptr = C.c_void_p()


def noalloc(ctx, ubuf, obuf, ncells, dt):
    e = C.c_double(-1.0)
    ctx._chk(ctx.lib.rs_ros23_step_noalloc(ctx.h, ubuf.ptr, obuf.ptr, ncells, dt, 1e-6, 1e-3,
                                           C.byref(e)))
    return e.value


def levels_prefix(p, k):
    """The problem restricted to its lowest k levels: its first nlon*nlat*k
    cells, species-major with that stride (emissions stay on level 0)."""
    g = P.Grid(p.grid.nlon, p.grid.nlat, k)
    n = g.nc
    u0 = p.u0.reshape(p.ns, p.nc)[:, :n].reshape(-1).copy()
    jf = None if p.jfac is None else p.jfac[:n].copy()
    return dataclasses.replace(p, grid=g, u0=u0, jfac=jf,
                               wind=P.solid_body_wind(g, p.dt_macro, 0.8))


def test_noalloc_attempt_matches_ros23_step(hip):
    p = P.config4(nlon=40, nlat=24, nlev=6)          # 5,760 cells: 90 workgroups
    p.jfac = np.random.default_rng(21).uniform(0.0, 1.0, p.nc)
    dt = 0.5
    with hip.context() as c:
        c.load(p)
        bufs = [DeviceBuffer(p.ns * p.nc) for _ in range(2)]
        try:
            # call 1: every cell, against the existing entry point on the same state
            e_ref = c.ros23_step(0.0, dt, 1e-6, 1e-3)
            u_ref = c.get_candidate()
            bufs[0].upload(p.u0)
            print(f"noalloc call 1: {p.nc} cells", flush=True)
            e1 = noalloc(c, bufs[0], bufs[1], p.nc, dt)
            assert np.array_equal(bufs[1].download(), u_ref)
            assert e1 == e_ref
            # the context's own state is untouched
            assert np.array_equal(c.get_state(), p.u0)
            # call 2, smaller: the lowest two levels, against rs_ros23_step on
            # a context whose grid is those levels
            q = levels_prefix(p, 2)
            with hip.context() as c2:
                c2.load(q)
                e2_ref = c2.ros23_step(0.0, dt, 1e-6, 1e-3)
                u2_ref = c2.get_candidate()
            small = [DeviceBuffer(q.ns * q.nc) for _ in range(2)]
            try:
                small[0].upload(q.u0)
                print(f"noalloc call 2: {q.nc} cells", flush=True)
                e2 = noalloc(c, small[0], small[1], q.nc, dt)
                assert np.array_equal(small[1].download(), u2_ref)
                assert e2 == e2_ref
            finally:
                for b in small:
                    b.free()
            # sizes beyond the grid are refused before any launch
            with pytest.raises(ReseactError) as err:
                noalloc(c, bufs[0], bufs[1], p.nc + 1, dt)
            assert err.value.code == RS_E_ARG
        finally:
            for b in bufs:
                b.free()
