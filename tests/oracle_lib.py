"""Loader of the CPU oracle (oracle/libreseact_cpu.so) for tests and the CPU
baseline.  The oracle is the checker, never the thing under test."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

from reseact_esm_amd.host import Library, _P, _D

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "libreseact_cpu.so")

EXTRA = {
    "rs_oracle_ros23_jvp": [_P, C.c_double, _D, _D, _D],
    "rs_oracle_ssprk43_jvp": [_P, C.c_double, _D, _D],
    "rs_oracle_transport_rhs": [_P, _D, _D],
    "rs_oracle_transport_rhs_adj": [_P, _D, _D],
    "rs_oracle_transport_rhs_jvp": [_P, _D, _D],
    "rs_oracle_chem_rhs": [_P, _D, _D],
    "rs_oracle_chem_jac": [_P, _D, C.c_size_t, _D],
    "rs_oracle_attempt_log": [_P, _D, C.c_longlong, C.POINTER(C.c_longlong)],
    "rs_oracle_adjoint_on_sequence": [_P, C.c_int, _D, C.POINTER(C.c_longlong),
                                      C.POINTER(C.c_longlong), _D, _D, C.c_void_p, C.c_int,
                                      C.c_int, _D, _D, _D, C.c_void_p],
    "rs_oracle_threads": [],
}


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_LIB = None


def oracle_library() -> Library:
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        _LIB = Library(ORACLE_LIB, EXTRA)
    return _LIB


def adjoint_on_sequence(ctx, p, opts, seqT, seqC, loss_kind=0, loss_species=0):
    """The oracle's time-loop adjoint on context `ctx` (problem p loaded) with
    the forward taking the accepted sequences seqT[m], seqC[m] (arrays of
    (t, dt) rows per macro step).  Returns (J, grad_u0, grad_emis, stats)."""
    import numpy as np
    from reseact_esm_amd.host import AdjointStats
    ts = np.ascontiguousarray(p.tstops(), np.float64)
    nT = (C.c_longlong * len(seqT))(*[len(s) for s in seqT])
    nC = (C.c_longlong * len(seqC))(*[len(s) for s in seqC])
    tT = np.ascontiguousarray(np.concatenate([np.reshape(s, (-1, 2)) for s in seqT]), np.float64)
    tC = np.ascontiguousarray(np.concatenate([np.reshape(s, (-1, 2)) for s in seqC]), np.float64)
    J = C.c_double()
    gu = np.empty(p.ns * p.nc)
    ge = np.empty(p.ns * p.grid.ncol)
    st = AdjointStats()
    P_ = C.POINTER(C.c_double)
    ctx._chk(ctx.lib.rs_oracle_adjoint_on_sequence(
        ctx.h, len(seqT), ts.ctypes.data_as(P_), nT, nC, tT.ctypes.data_as(P_),
        tC.ctypes.data_as(P_), C.byref(opts), loss_kind, loss_species, C.byref(J),
        gu.ctypes.data_as(P_), ge.ctypes.data_as(P_), C.byref(st)))
    return J.value, gu, ge, st
