"""ctypes binding of the C ABI declared in include/reseact_hip.h.

`Library` wraps any shared library that exports that ABI; `hip_library()`
loads the product library libreseact_hip.so built in this package directory
and raises if it is missing -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .problems import Mechanism, Problem

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(PKG_DIR, "libreseact_hip.so")

RS_OK, RS_E_ARG, RS_E_HIP, RS_E_NONFINITE, RS_E_MAXITERS, RS_E_STATE, RS_E_JIT = range(7)
_ERRNAMES = {1: "RS_E_ARG", 2: "RS_E_HIP", 3: "RS_E_NONFINITE", 4: "RS_E_MAXITERS",
             5: "RS_E_STATE", 6: "RS_E_JIT"}


class PICtrl(C.Structure):
    """PI controller settings (rx_traced_integrator.jl:86-95)."""
    _fields_ = [("beta1", C.c_double), ("beta2", C.c_double), ("qmin", C.c_double),
                ("qmax", C.c_double), ("gamma", C.c_double), ("qsteady_min", C.c_double),
                ("qsteady_max", C.c_double), ("qoldinit", C.c_double)]


def PICtrl_make(beta1, beta2, qsteady_max, qmin=1 / 5, qmax=10.0, gamma=9 / 10,
                qsteady_min=1.0, qoldinit=1e-4) -> PICtrl:
    return PICtrl(beta1, beta2, qmin, qmax, gamma, qsteady_min, qsteady_max, qoldinit)


def pictrl_ros23() -> PICtrl:
    """OrdinaryDiffEqCore defaults for Rosenbrock23 (rx_traced_integrator.jl:98)."""
    return PICtrl_make(7 / 20, 1 / 5, 6 / 5)


def pictrl_ssprk43() -> PICtrl:
    """OrdinaryDiffEqCore defaults for SSPRK43 (rx_traced_integrator.jl:99)."""
    return PICtrl_make(7 / 30, 2 / 15, 1.0)


class Counts(C.Structure):
    _fields_ = [("naccept", C.c_longlong), ("nreject", C.c_longlong),
                ("t_end", C.c_double), ("dt_end", C.c_double)]


class SplitOpts(C.Structure):
    _fields_ = [("ctrl_T", PICtrl), ("ctrl_C", PICtrl), ("atol_T", C.c_double),
                ("rtol_T", C.c_double), ("atol_C", C.c_double), ("rtol_C", C.c_double),
                ("clamp_nonneg", C.c_int), ("chem_percell", C.c_int),
                ("maxiters", C.c_longlong)]


class PercellStats(C.Structure):
    _fields_ = [("attempts_total", C.c_longlong), ("accepts_total", C.c_longlong),
                ("rejects_total", C.c_longlong), ("attempts_max", C.c_longlong),
                ("attempts_min", C.c_longlong), ("n_failed", C.c_longlong)]


class AdjointStats(C.Structure):
    _fields_ = [("naccept_T", C.c_longlong), ("naccept_C", C.c_longlong),
                ("nreject_T", C.c_longlong), ("nreject_C", C.c_longlong),
                ("nvjp_T", C.c_longlong), ("nvjp_C", C.c_longlong),
                ("seconds_forward", C.c_double), ("seconds_backward", C.c_double),
                ("seconds_replay", C.c_double), ("replay_maxrel", C.c_double),
                ("bad_op", C.c_int), ("bad_macro", C.c_int),
                ("bad_step", C.c_longlong), ("bad_nsteps", C.c_longlong),
                ("bad_t", C.c_double), ("bad_dt", C.c_double),
                ("bad_lam_in", C.c_longlong), ("bad_lam_out", C.c_longlong),
                ("bad_grad", C.c_longlong)]


class MechInfo(C.Structure):
    _fields_ = [("ns", C.c_int), ("nr", C.c_int), ("nnz_J", C.c_int), ("nnz_LU", C.c_int),
                ("ordering", C.c_int), ("flops_attempt", C.c_longlong),
                ("flops_vjp", C.c_longlong), ("regs_attempt", C.c_int),
                ("regs_percell", C.c_int), ("regs_vjp", C.c_int),
                ("scratch_attempt", C.c_longlong), ("scratch_percell", C.c_longlong),
                ("scratch_vjp", C.c_longlong)]


class TuneResult(C.Structure):
    _fields_ = [("transport_tile_levels", C.c_int), ("chem_block", C.c_int),
                ("transport_ms", C.c_double * 4), ("chem_ms", C.c_double * 3),
                ("transport_fused", C.c_int), ("transport_fused_threads", C.c_int),
                ("transport_fused_ms", C.c_double),
                ("chem_kpointer", C.c_int), ("pad2", C.c_int), ("chem_kp_ms", C.c_double)]


class EnsembleStats(C.Structure):
    _fields_ = [("seconds_forward", C.c_double), ("seconds_backward", C.c_double),
                ("graph_launches_forward", C.c_longlong), ("graph_launches_backward", C.c_longlong),
                ("graph_nodes_backward", C.c_longlong), ("backward_groups", C.c_int),
                ("pad", C.c_int), ("seconds_backward_setup", C.c_double),
                ("seconds_backward_run", C.c_double)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_longlong), ("total_ms", C.c_double)]


# clamp flags of the adaptive solves (include/reseact_hip.h): RS_CLAMP_NONNEG
# zeroes negative entries of every accepted state (the op_split.jl:108-109
# callback); RS_POSITIVE_DOMAIN rejects a trial state with a negative entry of
# the operator's own block and retries at half the step (PositiveDomain
# isoutofdomain, HELPERS.md:396-447)
RS_CLAMP_NONNEG = 1
RS_POSITIVE_DOMAIN = 2

# transport schemes (include/reseact_hip.h)
RS_TRANSPORT_UPWIND3 = 0
RS_TRANSPORT_PPM = 1


def _flags(clamp) -> int:
    """True/False -> RS_CLAMP_NONNEG or 0; an int passes as the flag set."""
    v = int(clamp)
    if v & ~(RS_CLAMP_NONNEG | RS_POSITIVE_DOMAIN):
        raise ValueError(f"unknown clamp flags {v:#x}")
    return v


def split_opts(p: Problem, clamp: bool = True, percell: bool = False,
               maxiters: int = 20000, positive_domain: bool = False) -> SplitOpts:
    flags = _flags(clamp) | (RS_POSITIVE_DOMAIN if positive_domain else 0)
    return SplitOpts(pictrl_ssprk43(), pictrl_ros23(), p.atol_T, p.rtol_T, p.atol_C,
                     p.rtol_C, flags, 1 if percell else 0, maxiters)


class ReseactError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_P = C.c_void_p

# name -> argtypes (every function returns int unless listed in _RET)
ABI = {
    "rs_create": [C.POINTER(_P), C.c_int],
    "rs_destroy": [_P],
    "rs_last_error": [_P],
    "rs_backend": [],
    "rs_set_grid": [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "rs_set_wind": [_P, _D, _D, _D],
    "rs_set_transport_scheme": [_P, C.c_int],
    "rs_set_ppm_geometry": [_P, C.c_int, C.c_double, _D, _D, _D],
    "rs_set_mass_fluxes": [_P, _D, _D, _D, _D],
    "rs_set_inflow": [_P, _D, _D, _D, _D],
    "rs_set_pressure_fixer": [_P, _D],
    "rs_diagnose_mz": [_P, _D, _D, _D, _D],
    "rs_set_pbl_mixing": [_P, _D, C.c_double],
    "rs_set_mechanism": [_P, C.c_int, _I, _I, _D, _I],
    "rs_set_mechanism_stoich": [_P, C.c_int, _I, _I, _I, _D, _D, _I],
    "rs_set_photolysis": [_P, _D],
    "rs_set_emissions": [_P, _D],
    "rs_set_state": [_P, _D],
    "rs_get_state": [_P, _D],
    "rs_clamp_state": [_P],
    "rs_ssprk43_step": [_P, C.c_double, C.c_double, C.c_double, C.c_double, _D],
    "rs_ros23_step": [_P, C.c_double, C.c_double, C.c_double, C.c_double, _D],
    "rs_get_candidate": [_P, _D],
    "rs_ros23_step_noalloc": [_P, C.c_void_p, C.c_void_p, C.c_longlong, C.c_double, C.c_double,
                              C.c_double, _D],
    "rs_accept_candidate": [_P, C.c_int],
    "rs_adaptive_transport": [_P, C.c_double, C.c_double, _D, C.POINTER(PICtrl), C.c_double,
                              C.c_double, C.c_int, C.c_longlong, C.POINTER(Counts)],
    "rs_adaptive_chem": [_P, C.c_double, C.c_double, _D, C.POINTER(PICtrl), C.c_double,
                         C.c_double, C.c_int, C.c_longlong, C.POINTER(Counts)],
    "rs_get_step_log": [_P, C.c_int, _D, C.c_longlong, C.POINTER(C.c_longlong)],
    "rs_chem_percell": [_P, C.c_double, C.c_double, C.c_double, _D, C.POINTER(PICtrl),
                        C.c_double, C.c_double, C.c_int, C.c_longlong, C.POINTER(PercellStats)],
    "rs_chem_percell_vjp": [_P, C.c_double, C.c_double, C.c_double, C.POINTER(PICtrl),
                            C.c_double, C.c_double, C.c_int, C.c_longlong,
                            C.POINTER(PercellStats)],
    "rs_macro_step": [_P, C.c_double, C.c_double, _D, _D, C.POINTER(SplitOpts),
                      C.POINTER(Counts), C.POINTER(Counts)],
    "rs_set_lambda": [_P, _D],
    "rs_get_lambda": [_P, _D],
    "rs_zero_grad": [_P],
    "rs_get_grad_emis": [_P, _D],
    "rs_ssprk43_vjp": [_P, C.c_double, C.c_double],
    "rs_keep_transport_stages": [_P, C.c_int],
    "rs_ros23_vjp": [_P, C.c_double, C.c_double],
    "rs_adjoint_run": [_P, C.c_int, _D, C.c_double, C.c_double, C.POINTER(SplitOpts), C.c_int,
                       C.c_int, _D, _D, _D, C.POINTER(AdjointStats)],
    "rs_replay_loss": [_P, C.c_int, C.c_int, _D],
    "rs_set_objective_weights": [_P, _D],
    "rs_get_recorded_steps": [_P, C.c_int, C.c_int, _D, C.c_longlong, C.POINTER(C.c_longlong)],
    "rs_set_lu_ordering": [_P, C.c_int],
    "rs_set_transport_kernel": [_P, C.c_int],
    "rs_mechanism_info": [_P, C.c_void_p],
    "rs_profile": [_P, C.c_int],
    "rs_profile_read": [_P, C.c_void_p, C.c_int, C.POINTER(C.c_int)],
    "rs_autotune": [_P, C.c_void_p],
    "rs_ensemble_create": [C.POINTER(_P), C.POINTER(_P), C.c_int],
    "rs_ensemble_destroy": [_P],
    "rs_ensemble_last_error": [_P],
    "rs_ensemble_autotune": [_P, C.c_void_p],
    "rs_ensemble_adjoint_run": [_P, C.c_int, _D, C.c_double, C.c_double, C.c_void_p, C.c_int,
                                C.c_int, C.c_double, _D, C.c_void_p, C.c_void_p],
}
_RET = {"rs_last_error": C.c_char_p, "rs_backend": C.c_char_p,
        "rs_ensemble_last_error": C.c_char_p}


def _dp(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _ip(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


class Library:
    """A loaded implementation of the ABI."""

    def __init__(self, path: str, extra: dict | None = None):
        if not os.path.exists(path):
            raise FileNotFoundError(f"shared library not built: {path}")
        self.path = path
        self.lib = C.CDLL(path)
        for name, args in {**ABI, **(extra or {})}.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = _RET.get(name, C.c_int)

    @property
    def backend(self) -> str:
        return self.lib.rs_backend().decode()

    def context(self, device: int = 0) -> "Context":
        return Context(self, device)


def hip_library() -> Library:
    """The product library.  Raises if it has not been built."""
    if not os.path.exists(HIP_LIB_PATH):
        raise RuntimeError(
            f"{HIP_LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(there is no CPU fallback)")
    return Library(HIP_LIB_PATH)


class Context:
    """One rs_ctx.  Methods mirror the C ABI one to one."""

    def __init__(self, library: Library, device: int = 0):
        self.L = library
        self.lib = library.lib
        self.h = _P()
        self._chk(self.lib.rs_create(C.byref(self.h), device))
        self.grid = None
        self.ns = self.nt = 0

    def close(self):
        if self.h:
            self.lib.rs_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc: int):
        if rc != RS_OK:
            msg = self.lib.rs_last_error(self.h).decode() if self.h else ""
            raise ReseactError(rc, msg)

    # ---- setup --------------------------------------------------------------
    def set_grid(self, grid, ns: int, nt: int):
        self._chk(self.lib.rs_set_grid(self.h, grid.nlon, grid.nlat, grid.nlev, ns, nt))
        self.grid, self.ns, self.nt = grid, ns, nt

    @property
    def nc(self) -> int:
        return self.grid.nc

    def set_wind(self, wind):
        self._wind = (np.ascontiguousarray(wind.cx, np.float64),
                      np.ascontiguousarray(wind.wz, np.float64),
                      np.ascontiguousarray(wind.wl, np.float64))
        self._chk(self.lib.rs_set_wind(self.h, *(_dp(a) for a in self._wind)))

    def set_mechanism(self, m: Mechanism):
        re = np.ascontiguousarray(m.reactants, np.int32).reshape(-1)
        pr = np.ascontiguousarray(m.products, np.int32).reshape(-1)
        k = np.ascontiguousarray(m.k, np.float64)
        ph = np.ascontiguousarray(m.photo, np.int32)
        self._chk(self.lib.rs_set_mechanism(self.h, m.nr, _ip(re), _ip(pr), _dp(k), _ip(ph)))

    def set_mechanism_stoich(self, m):
        """General products (an esm.CompiledMechanism, or anything with the
        same arrays): rs_set_mechanism_stoich."""
        re = np.ascontiguousarray(m.reactants, np.int32).reshape(-1)
        off = np.ascontiguousarray(m.prod_off, np.int32)
        sp = np.ascontiguousarray(m.prod_species, np.int32)
        co = np.ascontiguousarray(m.prod_coef, np.float64)
        k = np.ascontiguousarray(m.k, np.float64)
        ph = np.ascontiguousarray(m.photo, np.int32)
        self._chk(self.lib.rs_set_mechanism_stoich(self.h, m.nr, _ip(re), _ip(off), _ip(sp), _dp(co),
                                                   _dp(k), _ip(ph)))

    def set_lu_ordering(self, ordering: int):
        self._chk(self.lib.rs_set_lu_ordering(self.h, ordering))

    def set_transport_kernel(self, mode: int):
        """1: fused SSPRK43 launch per attempt (default where the grid allows; 1024-lane
        workgroups on whole periodic rows, else 512); 2: fused, 1024 lanes; 3: fused,
        512 lanes; 0: per stage."""
        self._chk(self.lib.rs_set_transport_kernel(self.h, mode))

    def mechanism_info(self) -> dict:
        m = MechInfo()
        self._chk(self.lib.rs_mechanism_info(self.h, C.byref(m)))
        return {f: getattr(m, f) for f, _ in MechInfo._fields_}

    def profile(self, enable: bool):
        self._chk(self.lib.rs_profile(self.h, 1 if enable else 0))

    def profile_read(self) -> dict:
        arr = (KernelTime * 16)()
        n = C.c_int()
        self._chk(self.lib.rs_profile_read(self.h, arr, 16, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}

    def autotune(self) -> TuneResult:
        r = TuneResult()
        self._chk(self.lib.rs_autotune(self.h, C.byref(r)))
        return r

    def set_photolysis(self, jfac):
        a = None if jfac is None else np.ascontiguousarray(jfac, np.float64)
        self._chk(self.lib.rs_set_photolysis(self.h, _dp(a)))

    def set_emissions(self, e):
        a = None if e is None else np.ascontiguousarray(e, np.float64)
        self._chk(self.lib.rs_set_emissions(self.h, _dp(a)))

    def set_transport_scheme(self, scheme: int):
        self._chk(self.lib.rs_set_transport_scheme(self.h, scheme))

    def set_ppm(self, ppm):
        """Monotone PPM transport: scheme, geometry, air-mass fluxes, inflow."""
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        self.set_transport_scheme(RS_TRANSPORT_PPM)
        self._ppm_keep = [f(x) for x in (ppm.coslat_e, ppm.dS_lat, ppm.dphi_lat, ppm.Mx, ppm.My,
                                         ppm.Mz, ppm.dp, ppm.qbc_w, ppm.qbc_e, ppm.qbc_s,
                                         ppm.qbc_n)]
        ce, dS, dphi, Mx, My, Mz, dp, qw, qe, qs, qn = self._ppm_keep
        self._chk(self.lib.rs_set_ppm_geometry(self.h, 1 if ppm.lon_periodic else 0,
                                               float(ppm.dlon), _dp(ce), _dp(dS), _dp(dphi)))
        self._chk(self.lib.rs_set_mass_fluxes(self.h, _dp(Mx), _dp(My), _dp(Mz), _dp(dp)))
        self._chk(self.lib.rs_set_inflow(self.h, _dp(qw), _dp(qe), _dp(qs), _dp(qn)))
        fix = None if ppm.dPSdt is None else np.ascontiguousarray(ppm.dPSdt, np.float64)
        self._ppm_keep.append(fix)
        self._chk(self.lib.rs_set_pressure_fixer(self.h, _dp(fix)))
        self.set_pbl_mixing(ppm.pbl_f, ppm.tau_pbl)

    def set_pressure_fixer(self, dPSdt):
        """Column pressure fixer of the air mass (dPSdt per column; None: off)."""
        a = None if dPSdt is None else np.ascontiguousarray(dPSdt, np.float64).reshape(-1)
        self._fix_keep = a
        self._chk(self.lib.rs_set_pressure_fixer(self.h, _dp(a)))

    def diagnose_mz(self, dB=None, dPSdt=None):
        """Diagnose the vertical air-mass flux from continuity with the column
        pressure fixer (rs_diagnose_mz); it becomes the context's Mz.  Returns
        (Mz[(nlev+1), nlat, nlon], largest |scan value at the lid|)."""
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        dB, dPSdt = f(dB), f(dPSdt)
        g = self.grid
        out = np.empty((g.nlev + 1) * g.nlat * g.nlon)
        resid = C.c_double(0.0)
        self._chk(self.lib.rs_diagnose_mz(self.h, _dp(dB), _dp(dPSdt), _dp(out),
                                          C.byref(resid)))
        return out.reshape(g.nlev + 1, g.nlat, g.nlon), resid.value

    def set_pbl_mixing(self, pbl_f, tau=900.0):
        """PBL mixing of the mixing ratios (pbl_f per cell, None: off)."""
        a = None if pbl_f is None else np.ascontiguousarray(pbl_f, np.float64).reshape(-1)
        self._pbl_keep = a
        self._chk(self.lib.rs_set_pbl_mixing(self.h, _dp(a), float(tau)))

    def set_state(self, u):
        u = np.ascontiguousarray(u, np.float64)
        assert u.size == self.ns * self.nc
        self._chk(self.lib.rs_set_state(self.h, _dp(u)))

    def get_state(self) -> np.ndarray:
        out = np.empty(self.ns * self.nc)
        self._chk(self.lib.rs_get_state(self.h, _dp(out)))
        return out

    def clamp_state(self):
        self._chk(self.lib.rs_clamp_state(self.h))

    def load(self, p: Problem, state: bool = True):
        self.set_grid(p.grid, p.ns, p.nt)
        if p.ppm is not None:
            self.set_ppm(p.ppm)
        else:
            self.set_wind(p.wind)
        if p.mech is not None:
            self.set_mechanism(p.mech)
        self.set_photolysis(p.jfac)
        self.set_emissions(p.emis)
        if state:
            self.set_state(p.u0)
        return self

    # ---- single attempts ------------------------------------------------------
    def ssprk43_step(self, t, dt, atol, rtol) -> float:
        e = C.c_double()
        self._chk(self.lib.rs_ssprk43_step(self.h, t, dt, atol, rtol, C.byref(e)))
        return e.value

    def ros23_step(self, t, dt, atol, rtol) -> float:
        e = C.c_double()
        self._chk(self.lib.rs_ros23_step(self.h, t, dt, atol, rtol, C.byref(e)))
        return e.value

    def get_candidate(self) -> np.ndarray:
        out = np.empty(self.ns * self.nc)
        self._chk(self.lib.rs_get_candidate(self.h, _dp(out)))
        return out

    def accept_candidate(self, clamp: bool):
        self._chk(self.lib.rs_accept_candidate(self.h, _flags(clamp)))

    # ---- adaptive -------------------------------------------------------------
    def adaptive_transport(self, t0, t1, dt, ctrl=None, atol=1e-6, rtol=1e-3, clamp=True,
                           maxiters=20000):
        d = C.c_double(dt)
        c = Counts()
        ctrl = pictrl_ssprk43() if ctrl is None else ctrl
        self._chk(self.lib.rs_adaptive_transport(self.h, t0, t1, C.byref(d), C.byref(ctrl), atol,
                                                 rtol, _flags(clamp), maxiters, C.byref(c)))
        return d.value, c

    def adaptive_chem(self, t0, t1, dt, ctrl=None, atol=1e-6, rtol=1e-3, clamp=True,
                      maxiters=20000):
        d = C.c_double(dt)
        c = Counts()
        ctrl = pictrl_ros23() if ctrl is None else ctrl
        self._chk(self.lib.rs_adaptive_chem(self.h, t0, t1, C.byref(d), C.byref(ctrl), atol,
                                            rtol, _flags(clamp), maxiters, C.byref(c)))
        return d.value, c

    def step_log(self, op: int) -> np.ndarray:
        """Accepted (t, dt) pairs of the last adaptive transport (op 0) or
        chemistry (op 1) solve, shape (n, 2)."""
        n = C.c_longlong(0)
        self._chk(self.lib.rs_get_step_log(self.h, op, None, 0, C.byref(n)))
        out = np.empty((n.value, 2))
        self._chk(self.lib.rs_get_step_log(self.h, op, _dp(out), n.value, C.byref(n)))
        return out

    def chem_percell(self, t0, t1, dt0, dt_cells=None, ctrl=None, atol=1e-6, rtol=1e-3,
                     clamp=True, maxiters=20000):
        st = PercellStats()
        ctrl = pictrl_ros23() if ctrl is None else ctrl
        dc = None if dt_cells is None else np.ascontiguousarray(dt_cells, np.float64)
        self._chk(self.lib.rs_chem_percell(self.h, t0, t1, dt0, _dp(dc), C.byref(ctrl), atol,
                                           rtol, _flags(clamp), maxiters, C.byref(st)))
        return st

    def chem_percell_vjp(self, t0, t1, dt0, ctrl=None, atol=1e-6, rtol=1e-3, clamp=True,
                         maxiters=20000):
        """lambda <- VJP of the per-cell chemistry over [t0, t1] from the
        current state (unchanged); the emission gradient accumulates."""
        st = PercellStats()
        ctrl = pictrl_ros23() if ctrl is None else ctrl
        self._chk(self.lib.rs_chem_percell_vjp(self.h, t0, t1, dt0, C.byref(ctrl), atol, rtol,
                                               _flags(clamp), maxiters, C.byref(st)))
        return st

    def macro_step(self, t0, t1, dtT, dtC, opts: SplitOpts):
        a, b = C.c_double(dtT), C.c_double(dtC)
        cT, cC = Counts(), Counts()
        self._chk(self.lib.rs_macro_step(self.h, t0, t1, C.byref(a), C.byref(b), C.byref(opts),
                                         C.byref(cT), C.byref(cC)))
        return a.value, b.value, cT, cC

    # ---- adjoint --------------------------------------------------------------
    def set_lambda(self, lam):
        lam = np.ascontiguousarray(lam, np.float64)
        self._chk(self.lib.rs_set_lambda(self.h, _dp(lam)))

    def get_lambda(self) -> np.ndarray:
        out = np.empty(self.ns * self.nc)
        self._chk(self.lib.rs_get_lambda(self.h, _dp(out)))
        return out

    def zero_grad(self):
        self._chk(self.lib.rs_zero_grad(self.h))

    def get_grad_emis(self) -> np.ndarray:
        out = np.empty(self.ns * self.grid.ncol)
        self._chk(self.lib.rs_get_grad_emis(self.h, _dp(out)))
        return out

    def ssprk43_vjp(self, t, dt):
        self._chk(self.lib.rs_ssprk43_vjp(self.h, t, dt))

    def keep_transport_stages(self, keep: bool):
        """PPM: keep the stages of rs_ssprk43_step for a following step VJP."""
        self._chk(self.lib.rs_keep_transport_stages(self.h, 1 if keep else 0))

    def ros23_vjp(self, t, dt):
        self._chk(self.lib.rs_ros23_vjp(self.h, t, dt))

    def set_objective_weights(self, w):
        """Weights of the linear objective J = <w, u(t_end)> (loss kind 1);
        None clears them."""
        w = None if w is None else np.ascontiguousarray(w, np.float64)
        self._chk(self.lib.rs_set_objective_weights(self.h, _dp(w)))

    def adjoint_run(self, tstops, dtT0, dtC0, opts: SplitOpts, loss_species=0,
                    want_grads=True, loss_kind=0):
        """Forward + discrete adjoint.  A failure raises ReseactError; for
        RS_E_NONFINITE / RS_E_STATE the error carries .stats (with the first
        non-finite VJP in its bad_* fields) and .result = (J, gu, ge)."""
        ts = np.ascontiguousarray(tstops, np.float64)
        J = C.c_double()
        gu = np.empty(self.ns * self.nc) if want_grads else None
        ge = np.empty(self.ns * self.grid.ncol) if want_grads else None
        st = AdjointStats()
        rc = self.lib.rs_adjoint_run(self.h, len(ts) - 1, _dp(ts), dtT0, dtC0, C.byref(opts),
                                     loss_kind, loss_species, C.byref(J), _dp(gu), _dp(ge),
                                     C.byref(st))
        if rc != RS_OK:
            err = ReseactError(rc, self.lib.rs_last_error(self.h).decode())
            err.stats = st
            err.result = (J.value, gu, ge)
            raise err
        return J.value, gu, ge, st

    def recorded_steps(self, macro_step: int, op: int) -> np.ndarray:
        """Accepted (t, dt) of operator op (0 transport, 1 chemistry) in one
        macro step of the last adjoint run, shape (n, 2)."""
        n = C.c_longlong(0)
        self._chk(self.lib.rs_get_recorded_steps(self.h, macro_step, op, None, 0, C.byref(n)))
        out = np.empty((n.value, 2))
        self._chk(self.lib.rs_get_recorded_steps(self.h, macro_step, op, _dp(out), n.value,
                                                 C.byref(n)))
        return out

    def replay_loss(self, loss_species=0, loss_kind=0) -> float:
        J = C.c_double()
        self._chk(self.lib.rs_replay_loss(self.h, loss_kind, loss_species, C.byref(J)))
        return J.value


class Ensemble:
    """rs_ensemble over already loaded member contexts."""

    def __init__(self, members: list[Context]):
        self.members = members
        self.lib = members[0].lib
        arr = (_P * len(members))(*[m.h for m in members])
        self.h = _P()
        rc = self.lib.rs_ensemble_create(C.byref(self.h), arr, len(members))
        if rc != RS_OK:
            raise ReseactError(rc, "rs_ensemble_create failed")

    def _chk(self, rc: int):
        if rc != RS_OK:
            raise ReseactError(rc, self.lib.rs_ensemble_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.rs_ensemble_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self) -> list:
        arr = (TuneResult * len(self.members))()
        self._chk(self.lib.rs_ensemble_autotune(self.h, arr))
        return list(arr)

    def adjoint_run(self, tstops, dtT0, dtC0, opts: SplitOpts, loss_species=0,
                    tape_budget_gib=64.0, loss_kind=0):
        ts = np.ascontiguousarray(tstops, np.float64)
        n = len(self.members)
        J = np.empty(n)
        st = (AdjointStats * n)()
        es = EnsembleStats()
        rc = self.lib.rs_ensemble_adjoint_run(self.h, len(ts) - 1, _dp(ts), dtT0, dtC0,
                                              C.byref(opts), loss_kind, loss_species,
                                              tape_budget_gib, _dp(J), st, C.byref(es))
        if rc != RS_OK:
            err = ReseactError(rc, self.lib.rs_ensemble_last_error(self.h).decode())
            err.stats = list(st)
            raise err
        return J, list(st), es


@dataclass
class ModelRun:
    """Convenience: a loaded context plus its problem."""
    ctx: Context
    problem: Problem
