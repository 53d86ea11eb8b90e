"""Builds the mechanism compiler's emitted code for the CPU (g++) so tests can
check it against the oracle without a GPU."""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(REPO, "reseact_esm_amd", "csrc")
WORK = os.path.join(tempfile.gettempdir(), f"reseact_hostchem_{os.getuid()}")
_D = C.POINTER(C.c_double)


def _cli() -> str:
    os.makedirs(WORK, exist_ok=True)
    exe = os.path.join(WORK, "mechgen_cli")
    srcs = [os.path.join(HERE, "mechgen_cli.cpp"), os.path.join(CSRC, "mechgen.cpp")]
    if not os.path.exists(exe) or any(os.path.getmtime(s) > os.path.getmtime(exe) for s in srcs
                                      + [os.path.join(CSRC, "mechgen.h")]):
        subprocess.run(["g++", "-O1", "-std=c++17", "-o", exe] + srcs, check=True)
    return exe


def generate(mech, ordering: int) -> tuple[str, str]:
    """Generated source for a problems.Mechanism (two product slots) or an
    esm.CompiledMechanism (general products)."""
    if hasattr(mech, "prod_off"):
        lines = [f"{mech.ns} {mech.nr} {-1 - ordering}"]
        for r in range(mech.nr):
            a, b = mech.reactants[r]
            prods = [(int(mech.prod_species[q]), float(mech.prod_coef[q]))
                     for q in range(mech.prod_off[r], mech.prod_off[r + 1])]
            lines.append(f"{a} {b} {int(mech.photo[r])} {len(prods)} " +
                         " ".join(f"{s_} {c!r}" for s_, c in prods))
    else:
        lines = [f"{mech.ns} {mech.nr} {ordering}"]
        for r in range(mech.nr):
            a, b = mech.reactants[r]
            p, q = mech.products[r]
            lines.append(f"{a} {b} {p} {q} {int(mech.photo[r])}")
    res = subprocess.run([_cli()], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True)
    return res.stdout, res.stderr


class HostChem:
    def __init__(self, mech, ordering: int = 1, exact_div: bool = False, sanitize: bool = False):
        src, info = generate(mech, ordering)
        self.info = dict(zip(info.split()[0::2], map(int, info.split()[1::2])))
        deps = "".join(open(os.path.join(d_, f)).read() for d_, f in
                       ((HERE, "host_chem.cpp"), (CSRC, "chem_kernels.hip"),
                        (CSRC, "chem_prelude.h"), (CSRC, "rs_launch.h")))
        key = hashlib.sha1((src + deps + str(exact_div) + str(sanitize)).encode()).hexdigest()[:16]
        d = os.path.join(WORK, key)
        os.makedirs(d, exist_ok=True)
        so = os.path.join(d, "hostchem.so")
        if not os.path.exists(so):
            with open(os.path.join(d, "gen_mech.h"), "w") as f:
                f.write(src)
            cmd = ["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                   "-I", d, "-I", CSRC, "-x", "c++", os.path.join(HERE, "host_chem.cpp"),
                   "-o", so + ".tmp"]
            if exact_div:
                cmd.insert(1, "-DRS_EXACT_DIV")
            if sanitize:   # host code only (g++): address and undefined-behaviour checks
                cmd[1:1] = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                            "-fno-omit-frame-pointer"]
            subprocess.run(cmd, check=True)
            os.replace(so + ".tmp", so)
        self.lib = C.CDLL(so)
        self.lib.hc_attempt.restype = C.c_double
        self.lib.hc_attempt.argtypes = [C.c_longlong, C.c_longlong, _D, _D, _D, _D, _D,
                                        C.c_double, C.c_double, C.c_double]
        self.lib.hc_attempt_wave.restype = C.c_longlong
        self.lib.hc_attempt_wave.argtypes = [C.c_longlong, C.c_longlong, _D, _D, _D, _D, _D,
                                             C.c_double, C.c_double, C.c_double, _D]
        self.lib.hc_vjp.restype = None
        self.lib.hc_vjp.argtypes = [C.c_longlong, C.c_longlong, _D, _D, _D, _D, C.c_double, _D, _D]
        self.mech = mech

    @staticmethod
    def _p(a):
        return None if a is None else np.ascontiguousarray(a, np.float64).ctypes.data_as(_D)

    def attempt(self, p, u, dt, atol, rtol):
        nc, ncol = p.grid.nc, p.grid.ncol
        out = np.empty_like(u)
        keep = [np.ascontiguousarray(x, np.float64) if x is not None else None
                for x in (self.mech.k, p.jfac, p.emis, u)]
        sse = self.lib.hc_attempt(nc, ncol, *(self._p(x) for x in keep), self._p(out), dt, atol, rtol)
        return out, float(np.sqrt(sse / (p.ns * nc)))

    def attempt_wave(self, p, u, dt, atol, rtol):
        """The attempt with the device's [3][NS][64] slab layout; returns
        (unew, EEst, number of slab slots of other lanes that were written)."""
        nc, ncol = p.grid.nc, p.grid.ncol
        out = np.empty_like(u)
        sse = C.c_double(0.0)
        keep = [np.ascontiguousarray(x, np.float64) if x is not None else None
                for x in (self.mech.k, p.jfac, p.emis, u)]
        bad = self.lib.hc_attempt_wave(nc, ncol, *(self._p(x) for x in keep), self._p(out), dt,
                                       atol, rtol, C.byref(sse))
        return out, float(np.sqrt(sse.value / (p.ns * nc))), int(bad)

    def vjp(self, p, u, dt, lam):
        nc, ncol = p.grid.nc, p.grid.ncol
        lam = np.array(lam, dtype=np.float64)
        gE = np.zeros(p.ns * ncol)
        keep = [np.ascontiguousarray(x, np.float64) if x is not None else None
                for x in (self.mech.k, p.jfac, p.emis, u)]
        self.lib.hc_vjp(nc, ncol, *(self._p(x) for x in keep), dt, self._p(lam), self._p(gE))
        return lam, gE
