"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

Each .npz holds the inputs of one call and the oracle's outputs:
  transport_step_6x4x3.npz  one SSPRK43 attempt, 3 tracers on a 6x4x3 grid
  toy_ros23_3sp_5cells.npz  one ROS23 attempt and its VJP on the 3-species x
                            5-cell toy mechanism (A+B->C, C->A+B, 2C->B)
  coupled_macro_9x6x4.npz   one coupled split step of the 10-species workload
                            on a 9x6x4 sub-grid (accepted/rejected counts too)
Run: python tests/golden/make_golden.py   (rewrites the files)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from reseact_esm_amd import problems as P  # noqa: E402
from reseact_esm_amd.host import split_opts  # noqa: E402


def problem(name: str) -> P.Problem:
    if name == "transport_step_6x4x3":
        g = P.Grid(6, 4, 3)
        return P.Problem(name=name, grid=g, ns=3, nt=3, wind=P.solid_body_wind(g, 600.0, 0.8),
                         mech=None, u0=P.blob_state(g, 3, 77))
    if name == "toy_ros23_3sp_5cells":
        return P.toy_problem()
    if name == "coupled_macro_9x6x4":
        return P.config1(nlon=9, nlat=6, nlev=4, nmacro=1)
    raise KeyError(name)


NAMES = ["transport_step_6x4x3", "toy_ros23_3sp_5cells", "coupled_macro_9x6x4"]


def compute(lib, name: str) -> dict:
    p = problem(name)
    ctx = lib.context().load(p)
    out = {"u0": p.u0}
    if name == "transport_step_6x4x3":
        out["eest"] = np.array(ctx.ssprk43_step(0.0, 600.0, 1e-6, 1e-3))
        out["unew"] = ctx.get_candidate()
        lam = np.random.default_rng(1).standard_normal(p.u0.size)
        ctx.set_lambda(lam)
        ctx.ssprk43_vjp(0.0, 600.0)
        out["lam"] = lam
        out["lam_in"] = ctx.get_lambda()
    elif name == "toy_ros23_3sp_5cells":
        out["eest"] = np.array(ctx.ros23_step(100.0, 0.05, 1e-9, 1e-4))
        out["unew"] = ctx.get_candidate()
        lam = np.random.default_rng(11).standard_normal(p.u0.size)
        ctx.set_lambda(lam)
        ctx.zero_grad()
        ctx.ros23_vjp(100.0, 0.05)
        out["lam"] = lam
        out["lam_in"] = ctx.get_lambda()
    else:
        dtT, dtC, cT, cC = ctx.macro_step(0.0, p.dt_macro, p.dt0_T, p.dt0_C, split_opts(p))
        out["u1"] = ctx.get_state()
        out["counts"] = np.array([cT.naccept, cT.nreject, cC.naccept, cC.nreject])
        out["dt_next"] = np.array([dtT, dtC])
    return out


def main() -> None:
    from oracle_lib import oracle_library
    lib = oracle_library()
    for name in NAMES:
        np.savez(os.path.join(HERE, name + ".npz"), **compute(lib, name))
        print("wrote", name)


if __name__ == "__main__":
    main()
