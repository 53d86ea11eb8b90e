"""Benchmark of the operator-split step on one MI355X.

Workload (the default, and the `value` line): the coupled forward of the
144x91x47 grid, 20 species / 40 reactions, split steps of 900 s (SSPRK43
transport of all 20 species, then Rosenbrock23 chemistry with one global
error norm), from seeded synthetic inputs already resident in HBM.  A "step"
is one split step over the whole grid; value = cells * steps / seconds.

Also reported on the same run:
  * adjoint/forward ratio: checkpointed forward + backward sweep over the
    same steps, divided by the plain forward time;
  * transport HBM roofline: the 360x180x72, 20-tracer transport at Courant
    0.8 (fixed dt); the fused kernel (one launch per attempt) moves 16 B per
    tracer-cell algorithmically (u in, unew out), the per-stage chain 72 B;
  * chemistry FP64 roofline: the 360x180x72, 50-species / 120-reaction
    per-cell adaptive chemistry over 1800 s in one persistent launch;
  * the plain coupled forward of 72x46x10 (4 tracers, 10 species, 24 split
    steps of 600 s), whole run;
  * cpu_baseline: the CPU oracle (C++/OpenMP, same algorithm) on 7 of the 91
    latitude rows of the same 144x91x47 grid, one warm-up split step then one
    timed, on this host's cores; cpu_baselines: the same for the other
    workloads (72x46x10 whole run; transport, batched chemistry and ensemble
    on labelled subsets).
Kernel durations come from HIP events recorded by the library on its own
stream around every launch (rs_profile).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick] [--no-cpu]
                       [--no-ensemble] [--probe]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from reseact_esm_amd import hip_library, problems as P, split_opts  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# PPM transport: flops per mixing-ratio species-cell and stage, counted from the
# operator's statements with each face reconstructed once (ppm_roofline)
PPM_FLOPS_SPECIES_CELL_STAGE = 172.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_ms(prof: dict, name: str) -> tuple[int, float]:
    n, ms = prof.get(name, (0, 0.0))
    return n, ms


def coupled(lib, steps: int, warmup: int):
    p = P.config4(nmacro=max(steps, 1))
    opts = split_opts(p)
    ctx = lib.context().load(p)
    # setup, untimed: pick the transport kernel and chemistry launch shape on
    # this grid (rs_autotune; the state is not changed)
    tune = ctx.autotune()
    log(f"autotune: transport fused={tune.transport_fused} ({tune.transport_fused_ms:.4f} ms) tile "
        f"{tune.transport_tile_levels}; chemistry block {tune.chem_block} "
        f"{[round(x, 4) for x in tune.chem_ms]} ms, rate constants via pointer={tune.chem_kpointer}")
    info = ctx.mechanism_info()
    log(f"coupled {p.name}: nc={p.nc} mech nnz_LU={info['nnz_LU']} flops/attempt="
        f"{info['flops_attempt']} regs={info['regs_attempt']} scratch={info['scratch_attempt']}")
    dtT, dtC = p.dt0_T, p.dt0_C
    for m in range(warmup):
        dtT, dtC, _, _ = ctx.macro_step(m * p.dt_macro, (m + 1) * p.dt_macro, dtT, dtC, opts)
    def forward():
        ctx.set_state(p.u0)
        dtT, dtC = p.dt0_T, p.dt0_C
        nT = nC = rT = rC = 0
        t0 = time.perf_counter()
        for m in range(steps):
            dtT, dtC, cT, cC = ctx.macro_step(m * p.dt_macro, (m + 1) * p.dt_macro, dtT, dtC, opts)
            nT += cT.naccept; rT += cT.nreject; nC += cC.naccept; rC += cC.nreject
        return time.perf_counter() - t0, (nT, rT, nC, rC)

    # the timed run records no per-launch events; a second, identical run
    # with the library's HIP-event profiling gives the per-kernel times
    fwd_s, (nT, rT, nC, rC) = forward()
    ctx.profile(True)
    fwd_prof_s, counts2 = forward()
    prof = ctx.profile_read()
    ctx.profile(False)
    if counts2 != (nT, rT, nC, rC):
        raise RuntimeError(f"profiled forward took another step sequence: {counts2}")
    log(f"coupled forward {steps} steps {fwd_s:.3f} s (with per-launch events {fwd_prof_s:.3f} s); "
        f"T {nT}/{rT} C {nC}/{rC}; kernels {prof}")
    # adjoint over the same steps; one untimed run first sizes the checkpoint
    # store and the replay tape (device allocations kept by the context)
    ctx.set_state(p.u0)
    ctx.adjoint_run(p.tstops()[:steps + 1], p.dt0_T, p.dt0_C, opts)
    ctx.set_state(p.u0)
    J, gu, ge, st = ctx.adjoint_run(p.tstops()[:steps + 1], p.dt0_T, p.dt0_C, opts)
    ratio = (st.seconds_forward + st.seconds_backward) / fwd_s
    log(f"adjoint: fwd(ckpt) {st.seconds_forward:.3f} s bwd {st.seconds_backward:.3f} s "
        f"(replay {st.seconds_replay:.3f}) ratio {ratio:.3f} J={J:.6e} |gu|={np.linalg.norm(gu):.4e} "
        f"|gE|={np.linalg.norm(ge):.4e} replay_maxrel={st.replay_maxrel:.2e}")
    # launches after the controller finished exit at once; the work is one
    # attempt over every cell per accepted or rejected step
    n_launch, ms_att = kernel_ms(prof, "chem_attempt")
    n_att = nC + rC
    flops = info["flops_attempt"] * p.nc
    avg_ms = ms_att / max(n_att, 1)
    achieved = flops / (avg_ms * 1e-3) / 1e12 if n_att else 0.0
    # compute roof: FP64 vector (VALU) issue, 78.6 TFLOP/s on MI355X; no
    # matrix instructions are used
    roof = {"bound": "fp64-valu", "kernel": "chem_attempt",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "avg_ms": round(avg_ms, 4), "attempts": n_att, "launches": n_launch,
            "flops_per_launch": flops}
    roof["launch"] = {"chem_block": tune.chem_block, "rate_constants": "pointer" if tune.chem_kpointer
                      else "kernel argument", "transport": "fused" if tune.transport_fused
                      else f"per-stage, {tune.transport_tile_levels} levels"}
    return dict(p=p, fwd_s=fwd_s, steps=steps, ratio=ratio, adj=st, roof=roof, prof=prof,
                counts=(nT, rT, nC, rC), info=info, J=J)


def transport_roofline(lib, nsteps: int = 60, nt: int = 20):
    """C2 transport at Courant 0.8 with a fixed dt: the fused kernel (one launch
    per attempt) at its default workgroup size (512 lanes on this grid's
    longitude segments), the fused kernel forced to 1024-lane workgroups, and
    the per-stage kernels (four launches), on the same inputs."""
    p = P.config2(nt=nt, nsteps=nsteps)
    dt = p.extra["dt_fixed"]
    res = {}
    for mode, kernels in ((1, ("transport_fused",)), (2, ("transport_fused",)),
                          (0, ("transport_stage", "transport_last"))):
        ctx = lib.context()
        ctx.set_transport_kernel(mode)
        ctx.load(p)
        ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)      # warm-up
        ctx.profile(True)
        t0 = time.perf_counter()
        for _ in range(nsteps):
            ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
            ctx.accept_candidate(False)
        t1 = time.perf_counter()
        prof = ctx.profile_read()
        ctx.profile(False)
        ctx.close()
        launches = sum(kernel_ms(prof, k)[0] for k in kernels)
        ms = sum(kernel_ms(prof, k)[1] for k in kernels)
        res[mode] = dict(ms_step=ms / nsteps, wall=t1 - t0, launches=launches,
                         avg_launch_ms=ms / max(launches, 1))
    cells = nt * p.nc
    f, f1024, s = res[1], res[2], res[0]
    # algorithmic bytes: the fused kernel must read u and write unew (16 B per
    # tracer-cell); the per-stage chain moves 72 B (stages 1-3: 8 + 8, last
    # stage: u, u3 in, unew out)
    achieved = 16.0 * cells / (f["avg_launch_ms"] * 1e-3) / 1e9
    log(f"transport {p.name}: fused {f['ms_step']:.3f} ms/step ({achieved:.0f} GB/s at 16 B; "
        f"1024-lane workgroups {f1024['ms_step']:.3f} ms), per-stage {s['ms_step']:.3f} ms/step "
        f"({72.0 * cells / (s['ms_step'] * 1e-3) / 1e9:.0f} GB/s at 72 B)")
    return {"workload": p.name, "kernel": "transport_fused", "bound": "hbm",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": 16.0 * cells,
            "avg_launch_ms": round(f["avg_launch_ms"], 4), "ms_per_step": round(f["ms_step"], 4),
            "tracer_cell_steps_per_s": round(cells / (f["ms_step"] * 1e-3), 1),
            "fused_1024_lanes": {"ms_per_step": round(f1024["ms_step"], 4),
                                 "achieved_GBs": round(16.0 * cells / (f1024["ms_step"] * 1e-3) / 1e9, 1)},
            "per_stage": {"ms_per_step": round(s["ms_step"], 4), "bytes_per_step": 72.0 * cells,
                          "achieved_GBs": round(72.0 * cells / (s["ms_step"] * 1e-3) / 1e9, 1),
                          "launches_per_step": round(s["launches"] / nsteps, 2)},
            "fused_speedup": round(s["ms_step"] / f["ms_step"], 3)}


def ppm_roofline(lib, nsteps: int = 10, nq: int = 20):
    """The reference's transport: monotone PPM in CWC mixing-ratio form on the
    global 360x180x72 grid, air mass + 20 mixing ratios, one SSPRK43 attempt
    per step (four stage launches) from the same state."""
    p = P.ppm_global(360, 180, 72, nq=nq)
    nt = p.nt
    ctx = lib.context().load(p)
    dt = 60.0
    ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)          # warm-up
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    # the step VJP at the same state: three forward stages recompute the stage
    # states (and the tendencies), a tendency-only stage covers u4, four
    # adjoint applications (ppm_vjp: z and y sweeps, longitude tiles, air-mass
    # row) reverse them
    # (set_state before each, so that every VJP recomputes its stages)
    ctx.set_lambda(np.ones(p.ns * p.nc))
    ctx.ssprk43_vjp(0.0, dt)                       # warm-up
    ctx.profile(True)                              # counters restart here
    nvjp = max(2, nsteps // 2)
    for _ in range(nvjp):
        ctx.set_state(p.u0)
        ctx.ssprk43_vjp(0.0, dt)
    profv = ctx.profile_read()
    ctx.profile(False)
    # the pullback form: the step keeps its stage states and tendencies
    # (rs_keep_transport_stages) and the VJP of the same step reuses them
    ctx.keep_transport_stages(True)
    ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)          # warm-up
    ctx.ssprk43_vjp(0.0, dt)
    ctx.profile(True)
    for _ in range(nvjp):
        ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
        ctx.ssprk43_vjp(0.0, dt)
    profk = ctx.profile_read()
    ctx.profile(False)
    ctx.keep_transport_stages(False)
    # the same attempts with the PBL mixing term on (a 400-1600 m boundary
    # layer; its column means are a pre-pass inside each stage's timing)
    P.with_pbl_mixing(p)
    ctx.set_pbl_mixing(p.ppm.pbl_f, p.ppm.tau_pbl)
    ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
    ctx.profile(True)
    for _ in range(nsteps):
        ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
    profp = ctx.profile_read()
    ctx.profile(False)
    ctx.close()
    _, ms_ps = kernel_ms(profp, "transport_stage")
    _, ms_pl = kernel_ms(profp, "transport_last")
    pbl_ms = (ms_ps + ms_pl) / nsteps
    nv_, ms_v = kernel_ms(profv, "ppm_vjp")
    nvs_, ms_vs = kernel_ms(profv, "transport_stage")
    vjp_ms = (ms_v + ms_vs) / nvjp
    ns_, ms_s = kernel_ms(prof, "transport_stage")
    nl_, ms_l = kernel_ms(prof, "transport_last")
    ms = (ms_s + ms_l) / nsteps
    _, ms_ks = kernel_ms(profk, "transport_stage")
    _, ms_kl = kernel_ms(profk, "transport_last")
    nkv_, ms_kv = kernel_ms(profk, "ppm_vjp")
    keep_step_ms = (ms_ks + ms_kl) / nvjp
    keep_vjp_ms = ms_kv / nvjp
    g = p.grid
    # algorithmic bytes per attempt: per species-cell 16 B in stages 1-2
    # (read, write), 32 B in stage 3 (reads a and u, writes u3 and a4) and 32 B
    # in the last (reads a4, u, u3, writes unew); per cell and stage the face
    # fluxes Mx, My, Mz and the vertical width factors (2 per slope row, 3 per
    # edge row), each read once
    per_cell_stage = 8.0 * ((g.nlon + 1) / g.nlon + (g.nlat + 1) / g.nlat + (g.nlev + 1) / g.nlev
                            + 2.0 * (g.nlev + 4) / g.nlev + 3.0 * (g.nlev + 3) / g.nlev)
    bytes_attempt = p.nc * (96.0 * nt + 4.0 * per_cell_stage)
    gbs = bytes_attempt / (ms * 1e-3) / 1e9
    # algorithmic flops per mixing-ratio species-cell and stage, counted from
    # the operator's statements (min/max and compares not counted), each face
    # reconstructed once (a cell owns one face per direction): two uniform
    # faces x 52 (3 slopes, 2 edges, limiter), one width-weighted vertical
    # face 51, three face fluxes and the divergence 12, the CWC tendency 3,
    # the stage update 2; the last stage adds 9 for the error norm; the air
    # mass takes 11 per stage (divergence and update)
    flops_attempt = p.nc * ((nt - 1) * (4.0 * PPM_FLOPS_SPECIES_CELL_STAGE + 9.0) + 4.0 * 11.0 + 9.0)
    tf = flops_attempt / (ms * 1e-3) / 1e12
    log(f"ppm {p.name}: {ms:.3f} ms/attempt (stage {ms_s / max(ns_, 1):.3f} ms, last "
        f"{ms_l / max(nl_, 1):.3f} ms), {tf:.2f} TFLOP/s, {gbs:.0f} GB/s algorithmic, "
        f"wall {t1 - t0:.2f} s; step VJP {vjp_ms:.3f} ms ({ms_v / max(nv_, 1):.3f} ms per "
        f"ppm_vjp launch), VJP/attempt {vjp_ms / ms:.2f}; with kept stages: step "
        f"{keep_step_ms:.3f} ms, VJP {keep_vjp_ms:.3f} ms ({keep_vjp_ms / ms:.2f} x the attempt; "
        f"step + VJP {(keep_step_ms + keep_vjp_ms) / ms:.2f} x, recomputing {(ms + vjp_ms) / ms:.2f} x)")
    return {"workload": p.name, "kernels": ["transport_stage", "transport_last"],
            "bound": "fp64-valu", "achieved": round(tf, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFLOPS, 4),
            "flops_per_attempt": flops_attempt,
            "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_attempt": bytes_attempt},
            "ms_per_attempt": round(ms, 4),
            "avg_stage_ms": round(ms_s / max(ns_, 1), 4), "avg_last_ms": round(ms_l / max(nl_, 1), 4),
            "tracer_cell_attempts_per_s": round(nt * p.nc / (ms * 1e-3), 1),
            "pbl_mixing": {"ms_per_attempt": round(pbl_ms, 4), "over_plain": round(pbl_ms / ms, 3)},
            "vjp": {"ms_per_step_vjp": round(vjp_ms, 4), "avg_ppm_vjp_ms": round(ms_v / max(nv_, 1), 4),
                    "ppm_vjp_launches_per_vjp": round(nv_ / nvjp, 2),
                    "vjp_over_attempt": round(vjp_ms / ms, 3),
                    "kept_stages": {"step_ms": round(keep_step_ms, 4), "vjp_ms": round(keep_vjp_ms, 4),
                                    "vjp_over_attempt": round(keep_vjp_ms / ms, 3),
                                    "step_plus_vjp_over_attempt": round((keep_step_ms + keep_vjp_ms) / ms, 3),
                                    "recomputing_step_plus_vjp_over_attempt": round((ms + vjp_ms) / ms, 3),
                                    "ppm_vjp_launches_per_vjp": round(nkv_ / nvjp, 2)}}}


def chemistry_roofline(lib):
    p = P.config3()
    ctx = lib.context().load(p)
    info = ctx.mechanism_info()
    ctx.profile(True)
    t0 = time.perf_counter()
    st = ctx.chem_percell(0.0, 1800.0, p.dt0_C, None, None, p.atol_C, p.rtol_C, True)
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    n, ms = kernel_ms(prof, "chem_percell")
    flops = st.attempts_total * info["flops_attempt"]
    achieved = flops / (ms * 1e-3) / 1e12
    log(f"chemistry {p.name}: attempts {st.attempts_total} (min {st.attempts_min} max "
        f"{st.attempts_max}) kernel {ms:.1f} ms wall {t1 - t0:.2f} s {achieved:.2f} TFLOP/s "
        f"regs {info['regs_percell']} scratch {info['scratch_percell']}")
    dropin = chemistry_dropin(lib)
    return {"workload": p.name, "bound": "fp64-valu", "kernel": "chem_percell",
            "dropin_global_norm": dropin,
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "kernel_ms": round(ms, 2),
            "attempts": st.attempts_total, "attempts_min": st.attempts_min,
            "attempts_max": st.attempts_max, "flops_per_attempt": info["flops_attempt"],
            "cell_steps_per_s": round(p.nc / (ms * 1e-3), 1)}


def chemistry_dropin(lib, nlon: int = 24):
    """The batched per-cell solve beside the plain drop-in on the same cells: a
    24-longitude subset of the 50-species grid (day and night columns), the
    persistent work-stealing kernel with per-cell error control against the
    global-norm adaptive solve (one controller over all cells, the reference's
    adaptive_solve), both over the 1800 s interval."""
    p = P.config3(nlon=nlon)
    res = {"sample": f"{p.grid.nlon}x{p.grid.nlat}x{p.grid.nlev} subset of the 50-species grid"}
    with lib.context() as ctx:
        ctx.load(p)
        t0 = time.perf_counter()
        st = ctx.chem_percell(0.0, 1800.0, p.dt0_C, None, None, p.atol_C, p.rtol_C, True)
        res["percell_seconds"] = round(time.perf_counter() - t0, 3)
        res["percell_attempts_mean"] = round(st.attempts_total / p.nc, 1)
    with lib.context() as ctx:
        ctx.load(p)
        t0 = time.perf_counter()
        _, c = ctx.adaptive_chem(0.0, 1800.0, p.dt0_C, None, p.atol_C, p.rtol_C, True)
        res["global_seconds"] = round(time.perf_counter() - t0, 3)
        res["global_attempts"] = c.naccept + c.nreject
    res["percell_cell_steps_per_s"] = round(p.nc / res["percell_seconds"], 1)
    res["global_cell_steps_per_s"] = round(p.nc / res["global_seconds"], 1)
    res["percell_speedup"] = round(res["global_seconds"] / res["percell_seconds"], 3)
    log(f"chemistry drop-in: {res}")
    return res


def ensemble_run(lib, ndom: int = 32, nmacro: int = 6, tape_budget_gib: float = 96.0):
    """32 regional domains (64x64x20 .. 256x192x47), 6 split steps forward plus
    adjoint, autotuned tiles, graph-captured (rs_ensemble_adjoint_run)."""
    from reseact_esm_amd.host import Ensemble
    probs = P.config5(ndom=ndom, nmacro=nmacro)
    members = [lib.context().load(p) for p in probs]
    ens = Ensemble(members)
    t0 = time.perf_counter()
    tunes = ens.autotune()
    t_tune = time.perf_counter() - t0
    p0 = probs[0]
    J, st, es = ens.adjoint_run(p0.tstops(), p0.dt0_T, p0.dt0_C, split_opts(p0), 0,
                                tape_budget_gib)
    cells = sum(p.nc for p in probs)
    hist_t: dict = {}
    hist_c: dict = {}
    for r in tunes:
        hist_t[r.transport_tile_levels] = hist_t.get(r.transport_tile_levels, 0) + 1
        hist_c[r.chem_block] = hist_c.get(r.chem_block, 0) + 1
    out = {"workload": f"ensemble-{ndom}-domains-20sp", "cells": cells, "split_steps": nmacro,
           "forward_cell_steps_per_s": round(cells * nmacro / es.seconds_forward, 1),
           "seconds_forward": round(es.seconds_forward, 3),
           "seconds_backward": round(es.seconds_backward, 3),
           "seconds_backward_setup": round(es.seconds_backward_setup, 3),
           "seconds_backward_run": round(es.seconds_backward_run, 3),
           "adjoint_ratio": round((es.seconds_forward + es.seconds_backward) / es.seconds_forward, 3),
           "graph_launches_forward": es.graph_launches_forward,
           "graph_launches_backward": es.graph_launches_backward,
           "graph_nodes_backward": es.graph_nodes_backward,
           "backward_groups": es.backward_groups,
           "autotune_seconds": round(t_tune, 2),
           "tile_levels_chosen": {str(k): v for k, v in sorted(hist_t.items())},
           "chem_block_chosen": {str(k): v for k, v in sorted(hist_c.items())},
           "max_replay_maxrel": max(s.replay_maxrel for s in st)}
    log(f"ensemble: {out}")
    for m in members:
        m.close()
    ens.close()
    # the plain drop-in beside it: the same domains one after another through
    # rs_adjoint_run, each in a fresh context (default kernels, no graphs)
    fwd = bwd = 0.0
    for p in probs:
        with lib.context() as ctx:
            ctx.load(p)
            _, _, _, s1 = ctx.adjoint_run(p.tstops(), p.dt0_T, p.dt0_C, split_opts(p), 0,
                                          want_grads=False)
            fwd += s1.seconds_forward
            bwd += s1.seconds_backward
    out["dropin_sequential"] = {
        "forward_cell_steps_per_s": round(cells * nmacro / fwd, 1),
        "seconds_forward": round(fwd, 3), "seconds_backward": round(bwd, 3),
        "adjoint_ratio": round((fwd + bwd) / fwd, 3),
        "graph_speedup_forward": round(fwd / es.seconds_forward, 3),
        "graph_speedup_forward_plus_adjoint": round((fwd + bwd) / (es.seconds_forward + es.seconds_backward), 3)}
    log(f"ensemble drop-in (sequential): {out['dropin_sequential']}")
    return out


def config1_leg(lib, warmup: int = 1):
    """The plain coupled forward of 72x46x10 (4 tracers, 10 species / 20
    reactions, 24 split steps of 600 s), whole run, after `warmup` untimed
    split steps; chemistry and transport kernel times from rs_profile."""
    p = P.config1()
    opts = split_opts(p)
    ctx = lib.context().load(p)
    ctx.autotune()
    info = ctx.mechanism_info()
    dtT, dtC = p.dt0_T, p.dt0_C
    for m in range(warmup):
        dtT, dtC, _, _ = ctx.macro_step(m * p.dt_macro, (m + 1) * p.dt_macro, dtT, dtC, opts)
    ts = p.tstops()

    def run():
        ctx.set_state(p.u0)
        dT, dC = p.dt0_T, p.dt0_C
        nC = 0
        t0 = time.perf_counter()
        for m in range(p.nmacro):
            dT, dC, cT, cC = ctx.macro_step(ts[m], ts[m + 1], dT, dC, opts)
            nC += cC.naccept + cC.nreject
        return time.perf_counter() - t0, nC
    sec, nC = run()
    ctx.profile(True)
    run()
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.close()
    _, ms_att = kernel_ms(prof, "chem_attempt")
    avg = ms_att / max(nC, 1)
    tf = info["flops_attempt"] * p.nc / (avg * 1e-3) / 1e12 if nC else 0.0
    out = {"workload": p.name, "split_steps": p.nmacro,
           "cell_steps_per_s": round(p.nc * p.nmacro / sec, 1),
           "seconds": round(sec, 4), "ms_per_step": round(sec / p.nmacro * 1e3, 3),
           "chem_attempt": {"bound": "fp64-valu", "avg_ms": round(avg, 5),
                            "achieved": round(tf, 3), "unit": "TFLOP/s",
                            "frac": round(tf / FP64_PEAK_TFLOPS, 4), "attempts": nC,
                            "flops_per_attempt_cell": info["flops_attempt"]}}
    log(f"config1 {p.name}: {out}")
    return out


def _oracle():
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_lib import oracle_library
    return oracle_library()


def host_info(lib) -> dict:
    return {"cores": int(lib.lib.rs_oracle_threads()), "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": cpu_model()}


def _macro_run(ctx, p, steps):
    opts = split_opts(p)
    dtT, dtC = p.dt0_T, p.dt0_C
    for m in range(steps):
        dtT, dtC, _, _ = ctx.macro_step(m * p.dt_macro, (m + 1) * p.dt_macro, dtT, dtC, opts)


def cpu_baseline(rows_step: int = 13):
    """The oracle on latitude rows 0, 13, ..., 78 (7 of 91) of the same
    144x91x47 coupled workload: one warm-up split step, then one timed."""
    lib = _oracle()
    p = P.latitude_rows(P.config4(nmacro=1), range(0, 91, rows_step))
    ctx = lib.context().load(p)
    _macro_run(ctx, p, 1)                                # warm-up
    ctx.set_state(p.u0)
    t0 = time.perf_counter()
    _macro_run(ctx, p, 1)
    sec = time.perf_counter() - t0
    ctx.close()
    rate = p.nc / sec
    h = host_info(lib)
    log(f"cpu oracle {p.name}: {sec:.2f} s, {rate:.1f} cell-steps/s, {h}")
    return {"value": round(rate, 2), "unit": "coupled cell-steps/s", "cores": h["cores"],
            "kind": "port",
            "sample": f"latitude rows 0, {rows_step}, ..., {rows_step * (p.grid.nlat - 1)} "
                      f"({p.grid.nlat} of 91) of the same 144x91x47 workload, all longitudes and "
                      f"levels ({p.nc} cells): one warm-up split step of 900 s, then one timed, "
                      f"{sec:.1f} s",
            **{k: v for k, v in h.items() if k != "cores"}}


def cpu_baselines() -> dict:
    """The oracle beside every other GPU leg, on this host's cores, each after
    one warm-up step: the 72x46x10 coupled run whole; the 360x180x72
    transport, the 50-species batched chemistry and the ensemble on labelled
    subsets."""
    lib = _oracle()
    h = host_info(lib)
    out = {"host": h}
    # C1: whole run
    p = P.config1()
    ctx = lib.context().load(p)
    _macro_run(ctx, p, 1)
    ctx.set_state(p.u0)
    t0 = time.perf_counter()
    _macro_run(ctx, p, p.nmacro)
    sec = time.perf_counter() - t0
    ctx.close()
    out["config1"] = {"value": round(p.nc * p.nmacro / sec, 2), "unit": "coupled cell-steps/s",
                      "cores": h["cores"], "kind": "port", "seconds": round(sec, 3),
                      "sample": f"the whole run: {p.grid} x {p.nmacro} split steps, after one "
                                f"warm-up step"}
    # C2: every 10th latitude row of the 20-tracer transport, 2 fixed steps
    p = P.latitude_rows(P.config2(), range(0, 180, 10))
    dt = p.extra["dt_fixed"]
    ctx = lib.context().load(p)
    ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
    t0 = time.perf_counter()
    for _ in range(2):
        ctx.ssprk43_step(0.0, dt, 1e-6, 1e-3)
        ctx.accept_candidate(False)
    sec = time.perf_counter() - t0
    ctx.close()
    out["config2"] = {"value": round(p.ns * p.nc * 2 / sec, 1), "unit": "tracer-cell-steps/s",
                      "cores": h["cores"], "kind": "port", "seconds": round(sec, 3),
                      "sample": f"latitude rows 0, 10, ..., 170 (18 of 180) of 360x180x72, 20 "
                                f"tracers, 2 steps at Courant 0.8 after one warm-up step"}
    # C3: stratified cells of the batched chemistry, per cell over 1800 s
    p3 = P.config3()
    cells = P.stratified_cells(p3.grid)
    q = P.cell_subset(p3, cells)
    w = P.cell_subset(p3, cells[:512])
    ctx = lib.context().load(w)
    ctx.chem_percell(0.0, 1800.0, w.dt0_C, None, None, w.atol_C, w.rtol_C, True)   # warm-up
    ctx.close()
    ctx = lib.context().load(q)
    t0 = time.perf_counter()
    st = ctx.chem_percell(0.0, 1800.0, q.dt0_C, None, None, q.atol_C, q.rtol_C, True)
    sec = time.perf_counter() - t0
    ctx.close()
    out["config3"] = {"value": round(cells.size / sec, 2), "unit": "cell-steps/s (1800 s interval)",
                      "attempts_per_s": round(st.attempts_total / sec, 1),
                      "cores": h["cores"], "kind": "port", "seconds": round(sec, 3),
                      "sample": f"{cells.size} cells of 360x180x72: every 16th longitude, 6th "
                                f"latitude and 6th level, per-cell ROS23 over 1800 s (after a "
                                f"warm-up run on 512 of them)"}
    # C5: every 6th latitude row of the smallest of the 32 domains, one split
    # step forward
    probs = P.config5(nmacro=1)
    p0 = min(probs, key=lambda q_: q_.nc)
    p = P.latitude_rows(p0, range(0, p0.grid.nlat, 6))
    ctx = lib.context().load(p)
    _macro_run(ctx, p, 1)
    ctx.set_state(p.u0)
    t0 = time.perf_counter()
    _macro_run(ctx, p, 1)
    sec = time.perf_counter() - t0
    ctx.close()
    out["ensemble"] = {"value": round(p.nc / sec, 2), "unit": "coupled cell-steps/s (forward)",
                       "cores": h["cores"], "kind": "port", "seconds": round(sec, 3),
                       "sample": f"every 6th latitude row ({p.grid.nlat} rows) of the smallest of "
                                 f"the 32 domains ({p0.grid.nlon}x{p0.grid.nlat}x{p0.grid.nlev}), one "
                                 f"split step of 900 s forward after one warm-up step"}
    log(f"cpu baselines: {out}")
    return out


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def probe(lib) -> dict:
    """Short run for hardware-counter passes (FETCH_SIZE, WRITE_SIZE): a
    1 GiB device-to-device hipMemcpy (the runtime's copy kernel) that
    calibrates the counters, then two fused transport attempts on the
    360x180x72 20-tracer grid and two chemistry attempts on the 144x91x47
    coupled grid.  The copy goes through the HIP runtime the library itself
    links (one runtime per process)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    n = 1 << 27
    a, b = C.c_void_p(), C.c_void_p()
    for q in (a, b):
        if hip.hipMalloc(C.byref(q), C.c_size_t(8 * n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, C.c_size_t(8 * n))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, C.c_size_t(8 * n), 3) != 0:      # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    for q in (a, b):
        hip.hipFree(q)
    p = P.config2(nt=20, nsteps=1)
    ctx = lib.context().load(p)
    for _ in range(2):
        ctx.ssprk43_step(0.0, p.extra["dt_fixed"], 1e-6, 1e-3)
    ctx.close()
    p4 = P.config4(nmacro=1)
    ctx = lib.context().load(p4)
    for _ in range(2):
        ctx.ros23_step(0.0, 1.0, p4.atol_C, p4.rtol_C)
    ctx.close()
    return {"copy_bytes": 8 * n, "transport": {"workload": p.name, "tracer_cells": 20 * p.nc},
            "chemistry": {"workload": p4.name, "cells": p4.nc}}


def probe_chem(lib) -> dict:
    """Short run for the FP64 counter passes of the chemistry kernels: two
    ROS23 attempts and two step VJPs over the 144x91x47 coupled grid, then the
    per-cell work-stealing solve of the 50-species mechanism over 1800 s on an
    8-longitude band of the 360x180x72 grid (every longitude index i with
    i % 45 == 0: day and night columns)."""
    p4 = P.config4(nmacro=1)
    ctx = lib.context().load(p4)
    for _ in range(2):
        ctx.ros23_step(0.0, 1.0, p4.atol_C, p4.rtol_C)
    ctx.set_lambda(np.ones(p4.ns * p4.nc))
    ctx.zero_grad()
    for _ in range(2):
        ctx.ros23_vjp(0.0, 1.0)
    info4 = ctx.mechanism_info()
    ctx.close()
    p3 = P.config3()
    band = np.arange(0, p3.grid.nlon, 45)
    cells = P.stratified_cells(p3.grid, di=45, dj=1, dk=1)
    q = P.cell_subset(p3, cells)
    ctx = lib.context().load(q)
    st = ctx.chem_percell(0.0, 1800.0, q.dt0_C, None, None, q.atol_C, q.rtol_C, True)
    info3 = ctx.mechanism_info()
    ctx.close()
    return {"chem_attempt": {"workload": p4.name, "cells": p4.nc, "launches": 2,
                             "flops_attempt_cell": info4["flops_attempt"]},
            "chem_vjp": {"workload": p4.name, "cells": p4.nc, "launches": 2,
                         "flops_vjp_cell": info4["flops_vjp"]},
            "chem_percell": {"workload": p3.name, "cells": int(cells.size),
                             "longitudes": band.tolist(), "attempts": st.attempts_total,
                             "flops_attempt_cell": info3["flops_attempt"]}}


def _pmc_pass(counter: str, outdir: str, timeout_s: int = 150) -> dict:
    """One rocprofv3 counter pass over `bench.py --probe` in a child process;
    returns {kernel name: [per-dispatch KiB]}."""
    import csv
    import shutil
    exe = shutil.which("rocprofv3")
    if exe is None:
        raise RuntimeError("rocprofv3 not found")
    os.makedirs(outdir, exist_ok=True)
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", outdir, "-o", counter.lower(),
           "--", sys.executable, os.path.join(REPO, "bench.py"), "--probe"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s,
                       env=dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp")))
    if r.returncode != 0:
        raise RuntimeError(f"rocprofv3 {counter} exited {r.returncode}")
    fn = os.path.join(outdir, f"{counter.lower()}_counter_collection.csv")
    out: dict = {}
    with open(fn) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                out.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return out


def traffic(outdir: str) -> dict:
    """HBM bytes per launch of the two roofline kernels from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the counter reads for
    the probe's 1 GiB device-to-device copy (bytes moved / counter bytes)."""
    fe = _pmc_pass("FETCH_SIZE", outdir)
    wr = _pmc_pass("WRITE_SIZE", outdir)
    copy = 8.0 * (1 << 27)
    cname = next(k for k in fe if "copyBuffer" in k)
    f_fetch = copy / (max(fe[cname]) * 1024.0)
    f_write = copy / (max(wr[cname]) * 1024.0)
    res = {"calibration": {"kernel": cname, "bytes": copy, "fetch_factor": round(f_fetch, 4),
                           "write_factor": round(f_write, 4)}}
    for key, pat in (("transport_fused", "rs_transport_fused<0,"), ("chem_attempt", "rs_chem_attempt")):
        kf = next(k for k in fe if pat in k)
        kw = next(k for k in wr if pat in k)
        rd = sum(fe[kf]) / len(fe[kf]) * 1024.0 * f_fetch
        wt = sum(wr[kw]) / len(wr[kw]) * 1024.0 * f_write
        res[key] = {"read_bytes": round(rd), "write_bytes": round(wt), "bytes": round(rd + wt)}
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="skip the transport/chemistry rooflines")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 32-domain ensemble")
    ap.add_argument("--probe", action="store_true",
                    help="only the short hardware-counter run (copy + one transport and chemistry)")
    ap.add_argument("--probe-chem", action="store_true",
                    help="only the short chemistry run for the FP64 counter passes")
    ap.add_argument("--ppm-only", action="store_true",
                    help="only the PPM transport leg (for profiling it)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure HBM traffic")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the split step shards by "
              "latitude bands across GPUs, which this build does not launch.", file=sys.stderr)
        return 2
    lib = hip_library()
    if a.probe:
        print(json.dumps({"probe": probe(lib)}), flush=True)
        return 0
    if a.probe_chem:
        print(json.dumps({"probe_chem": probe_chem(lib)}), flush=True)
        return 0
    if a.ppm_only:
        print(json.dumps({"ppm_transport": ppm_roofline(lib, nsteps=a.steps)}), flush=True)
        return 0
    c = coupled(lib, a.steps, a.warmup)
    extra = {}
    if not a.quick:
        extra["config1"] = config1_leg(lib)
        extra["transport_hbm"] = transport_roofline(lib)
        extra["chemistry_fp64"] = chemistry_roofline(lib)
        extra["ppm_transport"] = ppm_roofline(lib)
    if not a.quick and not a.no_ensemble:
        extra["ensemble"] = ensemble_run(lib)
    cpu = None if a.no_cpu else cpu_baseline()
    if not a.no_cpu and not a.quick:
        extra["cpu_baselines"] = cpu_baselines()
    tr = None
    if not a.no_traffic:
        try:
            tr = traffic(os.path.join(REPO, "gpurun_out", "bench_pmc"))
            log(f"traffic: {tr}")
        except Exception as e:  # the counters are a report, not the measurement
            log(f"traffic counters unavailable: {type(e).__name__} {e}")
    if "cpu_baselines" in extra:
        # GPU over CPU on each leg, in the leg's own unit
        cb = extra["cpu_baselines"]
        for leg, key, cpu_key in (("config1", "cell_steps_per_s", "config1"),
                                  ("transport_hbm", "tracer_cell_steps_per_s", "config2"),
                                  ("chemistry_fp64", "cell_steps_per_s", "config3"),
                                  ("ensemble", "forward_cell_steps_per_s", "ensemble")):
            if leg in extra and cb.get(cpu_key, {}).get("value"):
                extra[leg]["vs_cpu"] = round(extra[leg][key] / cb[cpu_key]["value"], 1)
    p = c["p"]
    if tr:
        c["roof"]["traffic"] = tr["chem_attempt"]["bytes"]
        c["roof"]["traffic_read_write"] = [tr["chem_attempt"]["read_bytes"],
                                           tr["chem_attempt"]["write_bytes"]]
        if "transport_hbm" in extra:
            extra["transport_hbm"]["traffic"] = tr["transport_fused"]["bytes"]
        extra["traffic_calibration"] = tr["calibration"]
    value = p.nc * c["steps"] / c["fwd_s"]
    line = {
        "metric": "coupled cell-steps/s fwd; adjoint/forward time ratio; % HBM & FP64 roofline",
        "value": round(value, 1),
        "unit": "coupled cell-steps/s",
        "n_gpus": 1,
        "steps": c["steps"],
        "warmup": a.warmup,
        "ms_per_step": round(c["fwd_s"] / c["steps"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal state, random mass-action mechanism, solid-body wind)",
        "config": {"workload": p.name, "grid": [p.grid.nlon, p.grid.nlat, p.grid.nlev],
                   "species": p.ns, "reactions": p.mech.nr, "split_dt_s": p.dt_macro,
                   "chemistry_norm": "global", "parallelism": "single GPU"},
        "roofline": c["roof"],
        "cpu_baseline": cpu,
        "adjoint_ratio": round(c["ratio"], 3),
        "adjoint": {"seconds_forward_ckpt": round(c["adj"].seconds_forward, 4),
                    "seconds_backward": round(c["adj"].seconds_backward, 4),
                    "seconds_replay": round(c["adj"].seconds_replay, 4),
                    "nvjp_T": c["adj"].nvjp_T, "nvjp_C": c["adj"].nvjp_C},
        "inner_steps": {"transport_acc_rej": c["counts"][:2], "chem_acc_rej": c["counts"][2:]},
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        **extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
