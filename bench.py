#!/usr/bin/env python
"""Benchmark of the run-time stepping physics on one MI355X.

Headline workload (one "step" = one pass of the hot path over one batch):
1M fresh tracks, e-:e+:gamma = 4:1:5 by track id, kinetic energy log-uniform in
[100 keV, 10 GeV], isotropic directions, couples uniform over
{G4_PbWO4, G4_Pb, G4_lAr, G4_Si} at a 0.7 mm production cut, safety
log-uniform in [1 um, 10 mm], one Philox stream per (seed, track id); one
HowFar + AlongStep + PostStep Perform of every track in FP64 (hepem_step:
particle-type compaction + fused physics kernel). Inputs are resident in HBM
before the timed region; every step starts from the same input batch and writes
a separate output batch.

Also reported: the FP32-table form of the same step, config-1 style lookups
(100k and 10M e- queries over G4_Pb/G4_lAr), 64 uneven batches (sizes
log-uniform in [1e3, 1e6], ~9.3M tracks) stepped by one tuned batch plan
replayed from a hipGraph in both forms, 10 GeV e- showers in 50 x (2.3 mm Pb +
5.7 mm lAr) (10k events, with the CPU harness on a sample), and the oracle's
CPU throughput on the headline workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SET1 = ["G4_Pb", "G4_lAr"]
SET2 = ["G4_PbWO4", "G4_Pb", "G4_lAr", "G4_Si"]
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(ptype: np.ndarray, nsec_total: int, out_of_place: bool = True) -> int:
    """HBM bytes a step must move: the per-track state it reads and writes
    (rng_stream included), the queue entry, the per-track results and the
    secondaries actually created (33 B each: type, energy, direction); SoA
    field sizes of include/g4hepem_amd.h."""
    common_r = 1 + 4 + 4 + 8 + 24 + 8 + 1 + 32 + 4 + 8 + 1 + 4  # ptype id imc ekin dir safety onb nial rng(4)
    msc = 3 * 8 + 1
    queue = 4
    state_w = 8 + 24 + 32 + 4 + 8 + 1  # ekin dir nial rng
    ident_w = (1 + 4 + 4 + 8 + 1 + 4) if out_of_place else 0
    res_w = 8 + 8 + 8 + 24 + 4 + 1 + 4  # true/geom step, edep, displacement, winner, stopped, nsec
    e_bytes = common_r + msc + queue + state_w + msc + ident_w + res_w
    g_bytes = common_r + queue + state_w + ident_w + res_w
    ne = int(np.count_nonzero(ptype != 2))
    return ne * e_bytes + (len(ptype) - ne) * g_bytes + 33 * int(nsec_total)


def measured_traffic(kernel_profile: str):
    """HBM bytes per launch of the step kernel from the committed hardware
    counter summary (FETCH_SIZE + WRITE_SIZE passes, each scaled by the same
    counters over a plain copy of known size); None when absent."""
    path = os.path.join(ROOT, "profiles", kernel_profile)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1]), "profiles/" + kernel_profile
    except OSError:
        pass
    return None, None


def time_steps(ctx, tin, tout, res, work, steps, warmup, seed):
    """Times `steps` back-to-back steps (prepare + run) with HIP events on the
    launch stream; the run (the physics kernels: selection, process sort and
    final states) of each step is bracketed by its own event pair. Argument structs are built once so the host enqueues ahead
    of the GPU."""
    import ctypes as C

    import torch
    from g4hepem_amd import lib
    L = lib()
    stream = torch.cuda.current_stream()
    sh = C.c_void_p(int(stream.cuda_stream))
    si, so, rr = tin.struct(), tout.struct(), res.struct()
    psi, pso, prr = C.byref(si), C.byref(so), C.byref(rr)
    wp = C.c_void_p(work.data_ptr())
    h = ctx.h
    sd = C.c_uint64(seed)

    def prepare():
        rc = L.hepem_step_prepare(h, psi, wp, sh)
        assert rc == 0, rc

    def run():
        rc = L.hepem_step_run(h, sd, psi, pso, prr, wp, sh)
        assert rc == 0, rc

    for _ in range(warmup):
        prepare()
        run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t_start, t_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for k in range(steps):
        prepare()
        ev[k][0].record(stream)
        run()
        ev[k][1].record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_time(t_end)
    run_ms = [a.elapsed_time(b) for a, b in ev]
    return total_ms / steps, float(np.mean(run_ms)), wall


def time_step_graph(ctx, tin, tout, res, work, steps, seed):
    """One step (hepem_step_prepare + hepem_step_run) captured in a hipGraph
    and replayed `steps` times, timed with HIP events on the replay stream.
    The step is out of place, so every replay repeats the same step: the last
    replay's outputs are compared with an eager step's afterwards. Returns
    (ms per step, identical)."""
    import torch
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_outputs_differ
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            ctx.step_prepare(tin, work, stream=s)
            ctx.step_run(seed, tin, tout, res, work, stream=s)
    torch.cuda.synchronize()
    for _ in range(2):
        g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        a.record(s)
        for _ in range(steps):
            g.replay()
        b.record(s)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    ref_t, ref_r = TrackSoA(tin.n, "torch", tin["ekin"].device), StepOutSoA(tin.n, "torch", tin["ekin"].device)
    ctx.step(seed, tin, ref_t, ref_r, work)
    torch.cuda.synchronize()
    same = not step_outputs_differ(tin["ptype"], tout, res, ref_t, ref_r)
    return ms, same


def time_split(ctx, tin, res, work, steps, warmup, seed):
    """Times hepem_howfar + hepem_perform on a copy of `tin` (the split calls
    step in place; the caller's geometry would run between them: here it
    changes nothing). The copy back to the original tracks runs before each
    timed pair, outside its events. Returns ms per HowFar + Perform."""
    import ctypes as C

    import torch
    from g4hepem_amd import lib
    from g4hepem_amd.tracks import HowFarSoA, TrackSoA
    L = lib()
    dev = tin["ekin"].device
    t = TrackSoA(tin.n, "torch", dev)
    hs = HowFarSoA(tin.n, "torch", dev)
    stream = torch.cuda.current_stream()
    sh = C.c_void_p(int(stream.cuda_stream))
    st, shs, rr = t.struct(), hs.struct(), res.struct()
    pt, phs, prr = C.byref(st), C.byref(shs), C.byref(rr)
    wp = C.c_void_p(work.data_ptr())
    sd = C.c_uint64(seed)
    pairs = []
    for k in range(warmup + steps):
        for name in t.a:
            t.a[name].copy_(tin.a[name])
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        rc = L.hepem_howfar(ctx.h, sd, pt, phs, wp, sh)
        assert rc == 0, rc
        rc = L.hepem_perform(ctx.h, sd, pt, phs, prr, wp, sh)
        assert rc == 0, rc
        b.record(stream)
        if k >= warmup:
            pairs.append((a, b))
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in pairs]))


def config3_batches(seed: int, nbatch: int = 64):
    """64 batches, sizes log-uniform in [1e3, 1e6], per-track distributions
    as the headline workload; track ids run on across batches."""
    from g4hepem_amd.tracks import step_tracks
    rng = np.random.Generator(np.random.Philox(seed ^ 0x3C3))
    sizes = np.exp(rng.uniform(math.log(1e3), math.log(1e6), nbatch)).astype(np.int64)
    out, first = [], 0
    for k, n in enumerate(sizes):
        out.append(step_tracks(int(n), seed + 17 * k, len(SET2), first_id=first))
        first += int(n)
    return out


def verify_batches(ctx, batches, seed) -> dict:
    """After the timed replays, outside the timed region: every batch's outputs
    from the plan against hepem_step on the same batch and seed, bit for bit
    (compared on the device)."""
    import torch
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_outputs_differ
    dev = batches[0][0]["ekin"].device
    work = ctx.work_buffer(max(b[0].n for b in batches), dev)
    bad = {}
    for k, (tin, tout, res) in enumerate(batches):
        t_ref, r_ref = TrackSoA(tin.n, "torch", dev), StepOutSoA(tin.n, "torch", dev)
        ctx.step(seed, tin, t_ref, r_ref, work)
        diff = step_outputs_differ(tin["ptype"], tout, res, t_ref, r_ref)
        if diff:
            bad[k] = diff
        del t_ref, r_ref
    torch.cuda.synchronize()
    return {"bit_identical_to_unbatched": not bad, "batches_checked": len(batches),
            "mismatches": {str(k): v for k, v in list(bad.items())[:4]}}


def _shape_name(block, tile, persistent):
    return f"b{block}-{'sort' + str(tile) if tile else 'fused'}-{'persistent' if persistent else 'dispatched'}"


def sweep_batched(plan, seed, reps=4) -> dict:
    """ms per direct run of `plan` (HIP events, default stream, after one
    untimed run) for every launch shape the tuner chooses from."""
    import torch
    from g4hepem_amd import BatchPlan
    out = {}
    for block, tile, pers in BatchPlan.shapes():
        plan.set_shape(block, tile, pers)
        plan.run(seed)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            plan.run(seed)
        b.record()
        torch.cuda.synchronize()
        out[_shape_name(block, tile, pers)] = a.elapsed_time(b) / reps
    return out


def time_batched(ctx, batches, seed, reps, sweep=False):
    """A batch plan over `batches` (device SoA triples): tune the launch shape,
    capture one run into a hipGraph and time `reps` replays with HIP events on
    the replay stream; then check the last replay's outputs against hepem_step
    on each batch. With `sweep`, first time every shape the tuner chooses from.
    Returns (ms per run, chosen shape, check, sweep or None)."""
    import torch
    from g4hepem_amd import BatchPlan
    plan = BatchPlan(ctx, batches)
    swept = sweep_batched(plan, seed) if sweep else None
    plan.tune(seed)
    shape = plan.get_shape()
    shape["name"] = _shape_name(shape["block"], shape["sort_tile"], shape["persistent"])
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            plan.run(seed, stream=s)
    torch.cuda.synchronize()
    for _ in range(2):
        g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        a.record(s)
        for _ in range(reps):
            g.replay()
        b.record(s)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    plan.close()
    return ms, shape, verify_batches(ctx, batches, seed), swept


def run_config3(ctx64, ctx32, seed: int, reps: int) -> dict:
    """The 64-batch workload in both forms: tuned plan, graph replays, the
    per-shape sweep and the bit-identity check against unbatched steps."""
    import torch
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA
    dev = torch.device("cuda:0")
    host_batches = config3_batches(seed)
    ntot = sum(len(b["ekin"]) for b in host_batches)
    dev_batches = [(TrackSoA.from_numpy(b, "torch", dev), TrackSoA(len(b["ekin"]), "torch", dev),
                    StepOutSoA(len(b["ekin"]), "torch", dev)) for b in host_batches]
    del host_batches
    ms_b64, shape64, chk64, sw64 = time_batched(ctx64, dev_batches, seed, reps, sweep=True)
    ms_b32, shape32, chk32, sw32 = time_batched(ctx32, dev_batches, seed, reps, sweep=True)
    del dev_batches
    torch.cuda.empty_cache()
    return {
        "batches": 64, "tracks": ntot, "launch": "hipGraph replay of a tuned batch plan",
        "fp64": {"value": ntot / (ms_b64 * 1e-3), "unit": "track-steps/s", "ms_per_run": ms_b64,
                 "shape": shape64, **chk64, "shape_sweep_ms": sw64},
        "fp32_table_form": {"value": ntot / (ms_b32 * 1e-3), "unit": "track-steps/s", "ms_per_run": ms_b32,
                            "shape": shape32, **chk32, "shape_sweep_ms": sw32},
        "speedup_fp32_vs_fp64": ms_b64 / ms_b32,
    }


def _chi2_means(a, b):
    """Per-slab agreement of two runs over the same events (rows: events,
    columns: slabs). The runs share seeds and streams, so their events are
    paired and correlated: the statistic is the paired one, per slab the mean
    of the per-event differences over its standard error, squared and summed
    over the slabs whose differences vary. Independent runs give the same
    expectation (ndf)."""
    d = a - b
    se2 = d.var(0, ddof=1) / len(d)
    m = se2 > 0
    return {"chi2": float(((d.mean(0) ** 2)[m] / se2[m]).sum()), "ndf": int(m.sum()), "test": "paired"}


def _slab_setup(woodcock_min_ekin: float = 0.0):
    from g4hepem_amd._native import slab_setup
    return slab_setup(50, [2.3, 5.7], [0, 1], 0, 1.0e4, capacity=1 << 23, events_per_chunk=4096, refill_below=1 << 21,
                      woodcock_min_ekin=woodcock_min_ekin)


def _timed_showers(ctx, su, nevents: int, seed: int):
    ctx.slab_run(su, min(64, nevents), seed + 1)  # untimed: first launches of the driver's kernels
    t0 = time.perf_counter()
    g = ctx.slab_run(su, nevents, seed, per_event=True)
    wall = time.perf_counter() - t0
    return g, wall, {"events": nevents, "wall_s": wall, "events_per_s": nevents / wall,
                     "track_steps_per_s": g["steps"] / wall, "steps": g["steps"], "tracks": g["tracks"],
                     "iterations": g["iterations"], "max_in_flight": g["max_in_flight"], "electron_substeps": g["substeps"]}


def run_config5(ctx, tables, seed: int, nevents: int, cpu_events: int, ctx32=None, woodcock: bool = True):
    """10 GeV e- into 50 x (2.3 mm Pb + 5.7 mm lAr) at 0.7 mm: GPU showers of
    `nevents` events (wall clock of the whole driver, host syncs included), and
    the oracle's CPU harness on the first `cpu_events` events for the per-slab
    chi2 and the fraction of events identical to the GPU's. With `ctx32` (the
    FP32-table context) the same events in that form, and in both cases the
    chi2 of its per-slab means against the FP64 run; then the FP64 run again
    with Woodcock tracking of the gammas above 0.2 MeV."""
    su = _slab_setup()
    g, wall, out = _timed_showers(ctx, su, nevents, seed)
    ev = g["edep_event"] / 1e9
    mean, rms = ev.mean(0), ev.std(0)
    out.update({"mean_edep_total_MeV": float(ev.sum(1).mean()), "mean_leak_MeV": float(g["leak_event"].mean() / 1e9),
                "layer_mean_MeV": [round(float(v), 4) for v in mean], "layer_rms_MeV": [round(float(v), 4) for v in rms]})
    head32 = None
    if ctx32 is not None:
        g32, _, o32 = _timed_showers(ctx32, su, nevents, seed)
        e32 = g32["edep_event"] / 1e9
        o32.update({"mean_edep_total_MeV": float(e32.sum(1).mean()), "vs_fp64": _chi2_means(ev, e32),
                    "speedup_vs_fp64": o32["events_per_s"] / out["events_per_s"],
                    "layer_mean_MeV": [round(float(v), 4) for v in e32.mean(0)],
                    "layer_rms_MeV": [round(float(v), 4) for v in e32.std(0)]})
        out["fp32_table_form"] = o32
        head32 = g32["edep_event"][:cpu_events].copy() if cpu_events > 0 else None
        del g32
    if woodcock:
        gw, _, ow = _timed_showers(ctx, _slab_setup(0.2), nevents, seed)
        ew = gw["edep_event"] / 1e9
        ow.update({"woodcock_min_ekin_MeV": 0.2, "mean_edep_total_MeV": float(ew.sum(1).mean()),
                   "vs_normal_tracking": _chi2_means(ev, ew),
                   "speedup_vs_normal": ow["events_per_s"] / out["events_per_s"]})
        out["woodcock_fp64"] = ow
        del gw
    if cpu_events > 0:
        from oracle.oracle_py import Oracle
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        t0 = time.perf_counter()
        dep, leak, steps, _ = Oracle().slab_shower(tables, su, cpu_events, seed, nthreads=max(1, min(int(cores), 64)))
        cwall = time.perf_counter() - t0
        a, b = g["edep_event"][:cpu_events] / 1e9, dep / 1e9
        out["cpu_harness"] = {"events": cpu_events, "wall_s": cwall, "track_steps_per_s": steps / cwall,
                              **_chi2_means(a, b),
                              "identical_events": float(np.all(g["edep_event"][:cpu_events] == dep, axis=1).mean())}
        if head32 is not None:
            # the FP32-table form's first events against the CPU harness's (FP64) same events
            out["cpu_harness"]["fp32_table_form_vs_cpu"] = _chi2_means(head32 / 1e9, dep / 1e9)
    return out


def fp32_errors(r64: dict, t64: dict, r32: dict, t32: dict, ptype: np.ndarray) -> dict:
    """Config 4's error report: the FP32-table step against the FP64 step on
    the same tracks and streams. Integer decisions as agreement fractions. A
    track that drew another number of random numbers (a rejection test that
    FP32 rounding decided the other way) sampled another final state, so the
    per-track errors are over the tracks whose winner, secondary count and
    draw count agree; the totals over all tracks compare the samples
    statistically."""
    n = len(r64["winner"])
    same = (r64["winner"] == r32["winner"]) & (r64["nsec"] == r32["nsec"]) & (t64["rng_ctr"] == t32["rng_ctr"])

    def rel(a, b, m):
        m = m & (b != 0)
        if not m.any():
            return None
        r = np.abs(a[m] - b[m]) / np.abs(b[m])
        return {"max": float(r.max()), "mean": float(r.mean())}

    def total(a, b, m=None):
        m = np.ones(len(a), bool) if m is None else m
        return float(abs(a[m].sum() - b[m].sum()) / abs(b[m].sum()))

    d64, d32 = t64["dir"].reshape(3, n), t32["dir"].reshape(3, n)
    ddir_all = np.abs(d64 - d32).max(axis=0)
    ddir = ddir_all[same]
    s64, s32 = r64["sec_ekin"].reshape(2, n)[0], r32["sec_ekin"].reshape(2, n)[0]
    has64, has32 = r64["nsec"] >= 1, r32["nsec"] >= 1

    # the same errors per particle kind and winner process, to trace the tails to a branch
    is_g = ptype == 2
    by_regime = {}
    for kind, km in (("e-/e+", ~is_g), ("gamma", is_g)):
        for w in np.unique(r64["winner"][km]):
            m = same & km & (r64["winner"] == w)
            if not m.any():
                continue
            key = f"{kind} winner {int(w)}"
            by_regime[key] = {"tracks": int(m.sum()), "edep": rel(r32["edep"], r64["edep"], m),
                              "post_step_ekin": rel(t32["ekin"], t64["ekin"], m),
                              "true_step": rel(r32["true_step"], r64["true_step"], m),
                              "post_step_dir_max_abs": float(ddir_all[m].max())}
    return {
        "winner_agreement": float((r64["winner"] == r32["winner"]).mean()),
        "nsec_agreement": float((r64["nsec"] == r32["nsec"]).mean()),
        "draw_count_agreement": float((t64["rng_ctr"] == t32["rng_ctr"]).mean()),
        "compared_tracks": int(same.sum()),
        "true_step": rel(r32["true_step"], r64["true_step"], same),
        "geom_step": rel(r32["geom_step"], r64["geom_step"], same),
        "edep": rel(r32["edep"], r64["edep"], same),
        "post_step_ekin": rel(t32["ekin"], t64["ekin"], same),
        "post_step_dir_max_abs": float(ddir.max()) if ddir.size else None,
        "post_step_dir_mean_abs": float(ddir.mean()) if ddir.size else None,
        "secondary0_ekin": rel(s32, s64, same & has64),
        "total_edep_rel": total(r32["edep"], r64["edep"]),
        "total_secondary0_ekin_rel": float(abs(s32[has32].sum() - s64[has64].sum()) / s64[has64].sum()),
        "by_regime": by_regime,
    }


def time_lookups(ctx, n, reps, seed=2026):
    import torch
    from g4hepem_amd.tracks import lookup_queries
    imc, e, le = lookup_queries(n, seed, len(SET1))
    d = torch.device("cuda")
    timc, te, tle = (torch.from_numpy(np.ascontiguousarray(x)).to(d) for x in (imc, e, le))
    out = torch.empty(5 * n, dtype=torch.float64, device=d)
    for _ in range(3):
        ctx.lookup(0, timc, te, tle, out)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        ctx.lookup(0, timc, te, tle, out)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    return n / (ms * 1e-3), ms


def time_lookups_device(ctx, n, reps, seed=2026):
    """time_lookups for large n with the queries drawn on the device (torch's
    Philox generator; same distributions as lookup_queries)."""
    import math
    import torch
    d = torch.device("cuda")
    gen = torch.Generator(device=d)
    gen.manual_seed(seed)
    lo, hi = math.log(1e-3), math.log(1e8)
    tle = torch.rand(n, generator=gen, device=d, dtype=torch.float64) * (hi - lo) + lo
    te = torch.exp(tle)
    timc = torch.randint(0, len(SET1), (n,), generator=gen, device=d, dtype=torch.int32)
    out = torch.empty(5 * n, dtype=torch.float64, device=d)
    for _ in range(3):
        ctx.lookup(0, timc, te, tle, out)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        ctx.lookup(0, timc, te, tle, out)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del timc, te, tle, out
    torch.cuda.empty_cache()
    return n / (ms * 1e-3), ms


def cpu_baseline(tables, tracks: dict, min_seconds: float, seed: int):
    """The path on the host cores: the reference's own run-time code compiled
    from its sources (oracle/_ref, kind "reference") when that library is
    present, else the oracle's C restatement (kind "port"); both are timed
    when both exist. Repeated passes over the first 250k tracks of the same
    batch until min_seconds of work are done."""
    from oracle.oracle_py import Oracle, Ref, ref_available
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(int(cores), 64))
    n_full = len(tracks["ekin"])
    n_sample = min(n_full, 250_000)
    sample = {}
    for k, v in tracks.items():
        c = len(v) // n_full
        sample[k] = np.concatenate([v[i * n_full: i * n_full + n_sample] for i in range(c)])

    def timed(engine):
        t0 = time.perf_counter()
        done = 0
        while True:
            engine.step(tables, seed, sample, nthreads=cores)
            done += n_sample
            if time.perf_counter() - t0 >= min_seconds:
                break
        return done, time.perf_counter() - t0

    runs = [("port", Oracle())]
    if ref_available():
        runs.insert(0, ("reference", Ref()))
    out = None
    for kind, engine in runs:
        done, dt = timed(engine)
        rec = {"value": done / dt, "unit": "track-steps/s", "cores": cores, "kind": kind,
               "sample": f"{done // n_sample} passes over the first {n_sample} tracks of the same batch "
                         f"({done} track-steps, {dt:.1f} s, {cores} threads)"}
        if out is None:
            out = rec
        else:
            out["also_" + kind] = {"value": rec["value"], "sample": rec["sample"]}
    return out


def log(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0x5EED2026)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline FP64 step only (for profiling)")
    ap.add_argument("--shower-events", type=int, default=10000, help="config-5 events (0: skip)")
    ap.add_argument("--shower-cpu-events", type=int, default=1000, help="config-5 events for the CPU harness")
    ap.add_argument("--shower-forms", default="fp64,fp32,woodcock",
                    help="config-5 runs: the FP64 form (always), the FP32-table form, FP64 with Woodcock tracking")
    ap.add_argument("--lookups-only", action="store_true",
                    help="config-1 lookups only, staged and from the table image (for profiling); prints their JSON object")
    ap.add_argument("--batched-only", action="store_true", help="config-3 batch plans only (tuner sweep included)")
    ap.add_argument("--fp32-only", action="store_true", help="the config-2 step in the FP32-table form only (for profiling)")
    ap.add_argument("--shower-only", action="store_true",
                    help="config-5 slab showers only, without the CPU harness (for profiling); prints their JSON object")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="before the steps, copy a buffer of this many MiB device-to-device "
                         "(reference kernel of known traffic for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by track batch, with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)

    import torch
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, Context, Tables
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_tracks, type_ordered

    assert torch.cuda.is_available(), "no GPU visible"
    dev = torch.device("cuda:0")
    if args.lookups_only:
        ctx1 = Context(Tables(SET1, 0.7), HEPEM_FP64)
        q10m, ms10m = time_lookups(ctx1, 10_000_000, 20)
        q100m, ms100m = time_lookups_device(ctx1, 100_000_000, 10)
        ctx1.set_lookup_staging(False)
        g10m, gms10m = time_lookups(ctx1, 10_000_000, 20)
        print(json.dumps({"config1_lookups": {"queries_per_s_10M": q10m, "ms_10M": ms10m,
                                              "queries_per_s_100M": q100m, "ms_100M": ms100m,
                                              "tables_in_image": {"queries_per_s_10M": g10m, "ms_10M": gms10m}}}),
              flush=True)
        return
    if args.batched_only:
        tables2 = Tables(SET2, 0.7)
        print(json.dumps({"config3_batched": run_config3(Context(tables2, HEPEM_FP64), Context(tables2, HEPEM_FP32T),
                                                         args.seed, max(3, args.steps // 4))}), flush=True)
        return
    if args.shower_only:
        tables_slab = Tables(SET1, 0.7)   # couples 0: G4_Pb, 1: G4_lAr
        forms = args.shower_forms.split(",")
        c5 = run_config5(Context(tables_slab, HEPEM_FP64), tables_slab, args.seed, args.shower_events, 0,
                         Context(tables_slab, HEPEM_FP32T) if "fp32" in forms else None, "woodcock" in forms)
        print(json.dumps({"config5_slab_shower": c5}), flush=True)
        return
    tables2 = Tables(SET2, 0.7)
    n = args.n
    tracks = step_tracks(n, args.seed, len(SET2))
    tin = TrackSoA.from_numpy(tracks, "torch", dev)
    tout = TrackSoA(n, "torch", dev)
    res = StepOutSoA(n, "torch", dev)
    work = Context.work_buffer(n, dev)

    if args.calibration_copy_mib > 0:
        nb = args.calibration_copy_mib * 1024 * 1024
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.fp32_only:
        ms32, run32, _ = time_steps(Context(tables2, HEPEM_FP32T), tin, tout, res, work, args.steps, args.warmup, args.seed)
        print(json.dumps({"fp32_table_form_step": {"tracks": n, "ms_per_step": ms32, "kernel_ms": run32,
                                                   "value": n / (ms32 * 1e-3), "unit": "track-steps/s"}}), flush=True)
        return
    log("headline FP64 step")
    ctx64 = Context(tables2, HEPEM_FP64)
    ms64, run64, wall64 = time_steps(ctx64, tin, tout, res, work, args.steps, args.warmup, args.seed)
    value = n / (ms64 * 1e-3)
    nsec_total = int(res["nsec"].sum().item())
    abytes = algorithmic_bytes(tracks["ptype"], nsec_total)
    achieved = abytes / (run64 * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("step_run_fp64_1M_pmc.md")
    line = {
        "metric": "track-steps/s on 1 MI355X, FP64 vs FP32-table form, + HBM GB/s vs peak",
        "value": value,
        "unit": "track-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox inputs; tables from the project's offline builder)",
        "config": {
            "workload": "config2: one HowFar+Perform step of 1M fresh tracks, e-:e+:gamma=4:1:5, "
                        "E log-U[100 keV,10 GeV], couples U{G4_PbWO4,G4_Pb,G4_lAr,G4_Si} @0.7 mm, FP64",
            "tracks": n, "wall_s": round(wall64, 4),
        },
        "roofline": {
            "bound": "hbm", "kernel": "hepem_step_run = k_step_select_local<double> + k_classify_tiles + k_interact_tiles<double>",
            "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "kernel_ms": run64, "algorithmic_bytes_per_launch": abytes,
        },
    }
    if not args.no_extras:
        log("FP32-table step")
        r64, t64 = res.to_numpy(), tout.to_numpy()   # results of the last FP64 step
        ctx32 = Context(tables2, HEPEM_FP32T)
        ms32, run32, _ = time_steps(ctx32, tin, tout, res, work, args.steps, args.warmup, args.seed)
        line["fp32_table_form"] = {"value": n / (ms32 * 1e-3), "unit": "track-steps/s", "ms_per_step": ms32,
                                   "kernel_ms": run32, "speedup_vs_fp64": ms64 / ms32,
                                   "errors_vs_fp64": fp32_errors(r64, t64, res.to_numpy(), tout.to_numpy(),
                                                                 tracks["ptype"])}
        del r64, t64
        log("config-2 tracks in type order")
        tin_o = TrackSoA.from_numpy(type_ordered(tracks), "torch", dev)
        ordered = {"layout": "the same tracks, e-/e+ first then gammas (as hepem_stack_advance lays out a batch)"}
        for form, ctx in (("fp64", ctx64), ("fp32_table_form", ctx32)):
            mso, runo, _ = time_steps(ctx, tin_o, tout, res, work, args.steps, args.warmup, args.seed)
            ordered[form] = {"value": n / (mso * 1e-3), "unit": "track-steps/s", "ms_per_step": mso, "kernel_ms": runo}
        line["config2_type_ordered"] = ordered
        del tin_o
        log("config-2 step replayed from a hipGraph")
        graph = {"launch": "hepem_step_prepare + hepem_step_run captured once, replayed"}
        try:
            for form, ctx in (("fp64", ctx64), ("fp32_table_form", ctx32)):
                msg, same = time_step_graph(ctx, tin, tout, res, work, args.steps, args.seed)
                graph[form] = {"value": n / (msg * 1e-3), "unit": "track-steps/s", "ms_per_step": msg,
                               "identical_to_eager": same}
        except RuntimeError as e:   # a capture the runtime refuses is reported, not fatal to the line
            graph["error"] = str(e)
            torch.cuda.synchronize()
        line["config2_graph_replay"] = graph
        log("config-2 tracks through the split HowFar and Perform calls")
        split = {"calls": "hepem_howfar, then hepem_perform (in place, no geometry between them)"}
        for form, ctx in (("fp64", ctx64), ("fp32_table_form", ctx32)):
            mss = time_split(ctx, tin, res, work, args.steps, args.warmup, args.seed)
            split[form] = {"value": n / (mss * 1e-3), "unit": "track-steps/s", "ms_per_step": mss}
        line["config2_split_calls"] = split
        log("config-1 lookups")
        tables1 = Tables(SET1, 0.7)
        ctx1 = Context(tables1, HEPEM_FP64)
        q100k, ms100k = time_lookups(ctx1, 100_000, 200)
        q10m, ms10m = time_lookups(ctx1, 10_000_000, 20)
        sweep = {}
        for nq in (30_000_000, 100_000_000):
            qs, mss = time_lookups_device(ctx1, nq, 10)
            sweep[f"{nq // 1_000_000}M"] = {"queries_per_s": qs, "ms": mss, "achieved_GBs": 60.0 * nq / (mss * 1e-3) / 1e9}
        ctx1.set_lookup_staging(False)  # the same lookups with the tables read from the image
        g10m, gms10m = time_lookups(ctx1, 10_000_000, 20)
        ctx1.set_lookup_staging(True)
        line["config1_lookups"] = {"queries_per_s_100k": q100k, "ms_100k": ms100k,
                                   "queries_per_s_10M": q10m, "ms_10M": ms10m,
                                   "achieved_GBs_10M": 60.0 * 10_000_000 / (ms10m * 1e-3) / 1e9,
                                   "sweep_device_drawn": sweep,
                                   "tables_in": "LDS",
                                   "tables_in_image": {"queries_per_s_10M": g10m, "ms_10M": gms10m}}
        del tin, tout, res
        torch.cuda.empty_cache()
        log("config-3 batches")
        line["config3_batched"] = run_config3(ctx64, ctx32, args.seed, max(3, args.steps // 4))
    if not args.no_extras and args.shower_events > 0:
        log("config-5 slab showers")
        tables_slab = Tables(SET1, 0.7)   # couples 0: G4_Pb, 1: G4_lAr
        forms = args.shower_forms.split(",")
        line["config5_slab_shower"] = run_config5(Context(tables_slab, HEPEM_FP64), tables_slab, args.seed,
                                                  args.shower_events, 0 if args.no_cpu else args.shower_cpu_events,
                                                  Context(tables_slab, HEPEM_FP32T) if "fp32" in forms else None,
                                                  "woodcock" in forms)
    if not args.no_cpu:
        log("CPU baseline (oracle)")
        line["cpu_baseline"] = cpu_baseline(tables2, tracks, args.cpu_seconds, args.seed)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
