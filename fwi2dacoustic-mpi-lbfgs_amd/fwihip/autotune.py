"""Launch-parameter tuner of the batched gradient.

For one workload shape (grid, shot count, receivers, nt) it times short
forward + adjoint passes of each candidate setting and keeps the fastest:

    kernels      0: 2-D tile kernels; 1: z-marching kernels
    strip_fwd    rows per wave of the tile kernels in the forward pass (4, 8)
    wgs_fwd      interior workgroups per CU beside the general strips, forward
    strip_bwd    rows per wave of the tile kernels in the backward pass (4, 8)
    wgs_bwd      the same for the backward pass
    tile_z       rows per workgroup tile of the z-marching kernels
    group        shots per workgroup of the z-marching kernels
    rec_group    shots per workgroup of the tile imaging kernel
    bwd_chunk    shots per backward pass (0: all in one pass; -1: the library's
                 choice by cache fit)
    bwd_lanes    backward chunks in flight (2: in pairs on two stream sets; 1)
    graph        time steps per captured HIP graph block

The probe runs the workload's own shots truncated to a few hundred steps, so
the per-step cost it measures is the one of the full run (the step cost does
not depend on nt).  The search is staged: first the tile shape of each pass
(strip rows with the interior cap; the forward pass and the backward pass
are timed apart, so each takes its own best), then, for that shape, the
imaging group, backward chunk and graph block; the z-marching kernels are
probed beside them.  Choices are cached in a JSON file keyed by the shape.
"""
import json
import os
import time

import numpy as np

from . import lib as L
from .workloads import Shot

DEFAULT = dict(kernels=0, strip_fwd=4, wgs_fwd=6, strip_bwd=4, wgs_bwd=4, tile_z=0, group=0, rec_group=8,
               bwd_chunk=-1, bwd_lanes=2, graph=100)
# stage 1: (rows, interior workgroups per CU) of a pass
PASS_SHAPES = ((8, 3), (8, 4), (4, 4), (4, 5), (4, 6))


def shape_candidates():
    """Stage 1: the same shape in both passes (each pass timed apart)."""
    return [dict(DEFAULT, strip_fwd=r, wgs_fwd=w, strip_bwd=r, wgs_bwd=w) for r, w in PASS_SHAPES]


def launch_candidates(nshots, best_shape):
    """Stage 2: imaging group, backward chunk (in pairs, and one after another
    at the library's chunk size) and graph block for the stage-1 choice; the
    z-marching kernels over tile height and shot group."""
    out = []
    chunks = sorted({-1, 0} | {c for c in (16, 32) if c < nshots})
    for rg in (4, 8, 16):
        for bc, lanes in [(c, 2) for c in chunks] + [(-1, 1)]:
            for gs in (50, 100):
                out.append(dict(best_shape, rec_group=rg, bwd_chunk=bc, bwd_lanes=lanes, graph=gs))
    for tz in (16, 32):
        for sg in (2, 4):
            out.append(dict(DEFAULT, kernels=1, tile_z=tz, group=sg, wgs_fwd=0, wgs_bwd=0))
    return out


def candidates(nshots):
    """Every setting the staged search can return."""
    out = []
    for f in PASS_SHAPES:
        for b in PASS_SHAPES:
            shape = dict(DEFAULT, strip_fwd=f[0], wgs_fwd=f[1], strip_bwd=b[0], wgs_bwd=b[1])
            for p in [shape] + launch_candidates(nshots, shape):
                if p not in out:
                    out.append(p)
    return out


def shape_key(geom, shots):
    nr = max(s.nr for s in shots)
    return f"{geom.nx}x{geom.nz}p{geom.npml}_s{len(shots)}_r{nr}"


def make_context(geom, shots, params, max_nt):
    nr = max(s.nr for s in shots)
    ctx = L.Context(geom, max_shots=len(shots), max_nt=max_nt, max_nr=nr, tile_z=params["tile_z"],
                    shots_per_group=params["group"])
    apply(ctx, params)
    return ctx


def apply(ctx, params):
    ctx.set_kernels(params["kernels"])
    ctx.set_rec_group(params["rec_group"])
    ctx.set_bwd_chunk(params["bwd_chunk"])
    ctx.set_bwd_lanes(params.get("bwd_lanes", DEFAULT["bwd_lanes"]))
    ctx.set_pass_shape(0, params.get("strip_fwd", DEFAULT["strip_fwd"]), params.get("wgs_fwd", DEFAULT["wgs_fwd"]))
    ctx.set_pass_shape(1, params.get("strip_bwd", DEFAULT["strip_bwd"]), params.get("wgs_bwd", DEFAULT["wgs_bwd"]))
    ctx.set_graph(params["graph"])


def _probe(geom, vel, shots, params, nt_probe, reps):
    short = [Shot(nt=min(s.nt, nt_probe), isx=s.isx, isz=s.isz, wav=s.wav[:min(s.nt, nt_probe)], irx=s.irx,
                  irz=s.irz) for s in shots]
    ctx = make_context(geom, short, params, nt_probe)
    try:
        ctx.set_model(vel)
        ctx.set_shots(short)
        zeros = np.zeros(ctx.ntrace(), np.float32)
        best_f = best_b = float("inf")
        for r in range(reps + 1):
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.forward(want_traces=False)
            ctx.synchronize()
            t1 = time.perf_counter()
            ctx.set_residual(zeros)
            ctx.adjoint(want=False)
            ctx.synchronize()
            if r:                          # the first pass captures the graphs
                best_f = min(best_f, t1 - t0)
                best_b = min(best_b, time.perf_counter() - t1)
        steps = max(s.nt for s in short)
        return best_f / steps, best_b / steps
    finally:
        ctx.close()


def tune(geom, vel, shots, nt_probe=300, reps=2, cache_path=None, verbose=False):
    """The fastest setting for this shape (seconds per time step of forward +
    adjoint), from the cache when present."""
    key = shape_key(geom, shots)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except (OSError, ValueError):
            cache = {}
    if key in cache:
        return dict(cache[key]["params"]), cache[key]["s_per_step"], True
    results = []   # (forward s/step, backward s/step, params)

    def run(cands):
        for p in cands:
            if any(p == q for _, _, q in results):
                continue
            try:
                tf, tb = _probe(geom, vel, shots, p, nt_probe, reps)
            except L.FwiError as e:       # e.g. a setting the device memory cannot hold
                if verbose:
                    print(f"autotune: {p} skipped ({e})", flush=True)
                continue
            results.append((tf, tb, p))
            if verbose:
                print(f"autotune: {p} forward {tf * 1e6:.1f} + backward {tb * 1e6:.1f} us/step", flush=True)

    run(shape_candidates())
    if not results:
        return dict(DEFAULT), None, False
    # each pass takes the shape it ran fastest with
    pf = min(results, key=lambda r: r[0])[2]
    pb = min(results, key=lambda r: r[1])[2]
    shape = dict(DEFAULT, strip_fwd=pf["strip_fwd"], wgs_fwd=pf["wgs_fwd"], strip_bwd=pb["strip_bwd"],
                 wgs_bwd=pb["wgs_bwd"])
    run(launch_candidates(len(shots), shape))
    tf, tb, p = min(results, key=lambda r: r[0] + r[1])
    t = tf + tb
    if cache_path:
        cache[key] = dict(params=p, s_per_step=t)
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, cache_path)
    return dict(p), t, False
