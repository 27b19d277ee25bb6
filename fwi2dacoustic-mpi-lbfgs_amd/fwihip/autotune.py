"""Launch-parameter tuner of the batched gradient.

For one workload shape (grid, shot count, receivers, nt) it times short
forward + adjoint passes of each candidate setting and keeps the fastest:

    kernels      0: 2-D tile kernels; 1: z-marching kernels
    tile_z       rows per workgroup tile of the z-marching kernels
    group        shots per workgroup of the z-marching kernels
    rec_group    shots per workgroup of the tile imaging kernel
    bwd_chunk    shots per backward pass (0: all in one pass)
    interior_wgs interior workgroups per CU beside the general strips (0: no cap)
    graph        time steps per captured HIP graph block

The probe runs the workload's own shots truncated to a few hundred steps, so
the per-step cost it measures is the one of the full run (the step cost does
not depend on nt).  Choices are cached in a JSON file keyed by the shape.
"""
import json
import os
import time

import numpy as np

from . import lib as L
from .workloads import Shot

DEFAULT = dict(kernels=0, tile_z=0, group=0, rec_group=8, bwd_chunk=0, interior_wgs=3, graph=100)


def candidates(nshots):
    """The settings tried: tile kernels over imaging group, backward chunk and
    graph block; z-marching kernels over tile height and shot group."""
    out = []
    chunks = sorted({0} | {c for c in (16, 32) if c < nshots})
    for rg in (4, 8, 16):
        for bc in chunks:
            for iw in (0, 3):
                for gs in (50, 100):
                    out.append(dict(kernels=0, tile_z=0, group=0, rec_group=rg, bwd_chunk=bc, interior_wgs=iw, graph=gs))
    for tz in (16, 32):
        for sg in (2, 4):
            out.append(dict(kernels=1, tile_z=tz, group=sg, rec_group=8, bwd_chunk=0, interior_wgs=0, graph=100))
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
    ctx.set_interior_wgs(params.get("interior_wgs", DEFAULT["interior_wgs"]))
    ctx.set_graph(params["graph"])


def _probe(geom, vel, shots, params, nt_probe, reps):
    short = [Shot(nt=min(s.nt, nt_probe), isx=s.isx, isz=s.isz, wav=s.wav[:min(s.nt, nt_probe)], irx=s.irx,
                  irz=s.irz) for s in shots]
    ctx = make_context(geom, short, params, nt_probe)
    try:
        ctx.set_model(vel)
        ctx.set_shots(short)
        zeros = np.zeros(ctx.ntrace(), np.float32)
        best = float("inf")
        for r in range(reps + 1):
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.forward(want_traces=False)
            ctx.set_residual(zeros)
            ctx.adjoint(want=False)
            ctx.synchronize()
            if r:                          # the first pass captures the graphs
                best = min(best, time.perf_counter() - t0)
        steps = max(s.nt for s in short)
        return best / steps
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
    results = []
    for p in candidates(len(shots)):
        try:
            t = _probe(geom, vel, shots, p, nt_probe, reps)
        except L.FwiError as e:       # e.g. a setting the device memory cannot hold
            if verbose:
                print(f"autotune: {p} skipped ({e})", flush=True)
            continue
        results.append((t, p))
        if verbose:
            print(f"autotune: {p} {t * 1e6:.1f} us/step", flush=True)
    if not results:
        return dict(DEFAULT), None, False
    t, p = min(results, key=lambda r: r[0])
    if cache_path:
        cache[key] = dict(params=p, s_per_step=t)
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, cache_path)
    return dict(p), t, False
