"""Benchmark of the batched FWI gradient on one MI355X.

A step is one FWI gradient over the whole shot batch of the 1024x320
workload (dx=dz=10 m, dt=0.8 ms, nt=4000, 64 shots, 1024 receivers, 20-cell
CPML, free surface): model upload, forward propagation of every shot,
L2 misfit and residual on the device (observed data resident in HBM),
exact-adjoint back-propagation with source reconstruction and FP64 imaging,
gradient copied back.  Synthetic seeded layered/faulted model.

Prints ONE JSON line (rank 0):
  value       shot-gradients/s (whole job)
  roofline    dominant kernel's algorithmic bytes / its HIP-event duration,
              against the 8 TB/s HBM3E peak.  Algorithmic bytes: 16 B per
              wavefield cell-update (read p^n, p^{n-1}, ekp; write p^{n+1}),
              over the cells the kernel updates (propagation grid for the
              forward and adjoint, physical grid for the reconstruction), times
              the shots; CPML memory, frames, traces and the image are not
              credited (they are listed as min_traffic_bytes)
  cpu_baseline  the FP32 OpenMP oracle (a restatement of the reference CPU
              path) on a bounded sample of the same workload, on this host
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fwi2dacoustic-mpi-lbfgs_amd")
sys.path.insert(0, PKG)

HBM_PEAK = 8.0e12          # MI355X datasheet HBM3E bandwidth, B/s
BYTES_PER_CELL_UPDATE = 16  # read p^n, p^{n-1}, ekp; write p^{n+1} (FP32)
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_1024x320_64shots.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=None, help="shots (default: 64; 16 with --config 4; 96 uneven)")
    ap.add_argument("--nt", type=int, default=None, help="time steps (default: 4000; 10000 with --config 4)")
    ap.add_argument("--config", type=int, default=2, choices=(2, 4),
                    help="batched workload: 2 = 1024x320 (the headline), 4 = 4096x1024, dx=dz=5 m, dt=0.4 ms")
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--group", type=int, default=0, help="shots per workgroup of the z-marching kernels")
    ap.add_argument("--kernels", type=int, default=0, help="0: 2-D tile kernels, 1: z-marching kernels")
    ap.add_argument("--rec-group", type=int, default=0, help="shots per workgroup of the tile imaging kernel")
    ap.add_argument("--bwd-chunk", type=int, default=-1,
                    help="shots per backward pass (0: all; -1: the library's choice by cache fit)")
    ap.add_argument("--bwd-gen-rows", type=int, default=-1,
                    help="rows per wave of the backward step's general side under 4-row strips (4 or 8; -1: default)")
    ap.add_argument("--bwd-lanes", type=int, default=-1,
                    help="backward chunks in flight (2: in pairs on two stream sets, 1: one after another; -1: default)")
    ap.add_argument("--bwd-overlap", type=int, default=-1, help="adjoint beside imaging on two streams (1/0; -1: default)")
    ap.add_argument("--concurrent", type=int, default=-1,
                    help="general strips on a second stream beside the interior launch (1/0; -1: default)")
    ap.add_argument("--interior-wgs", type=int, default=-1,
                    help="interior workgroups per CU beside the general strips, both passes (0: no cap; -1: the "
                         "defaults, 6 forward and 4 backward)")
    ap.add_argument("--fwd-shape", type=int, nargs=2, default=None, metavar=("ROWS", "WGS"),
                    help="forward pass: rows per wave of the tile kernels (4 or 8) and interior workgroups per CU")
    ap.add_argument("--bwd-shape", type=int, nargs=2, default=None, metavar=("ROWS", "WGS"),
                    help="backward pass: rows per wave of the tile kernels (4 or 8) and interior workgroups per CU")
    ap.add_argument("--fwd-mixed", type=int, default=-1,
                    help="the forward step as one launch of interior and general strips (1/0; -1: default)")
    ap.add_argument("--bwd-mixed", type=int, default=-1,
                    help="the fused backward step as one launch (1/0; -1: default)")
    ap.add_argument("--bwd-fused", type=int, default=-1,
                    help="one fused adjoint + reconstruction + imaging kernel per backward step (1/0; -1: default 1)")
    ap.add_argument("--ckpt-every", type=int, default=100, help="snapshot interval of --workload checkpoint")
    ap.add_argument("--misfit", default="ncc", choices=("l2", "ncc", "eot"), help="misfit of --workload inversion")
    ap.add_argument("--iters", type=int, default=20, help="L-BFGS iterations of --workload inversion")
    ap.add_argument("--storage", default="fp32", choices=("fp32", "fp16", "bf16"),
                    help="snapshot storage of --workload inversion (with --ckpt-every; 0 there: boundary saving)")
    ap.add_argument("--rec-step", type=int, default=1, help="receiver spacing (columns) of --workload inversion")
    ap.add_argument("--jump", type=int, default=1, help="data sampled every jump steps in --workload inversion")
    ap.add_argument("--workload", default="batched", choices=("batched", "uneven", "checkpoint", "inversion"),
                    help="batched: the 1024x320 64-shot gradient (the headline line); uneven: 1536x400, "
                         "96 shots with nt 2000-6000, apertures 256-1536, host Butterworth band-pass")
    ap.add_argument("--autotune", action="store_true",
                    help="time short passes of the candidate launch settings first and use the fastest")
    ap.add_argument("--log-dir", default=None,
                    help="folder of the per-iteration log of --workload inversion (default: the temporary folder)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-single", action="store_true", help="skip the one-shot-per-launch comparison")
    ap.add_argument("--cpu-nt", type=int, default=0, help="time steps of each CPU sample shot (0: the workload's nt)")
    ap.add_argument("--cpu-shots", type=int, default=8, help="shots of the CPU sample, spread along the line")
    ap.add_argument("--no-boundary", action="store_true",
                    help="skip the same gradient through the drop-in grad_all_shots (host misfit, trace copies)")
    ap.add_argument("--no-ckpt", action="store_true", help="skip the checkpointed (FP16/BF16 snapshot) backward pass")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.shots is None:
        args.shots = 96 if args.workload == "uneven" else (16 if args.config == 4 else 64)
    if args.nt is None:
        args.nt = 10000 if args.config == 4 else 4000
    return args


def algorithmic_bytes(g, nshots):
    """Credited bytes of one launch of each step kernel: 16 B per cell-update."""
    f = g.pnx * g.pnz
    b = g.nx * g.nz
    return (BYTES_PER_CELL_UPDATE * f * nshots, BYTES_PER_CELL_UPDATE * f * nshots,
            BYTES_PER_CELL_UPDATE * b * nshots)


def min_traffic_bytes(g, nshots, nr):
    """Minimum HBM bytes of one launch of each step kernel (FP32 fields):
    forward   per propagation-grid cell and shot: read p^n, p^{n-1}, write
              p^{n+1} (12 B); ekp once per cell (shared by the shots); CPML
              memory read + write (8 B per live node); frame strips p and X
              written (8 B per frame cell); one trace sample per receiver;
    adjoint   the same for mu, without frames (the residual is read instead);
    imaging   per physical cell and shot: read p^k, p^{k+1}, mu^{k+1}, write
              p^{k-1} (16 B); ekp once; frames read (8 B per frame cell); the
              FP64 image accumulator read + written once (16 B per cell)."""
    f = g.pnx * g.pnz
    b = g.nx * g.nz
    n1 = g.npml + 1
    cpml = (2 * n1 + 2 * g.npml) * g.pnz + (n1 + g.npml) * g.pnx
    frame = 6 * g.nx + 6 * g.nz
    fwd = nshots * (12 * f + 8 * cpml + 8 * frame + 4 * nr) + 4 * f
    adj = nshots * (12 * f + 8 * cpml + 4 * nr) + 4 * f
    rec = nshots * (16 * b + 8 * frame) + 4 * b + 16 * b
    return fwd, adj, rec


def cpu_baseline(g, vel, shots, nt_sample, nt_full, nsample=4):
    """FP32 oracle on the host cores (OpenMP over all of them: the reference's
    unfused loop structure for the forward operator, its exact transpose for
    the adjoint, the reconstruction and imaging loops parallel over cells) on
    `nsample` shots spread along the line, each with nt cut to nt_sample.
    Shots are independent, so the mean time per shot, scaled from nt_sample
    to nt_full steps, gives the shot-gradients/s of the full workload.
    Returns the rate, the per-phase Gcell-updates/s and the sample's wall
    time."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    from fwihip.workloads import Shot
    o = Oracle(np.float32)
    idx = sorted(set(int(round(i)) for i in np.linspace(0, len(shots) - 1, nsample)))
    t_fwd = t_msf = t_bwd = 0.0
    for i in idx:
        sh = shots[i]
        s = Shot(nt=nt_sample, isx=sh.isx, isz=sh.isz, wav=sh.wav[:nt_sample], irx=sh.irx, irz=sh.irz)
        obs = np.zeros((s.nr, nt_sample), np.float32)
        t0 = time.perf_counter()
        f = o.forward(g, vel, s, want_strips=True)
        t1 = time.perf_counter()
        _, res = o.misfit_l2(f["syn"], obs)
        t2 = time.perf_counter()
        o.backward(g, vel, s, res, f)
        t3 = time.perf_counter()
        t_fwd += t1 - t0
        t_msf += t2 - t1
        t_bwd += t3 - t2
    n = len(idx)
    per_shot = (t_fwd + t_msf + t_bwd) / n * nt_full / nt_sample
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    f_cells = g.pnx * g.pnz * nt_sample * n
    b_cells = (g.pnx * g.pnz + g.nx * g.nz) * nt_sample * n
    return dict(sg_per_s=1.0 / per_shot, shots=n, threads=threads, wall_s=t_fwd + t_msf + t_bwd,
                fwd_gcell=f_cells / t_fwd / 1e9, bwd_gcell=b_cells / t_bwd / 1e9,
                gcell=(f_cells + b_cells) / (t_fwd + t_msf + t_bwd) / 1e9)


def boundary_leg(g, start, shots, obs_packed, steps, warmup):
    """The same gradient through the drop-in boundary: grad_all_shots over one
    reference-layout plan per shot (fwi_plan.h), with the model and the
    observed data in host memory.  Each call uploads the model, runs the
    forward pass with each block of 100 steps' traces copied to the host
    while the next block propagates and the host misfit step (L2 in FP64,
    threaded over shots) run on each block as it arrives, hands the adjoint
    source over (uploaded block by block beside the adjoint steps), runs the
    adjoint pass and copies the gradient to plans[0].  Returns the rate and
    the mean wall time of each phase (the misfit of the plain L2 is inside
    the forward phase, the residual upload inside the adjoint phase)."""
    import ctypes
    import numpy as np
    from fwihip import plans as PC
    h = PC.lib()
    plans, o = [], 0
    for s in shots:
        n = s.nr * s.nt
        plans.append(PC.make_plan(h, g, s, obs_packed[o:o + n].reshape(s.nr, s.nt), start))
        o += n
    arr = (ctypes.POINTER(PC.FwiPlan) * len(plans))(*plans)
    ph = (ctypes.c_double * 5)()
    try:
        for _ in range(warmup):
            assert h.grad_all_shots(arr, len(plans), 0) == 0, h.grad4_last_error()
        acc = np.zeros(5)
        t0 = time.perf_counter()
        for _ in range(steps):
            assert h.grad_all_shots(arr, len(plans), 0) == 0, h.grad4_last_error()
            h.grad4_phase_ms(ph, 5)
            acc += np.array(ph[:])
        wall = time.perf_counter() - t0
    finally:
        for p in plans:
            h.fwi_plan_free(p)
    acc /= steps
    names = ("setup", "forward_trace_copy_and_l2", "host_misfit_rest", "residual_hand_over", "adjoint_residual_copy_and_gradient")
    return dict(value=len(shots) * steps / wall, unit="shot-gradients/s", ms_per_step=wall / steps * 1e3,
                steps=steps, warmup=warmup, phases_ms={k: round(float(v), 2) for k, v in zip(names, acc)})


def checkpoint_leg(g, vel, start, shots, obs_packed, nt, every, storage, steps, warmup, snaps=0):
    """The headline step with a checkpointed backward pass instead of the
    boundary frames: snapshots of every shot every `every` steps, or (snaps >
    0) the binomial schedule with `snaps` snapshot slots, stored in `storage`
    (per-shot max-abs scaling for 16 bit), states recomputed in the adjoint
    pass."""
    from fwihip import lib
    nr = max(s.nr for s in shots)
    ctx = lib.Context(g, max_shots=len(shots), max_nt=nt, max_nr=nr)
    try:
        if snaps:
            ctx.set_revolve(snaps, storage)
        else:
            ctx.set_checkpointing(every, storage)
        ctx.set_shots(shots)
        ctx.set_observed(obs_packed)

        def step():
            ctx.set_model(start)
            ctx.forward(want_traces=False)
            ctx.misfit_l2()
            ctx.adjoint()

        for _ in range(warmup):
            step()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        ctx.synchronize()
        wall = time.perf_counter() - t0
        out = dict(storage=storage, value=len(shots) * steps / wall, unit="shot-gradients/s",
                   ms_per_step=wall / steps * 1e3, device_gbytes=round(ctx.bytes / 1e9, 2))
        if snaps:
            out.update(method="binomial", snaps=snaps, recomputed_forward_steps=ctx.backward_plan()[2])
        else:
            out.update(method="uniform snapshots", every=every)
        return out
    finally:
        ctx.close()


def run_uneven(args):
    """The uneven batch: per-shot nt and receiver apertures in one batch; the
    misfit step on the host band-passes the synthetic traces (4-pole zero-phase
    Butterworth, 3-25 Hz, the observed data filtered once), takes the L2
    residual and applies the transposed filter, over a thread pool (the C
    calls release the interpreter lock)."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    from fwihip import lib
    from fwihip.workloads import config3, triangle_smooth

    g, vel, shots = config3(nshots=args.shots)
    start = triangle_smooth(vel, 25)
    nr = max(s.nr for s in shots)
    ntmax = max(s.nt for s in shots)
    h = lib.load()
    PF = ctypes.POINTER(ctypes.c_float)
    for name in ("sfbandpass", "sfbandpass_adjoint"):
        fn = getattr(h, name)
        fn.argtypes = [PF, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_bool, ctypes.c_int, ctypes.c_int,
                       ctypes.c_float, ctypes.c_float, PF]
        fn.restype = ctypes.c_int
    flo, fhi = 3.0, 25.0

    def band(x, adjoint=False):
        y = np.ascontiguousarray(x, np.float32).copy()
        fn = h.sfbandpass_adjoint if adjoint else h.sfbandpass
        rc = fn(y.ctypes.data_as(PF), y.shape[1], y.shape[0], g.dt, False, 4, 4, fhi, flo, y.ctypes.data_as(PF))
        assert rc == 0
        return y

    pool = ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))
    ctx = lib.Context(g, max_shots=len(shots), max_nt=ntmax, max_nr=nr)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    obs_f = list(pool.map(band, ctx.split(ctx.forward())))
    timing = {"host_s": 0.0}

    def one(a):
        syn, ob = a
        r = band(syn) - ob
        return 0.5 * float((r.astype(np.float64) ** 2).sum()), band(r, adjoint=True)

    def step():
        ctx.set_model(start)
        syn = ctx.split(ctx.forward())
        t0 = time.perf_counter()
        out = list(pool.map(one, zip(syn, obs_f)))
        timing["host_s"] += time.perf_counter() - t0
        ctx.set_residual(ctx.pack([r for _, r in out]))
        ctx.adjoint(want=False)
        return sum(j for j, _ in out)

    for _ in range(args.warmup):
        step()
    timing["host_s"] = 0.0
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        J = step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    cells = sum((g.pnx * g.pnz * 2 + g.nx * g.nz) * s.nt for s in shots)
    out = {
        "metric": "shot-gradients/s & Gcell-updates/s, uneven batch 1536x400, 96 shots, nt 2000-6000",
        "value": len(shots) * args.steps / wall, "unit": "shot-gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered/faulted model, Ricker shots of 5-15 Hz, observed data from the true model)",
        "config": {"workload": f"uneven FWI gradient {g.nx}x{g.nz} dt={g.dt * 1e3:g} ms, {len(shots)} shots, "
                               f"nt {min(s.nt for s in shots)}-{ntmax}, apertures "
                               f"{min(s.nr for s in shots)}-{nr} receivers, host 4-pole Butterworth {flo:g}-{fhi:g} Hz",
                   "nx": g.nx, "nz": g.nz, "nshots": len(shots)},
        "gcell_updates_per_s": cells * args.steps / wall / 1e9,
        "host_misfit_ms_per_step": timing["host_s"] / args.steps * 1e3,
        "misfit": J,
    }
    print(json.dumps(out))


def run_checkpoint(args):
    """Boundary-saving reconstruction against checkpointed recomputation with
    FP32, FP16 and BF16 snapshots on the batched workload: shot-gradients/s
    and device bytes of each, and the gradient error against the FP64 oracle
    on a 2-shot subset with nt <= 1500 (one fixed residual)."""
    import numpy as np
    from fwihip import lib
    from fwihip.workloads import config2, triangle_smooth, Shot

    g, vel, shots = config2(nshots=args.shots, nt=args.nt)
    start = triangle_smooth(vel, 25)
    nr = max(s.nr for s in shots)
    K = args.ckpt_every
    # (every, storage, binomial slots)
    modes = [(0, "fp32", 0), (K, "fp32", 0), (K, "fp16", 0), (K, "bf16", 0), (0, "fp32", 32), (0, "fp16", 32),
             (0, "fp16", 128)]
    ref = lib.Context(g, max_shots=len(shots), max_nt=args.nt, max_nr=nr)
    ref.set_model(vel)
    ref.set_shots(shots)
    obs = ref.forward()
    ref.close()
    rows = []
    for every, storage, snaps in modes:
        ctx = lib.Context(g, max_shots=len(shots), max_nt=args.nt, max_nr=nr)
        if snaps:
            ctx.set_revolve(snaps, storage)
        else:
            ctx.set_checkpointing(every, storage)
        ctx.set_model(start)
        ctx.set_shots(shots)
        ctx.set_observed(obs)

        def step():
            ctx.forward(want_traces=False)
            ctx.misfit_l2()
            ctx.adjoint(want=False)

        for _ in range(args.warmup):
            step()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        wall = time.perf_counter() - t0
        rows.append(dict(method=(f"binomial, {snaps} slots" if snaps else
                                 ("boundary-saving" if every == 0 else f"checkpoint every {every}")),
                         storage=storage, shot_gradients_per_s=len(shots) * args.steps / wall,
                         device_gbytes=ctx.bytes / 1e9, recomputed_forward_steps=ctx.backward_plan()[2]))
        ctx.close()
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    # gradient error against the FP64 oracle on a small subset
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    nt_acc = min(args.nt, 1500)
    sub = [Shot(nt=nt_acc, isx=s.isx, isz=s.isz, wav=s.wav[:nt_acc], irx=s.irx, irz=s.irz) for s in shots[:2]]
    # the L2 residual of the start model against the data of the true model
    cr = lib.Context(g, max_shots=2, max_nt=nt_acc, max_nr=nr)
    cr.set_shots(sub)
    cr.set_model(vel)
    ob = cr.split(cr.forward())
    cr.set_model(start)
    res = [(sy - o).astype(np.float32) for sy, o in zip(cr.split(cr.forward()), ob)]
    cr.close()
    o64 = Oracle(np.float64)
    t0 = time.perf_counter()
    g64 = np.zeros((g.nz, g.nx))
    for s, r in zip(sub, res):
        f = o64.forward(g, start.astype(np.float64), s, want_strips=True)
        g64 += o64.backward(g, start.astype(np.float64), s, r.astype(np.float64), f)["grad"]
    t_oracle = time.perf_counter() - t0
    for row, (every, storage, snaps) in zip(rows, modes):
        ctx = lib.Context(g, max_shots=2, max_nt=nt_acc, max_nr=nr)
        if snaps:
            ctx.set_revolve(snaps, storage)
        else:
            ctx.set_checkpointing(min(every, nt_acc) if every else 0, storage)
        ctx.set_model(start)
        ctx.set_shots(sub)
        ctx.forward(want_traces=False)
        ctx.set_residual(ctx.pack(res))
        grad, _ = ctx.adjoint()
        ctx.close()
        row["grad_rel_l2_vs_fp64"] = float(np.linalg.norm(grad - g64) / np.linalg.norm(g64))
    out = {"metric": "checkpointed vs boundary-saving backward pass at 1024x320, nt=%d, %d shots" % (args.nt, len(shots)),
           "unit": "shot-gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "rows": rows, "accuracy_sample": f"2 shots, nt={nt_acc}, the L2 residual of the start model, FP64 oracle "
                                            f"({t_oracle:.0f} s)",
           "dtype": "f32 arithmetic, snapshot storage per row"}
    print(json.dumps(out))


def run_inversion(args):
    """Full-waveform inversion on the batched workload through the C driver:
    fwi_invert (the L-BFGS loop of the reference's inversion driver, m = 5)
    over grad_all_shots with the chosen misfit on the host, gradient
    preconditioning, and optional checkpointed 16-bit snapshots.  Reports the
    misfit and the model's relative L2 error against the true model per
    iteration."""
    import ctypes
    import numpy as np
    from fwihip import plans as PC
    from fwihip import lib
    from fwihip.workloads import Geometry, layered_faulted, surface_shots, triangle_smooth

    g = Geometry(nx=1024, nz=320, npml=20, dx=10.0, dz=10.0, dt=0.8e-3)
    vel = layered_faulted(g, 2026)
    shots = surface_shots(g, args.shots, args.nt, seed=2027, rec_step=args.rec_step)
    start = triangle_smooth(vel, 25)
    jump = args.jump
    if (args.nt - 1) % jump:
        raise SystemExit(f"--jump {jump} must divide nt - 1 = {args.nt - 1}")
    nr = max(s.nr for s in shots)
    ctx = lib.Context(g, max_shots=len(shots), max_nt=args.nt, max_nr=nr)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    obs = ctx.split(ctx.forward())
    ctx.close()
    h = PC.lib()
    h.fwi_plan_set_grad.argtypes = [ctypes.POINTER(PC.FwiPlan), ctypes.c_float, ctypes.c_float] + [ctypes.c_int] * 6
    h.grad4_set_checkpointing.argtypes = [ctypes.c_int, ctypes.c_int]

    class LbfgsParam(ctypes.Structure):
        _fields_ = [("m", ctypes.c_int), ("epsilon", ctypes.c_float), ("max_iterations", ctypes.c_int),
                    ("max_linesearch", ctypes.c_int), ("ftol", ctypes.c_float), ("wolfe", ctypes.c_float)]

    h.fwi_lbfgs_default_param.argtypes = [ctypes.POINTER(LbfgsParam)]
    h.fwi_lbfgs_default_param.restype = None
    h.fwi_invert.argtypes = [ctypes.POINTER(ctypes.POINTER(PC.FwiPlan)), ctypes.c_int, ctypes.POINTER(LbfgsParam),
                             ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_char_p,
                             ctypes.POINTER(ctypes.c_int)]
    h.fwi_invert.restype = ctypes.c_int
    h.fwi_lbfgs_last_error.restype = ctypes.c_char_p
    plans = []
    for s, o in zip(shots, obs):
        p = PC.make_plan(h, g, s, np.ascontiguousarray(o[:, ::jump]), start, jump=jump)
        xyz = PC.view(p.contents.sou.contents.xyz_sr, (1 + s.nr, 2))
        xyz[0] = (s.isx * g.dx, (s.isz + 0.5) * g.dz)
        xyz[1:, 0] = s.irx * g.dx
        xyz[1:, 1] = (s.irz + 0.5) * g.dz
        plans.append(p)
    arr = (ctypes.POINTER(PC.FwiPlan) * len(plans))(*plans)
    kinds = {"l2": 0, "ncc": 1, "eot": 4}
    par = None
    if args.misfit == "eot":
        par = PC.EotPar(0, 0, 500, None, None, None, float((8 * g.dt * jump) ** 2), 1e-6, 0.0)
    assert h.grad4_set_misfit(kinds[args.misfit], ctypes.byref(par) if par is not None else None, 1.0, 0.0) == 0
    storage = {"fp32": 0, "fp16": 1, "bf16": 2}[args.storage]
    every = args.ckpt_every if args.storage != "fp32" or args.ckpt_every else 0
    assert h.grad4_set_checkpointing(every, storage) == 0
    # gradient scale: the first trial step (0.8 x the preconditioned gradient)
    # moves the model by about 40 m/s
    t0 = time.perf_counter()
    assert h.grad_all_shots(arr, len(plans), 0) == 0, h.grad4_last_error()
    t_grad = time.perf_counter() - t0
    gmax = float(np.abs(PC.view(plans[0].contents.grad.contents.body, (g.nz, g.nx))).max())
    assert h.fwi_plan_set_grad(plans[0], 0.0, 50.0 / gmax, 3, 3, 1, 2, 2, 1) == 0
    ib = np.ctypeslib.as_array(plans[0].contents.model.contents.ibtry, (g.nx,))
    ib[:] = 3                                    # sources and receivers sit in the top rows
    lp = LbfgsParam()
    h.fwi_lbfgs_default_param(ctypes.byref(lp))
    lp.m, lp.max_iterations = 5, args.iters
    log = os.path.join(args.log_dir or tempfile.gettempdir(), f"inversion_{args.misfit}_{args.storage}.jsonl")
    os.makedirs(os.path.dirname(log), exist_ok=True)
    if os.path.exists(log):
        os.remove(log)
    st = ctypes.c_int(-1)
    truth = np.ascontiguousarray(vel, np.float32)
    t0 = time.perf_counter()
    rc = h.fwi_invert(arr, len(plans), ctypes.byref(lp), 1, truth.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                      log.encode(), ctypes.byref(st))
    wall = time.perf_counter() - t0
    if rc:
        raise SystemExit(f"fwi_invert: {h.fwi_lbfgs_last_error().decode()}")
    recs = [json.loads(line) for line in open(log)]
    for p in plans:
        h.fwi_plan_free(p)
    out = {"metric": f"L-BFGS FWI, {args.misfit} misfit, 1024x320 nt={args.nt}, {len(shots)} shots x {nr} receivers",
           "unit": "per-iteration misfit and model rel-L2 error", "n_gpus": 1,
           "lbfgs": {"m": 5, "iterations": args.iters, "status": st.value}, "wall_s": wall,
           "gradient_s": t_grad, "snapshots": {"every": every, "storage": args.storage},
           "data_jump": jump, "iterations": recs,
           "data": "synthetic (seeded layered/faulted model; start = 25-cell triangle smoothing)"}
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); "
              "the shot batch would shard across GPUs with one gradient all-reduce, not measured here.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 5}))
        return
    if args.workload == "uneven":
        run_uneven(args)
        return
    if args.workload == "checkpoint":
        run_checkpoint(args)
        return
    if args.workload == "inversion":
        run_inversion(args)
        return
    import numpy as np
    from fwihip import lib
    from fwihip.workloads import config2, config4, triangle_smooth

    if args.config == 4:
        g, vel, shots = config4(nshots=args.shots, nt=args.nt)
    else:
        g, vel, shots = config2(nshots=args.shots, nt=args.nt)
    start = triangle_smooth(vel, 25)
    nr = max(s.nr for s in shots)
    launch = None
    if args.autotune:
        from fwihip import autotune
        t0 = time.perf_counter()
        launch, s_step, _ = autotune.tune(g, start, shots, verbose=True)
        launch["tuning_s"] = round(time.perf_counter() - t0, 1)
        launch["probe_us_per_step"] = round(s_step * 1e6, 1) if s_step else None
        ctx = autotune.make_context(g, shots, launch, args.nt)
    else:
        ctx = lib.Context(g, max_shots=len(shots), max_nt=args.nt, max_nr=nr, tile_z=args.tile_z,
                          shots_per_group=args.group)
        ctx.set_kernels(args.kernels)
        if args.rec_group:
            ctx.set_rec_group(args.rec_group)
        if args.bwd_chunk >= 0:
            ctx.set_bwd_chunk(args.bwd_chunk)
        if args.bwd_lanes >= 0:
            ctx.set_bwd_lanes(args.bwd_lanes)
        if args.bwd_gen_rows >= 0:
            ctx.set_bwd_general_rows(args.bwd_gen_rows)
        if args.bwd_overlap >= 0:
            ctx.set_bwd_overlap(args.bwd_overlap)
        if args.bwd_fused >= 0:
            ctx.set_bwd_fused(args.bwd_fused)
        if args.concurrent >= 0:
            ctx.set_concurrent(args.concurrent)
        if args.interior_wgs >= 0:
            ctx.set_interior_wgs(args.interior_wgs)
        if args.fwd_mixed >= 0:
            ctx.set_fwd_mixed(args.fwd_mixed)
        if args.bwd_mixed >= 0:
            ctx.set_bwd_mixed(args.bwd_mixed)
        if args.fwd_shape:
            ctx.set_pass_shape(0, *args.fwd_shape)
        if args.bwd_shape:
            ctx.set_pass_shape(1, *args.bwd_shape)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    obs = ctx.forward()
    ctx.set_observed(obs)  # observed data resident in HBM from here on

    def step():
        ctx.set_model(start)
        ctx.forward(want_traces=False)
        J = ctx.misfit_l2()
        grad, _ = ctx.adjoint()
        return J, grad

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        J, grad = step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    nshots = len(shots)
    cells_shot = (g.pnx * g.pnz * 2 + g.nx * g.nz) * args.nt
    sg_per_s = nshots * args.steps / wall
    gcell = cells_shot * nshots * args.steps / wall / 1e9

    # per-step durations (HIP events on the context stream), one extra step
    kern = {}
    kind = launch["kernels"] if launch else args.kernels
    names = (("k_fwd_tile", "k_adj_tile", "k_rec_tile") if kind == 0
             else ("k_fwd_step", "k_adj_step", "k_rec_step"))
    alg = list(algorithmic_bytes(g, nshots))
    mint = list(min_traffic_bytes(g, nshots, nr))
    fused = kind == 0 and args.bwd_fused != 0 and args.workload == "batched"
    if fused:
        # events around the whole forward and backward passes, run as in the
        # timed steps (graphs, second stream): time per step of each pass.  The
        # backward step credits both wavefields' cell-updates; its minimum
        # traffic skips the re-read of mu^{k+1} and the second ekp read
        ctx.set_timing(2)
        b = g.nx * g.nz
        names = ("k_fwd_tile", "k_bwd_tile")
        alg = [alg[0], alg[1] + alg[2]]
        mint = [mint[0], mint[1] + mint[2] - 4 * b * nshots - 4 * b]
    else:
        ctx.set_timing(1)  # an event pair around every launch
    step()
    for w, name in enumerate(names):
        ms, n = ctx.kernel_time(w)
        avg = ms / max(n, 1) * 1e-3
        kern[name] = dict(launches=n, avg_us=avg * 1e6, algorithmic_bytes=alg[w], min_traffic_bytes=mint[w],
                          achieved_GBps=alg[w] / avg / 1e9 if avg > 0 else None)
        if fused:
            kern[name]["per"] = "time step (all launches of the step, graph replay)"
    ctx.set_timing(0)
    dom = max(names, key=lambda k: kern[k]["avg_us"] * kern[k]["launches"])
    traffic, traffic_kind = None, None
    if args.config == 2 and args.shots == 64 and os.path.exists(PROFILE_TRAFFIC):
        try:
            prof = json.load(open(PROFILE_TRAFFIC))
            traffic, traffic_kind = prof.get(dom), prof.get("unit")
        except Exception:
            traffic = None
    roof = dict(bound="hbm", kernel=dom, achieved=kern[dom]["achieved_GBps"], peak=HBM_PEAK / 1e9, unit="GB/s",
                frac=kern[dom]["achieved_GBps"] / (HBM_PEAK / 1e9), traffic=traffic, traffic_kind=traffic_kind)

    single = None
    if not args.no_single:
        # the plain one-shot-per-launch GPU form, on the first 4 shots
        sub = shots[:4]
        c1 = lib.Context(g, max_shots=1, max_nt=args.nt, max_nr=nr, tile_z=args.tile_z)
        c1.set_kernels(args.kernels)
        c1.set_model(vel)
        t1 = 0.0
        for s in sub:
            c1.set_shots([s])
            o1 = c1.forward()
            c1.set_observed(o1)
            c1.set_model(start)
            c1.synchronize()
            ts = time.perf_counter()
            c1.forward(want_traces=False)
            c1.misfit_l2()
            c1.adjoint()
            c1.synchronize()
            t1 += time.perf_counter() - ts
            c1.set_model(vel)
        single = len(sub) / t1
        c1.close()

    boundary = ckpt = None
    if args.workload == "batched" and not args.autotune:
        obs_host = obs
        ctx.close()                        # its device memory is not needed any more
        xs, xw = min(args.steps, 5), min(args.warmup, 1)   # the extra legs: a few steps each
        if not args.no_boundary:
            boundary = boundary_leg(g, start, shots, obs_host, xs, xw)
            boundary["over_resident"] = boundary["value"] / sg_per_s
        if not args.no_ckpt:
            ckpt = [checkpoint_leg(g, vel, start, shots, obs_host, args.nt, 100, st, xs, xw)
                    for st in ("fp16", "bf16")]
            ckpt.append(checkpoint_leg(g, vel, start, shots, obs_host, args.nt, 0, "fp16", xs, xw, snaps=32))

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(g, start, shots, args.cpu_nt or args.nt, args.nt, args.cpu_shots)
        cpu = dict(value=c["sg_per_s"], unit="shot-gradients/s", cores=c["threads"], kind="port",
                   sample=f"{c['shots']} of the {nshots} shots, spread along the line, nt={args.cpu_nt or args.nt} "
                          f"each ({c['wall_s']:.1f} s, FP32 oracle, OpenMP on {c['threads']} threads): forward "
                          f"{c['fwd_gcell']:.3f}, backward (adjoint + reconstruction + imaging) "
                          f"{c['bwd_gcell']:.3f}, whole {c['gcell']:.3f} Gcell-updates/s; shots are "
                          "independent, so the mean time per shot gives the rate",
                   gcell_updates_per_s=c["gcell"], forward_gcell_updates_per_s=c["fwd_gcell"],
                   backward_gcell_updates_per_s=c["bwd_gcell"])

    out = {
        "metric": f"shot-gradients/s & Gcell-updates/s at {g.nx}x{g.nz}, nt={args.nt}; % HBM peak",
        "value": sg_per_s,
        "unit": "shot-gradients/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered/faulted model, Ricker shots, observed data from the true model)",
        "config": {"workload": f"batched FWI gradient {g.nx}x{g.nz} dx=dz={g.dx:g} m dt={g.dt * 1e3:g} ms "
                               f"nt={args.nt}, {nshots} shots x {nr} receivers, npml={g.npml}, free surface",
                   "nx": g.nx, "nz": g.nz, "nt": args.nt, "nshots": nshots, "nr": nr, "npml": g.npml,
                   "launch": launch or {"kernels": args.kernels, "rec_group": args.rec_group or 8,
                                        "bwd_chunk": "auto" if args.bwd_chunk < 0 else args.bwd_chunk,
                                        "bwd_lanes": 2 if args.bwd_lanes < 0 else args.bwd_lanes,
                                        "graph": 100,
                                        "fwd_shape": list(args.fwd_shape or (4, 6)),
                                        "bwd_shape": list(args.bwd_shape or (4, 4))}},
        "gcell_updates_per_s": gcell,
        "hbm_frac_whole_job": gcell * 1e9 * BYTES_PER_CELL_UPDATE / HBM_PEAK,
        "misfit": J,
        "roofline": roof,
        "kernels": kern,
        "method": "boundary-saving reconstruction (frames in HBM), L2 misfit on the device with the observed "
                  "data resident in HBM",
        "drop_in_boundary": boundary,
        "checkpointed": ckpt,
        "one_shot_per_launch_shot_gradients_per_s": single,
        "batched_over_single": (sg_per_s / single) if single else None,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (sg_per_s / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
