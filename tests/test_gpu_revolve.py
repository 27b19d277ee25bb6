"""Binomial (Revolve-style) checkpointing on the GPU, the HBM-budget choice
between it and boundary saving, and the differentiation checks of the
large-grid workload (needs an MI355X).

  * small uneven batch: the binomial gradient with 1, 2, 3 and 5 slots equals
    the boundary-saving gradient to FP32 roundoff and the FP64 oracle within
    1e-4; the forward steps it recomputes are the schedule's;
  * a context whose HBM budget is too small for the boundary frames takes
    binomial checkpoints on its own and gives the same gradient;
  * 4096x1024, 16 shots (the large-grid workload, nt cut to 800): binomial
    checkpoints against boundary saving (<= 1e-5), a Taylor ladder of the
    misfit along a seeded smooth direction at steps 1e-1 .. 1e-6 (the
    first-order remainder falls as h^2 until it meets the FP32 rounding floor
    of the misfit, about 1e-6 J, below h = 1e-2; the central difference
    meets the gradient within 1e-3 at its best step), and the FP32/FP64
    oracle on a copy of the model downsampled four times.
"""
import numpy as np
import pytest

from oracle import Oracle
from fwihip.workloads import Geometry, Shot, ricker, config4, triangle_smooth

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    from fwihip import lib as L
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


def small_case():
    g = Geometry(nx=150, nz=72, npml=10, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(41)
    vel = triangle_smooth(np.clip(2200 + 300 * rng.standard_normal((g.nz, g.nx)), 1500, 3200), 3)
    shots = []
    for j, nt in enumerate((233, 190, 211)):               # uneven nt: the schedule runs over the longest
        irx = np.arange(j % 2, g.nx, 2, dtype=np.int32)
        shots.append(Shot(nt=nt, isx=25 + 45 * j, isz=2 + 9 * j, wav=ricker(nt, g.dt, 20.0), irx=irx,
                          irz=np.full(len(irx), 2 + 3 * j, np.int32)))
    res = [np.random.default_rng(42 + j).standard_normal((s.nr, s.nt)).astype(np.float32)
           for j, s in enumerate(shots)]
    return g, vel, shots, res


def gradient(ctx, res):
    ctx.forward(want_traces=False)
    ctx.set_residual(ctx.pack(res))
    return ctx.adjoint()[0]


def test_binomial_gradient_small_batch(lib):
    g, vel, shots, res = small_case()
    ctx = lib.Context(g, max_shots=3, max_nt=233, max_nr=75)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    ref = gradient(ctx, res)
    assert ctx.backward_plan()[0] == 0
    o64 = Oracle(np.float64)
    g64 = np.zeros_like(ref)
    for s, r in zip(shots, res):
        f = o64.forward(g, vel.astype(np.float64), s, want_strips=True)
        g64 += o64.backward(g, vel.astype(np.float64), s, r.astype(np.float64), f)["grad"]
    print(f"boundary saving vs FP64 {rel(ref, g64):.2e}")
    try:
        for snaps in (1, 2, 3, 5):
            ctx.set_revolve(snaps)
            gr = gradient(ctx, res)
            method, sn, rec = ctx.backward_plan()
            _, fsteps = lib.revolve_schedule(233, snaps)
            print(f"{snaps} slots: vs boundary saving {rel(gr, ref):.2e}, vs FP64 {rel(gr, g64):.2e}, "
                  f"recomputed {rec} forward steps")
            assert (method, sn) == (2, snaps)
            assert rec == fsteps - 232            # the first sweep's 232 advance steps ran in the forward pass
            assert rel(gr, ref) < 1e-5 and rel(gr, g64) < 1e-4
            assert np.array_equal(gradient(ctx, res), gr)   # the schedule replays the same way
        for storage, tol in (("fp16", 1e-2), ("bf16", 5e-2)):
            ctx.set_revolve(3, storage)
            gr = gradient(ctx, res)
            print(f"3 slots in {storage}: vs FP64 {rel(gr, g64):.2e}")
            assert rel(gr, g64) < tol
        ctx.set_revolve(0)
        np.testing.assert_array_equal(gradient(ctx, res), ref)
        assert ctx.backward_plan()[0] == 0
    finally:
        ctx.close()


def test_budget_too_small_for_frames_takes_checkpoints(lib):
    g, vel, shots, res = small_case()
    method, snaps, need = lib.plan_memory(g, 3, 233, 75, 0)
    assert method == 0
    budget = need - (1 << 20)                               # 1 MiB short of the boundary frames
    m2, sn2, b2 = lib.plan_memory(g, 3, 233, 75, budget)
    assert m2 == 2 and sn2 >= 3 and b2 <= budget
    free = lib.Context(g, max_shots=3, max_nt=233, max_nr=75)
    free.set_model(vel)
    free.set_shots(shots)
    ref = gradient(free, res)
    free.close()
    ctx = lib.Context(g, max_shots=3, max_nt=233, max_nr=75, hbm_budget=budget)
    try:
        ctx.set_revolve(-1)
        ctx.set_model(vel)
        ctx.set_shots(shots)
        gr = gradient(ctx, res)
        method, sn, rec = ctx.backward_plan()
        print(f"budget {budget} bytes: method {method}, {sn} slots, {ctx.bytes} bytes held, "
              f"gradient vs boundary saving {rel(gr, ref):.2e}")
        # the planner counts every shot at max_nt: the context may fit a slot more
        assert method == 2 and sn >= sn2 and ctx.bytes <= budget
        assert rel(gr, ref) < 1e-5
    finally:
        ctx.close()
    roomy = lib.Context(g, max_shots=3, max_nt=233, max_nr=75, hbm_budget=need + (1 << 20))
    try:
        roomy.set_revolve(-1)
        roomy.set_model(vel)
        roomy.set_shots(shots)
        np.testing.assert_array_equal(gradient(roomy, res), ref)   # the frames fit: boundary saving
        assert roomy.backward_plan()[0] == 0
    finally:
        roomy.close()


# --------------------------------------------------------------- 4096 x 1024
NT4 = 800


@pytest.fixture(scope="module")
def large():
    g, vel, shots = config4(nshots=16, nt=10000)
    shots = [Shot(nt=NT4, isx=s.isx, isz=s.isz, wav=s.wav[:NT4], irx=s.irx, irz=s.irz, f0=s.f0) for s in shots]
    # 3 % slow as well: the smoothed model equals the true one in the thick
    # top layer, which is most of what these short records see
    start = (0.97 * triangle_smooth(vel, 25)).astype(np.float32)
    return g, vel, start, shots


def test_config4_binomial_vs_boundary_and_taylor_ladder(lib, large):
    g, vel, start, shots = large
    ctx = lib.Context(g, max_shots=16, max_nt=NT4, max_nr=4096)
    try:
        ctx.set_model(vel)
        ctx.set_shots(shots)
        obs = ctx.forward()
        ctx.set_observed(obs)

        def misfit(model):
            ctx.set_model(model)
            ctx.forward(want_traces=False)
            return ctx.misfit_l2()

        J0 = misfit(start)
        grad = ctx.adjoint()[0]
        ctx.set_revolve(16)
        J1 = misfit(start)
        gr = ctx.adjoint()[0]
        method, sn, rec = ctx.backward_plan()
        print(f"4096x1024 x 16 shots, nt {NT4}: binomial ({sn} slots, {rec} forward steps recomputed) vs boundary "
              f"saving: gradient {rel(gr, grad):.2e}, misfit {J1} vs {J0}", flush=True)
        assert method == 2 and J1 == J0 and rel(gr, grad) < 1e-5
        ctx.set_revolve(0)
        # Taylor ladder along a seeded smooth direction (10 % of the model at
        # h = 1: h = 1e-6 moves the model by about the FP32 resolution of a velocity)
        rng = np.random.default_rng(2028)
        dm = triangle_smooth(rng.standard_normal((g.nz, g.nx)), 25)
        dm = (dm / np.abs(dm).max() * 0.1 * start).astype(np.float32)
        G = float(np.sum(grad * dm))
        rows = []
        for h in (1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
            jp = misfit((start + h * dm).astype(np.float32))
            jm = misfit((start - h * dm).astype(np.float32))
            rows.append((h, abs(jp - J0 - h * G), abs((jp - jm) / (2 * h) - G) / abs(G)))
            print(f"  h {h:.0e}: first-order remainder {rows[-1][1]:.3e}, central difference rel err "
                  f"{rows[-1][2]:.2e}", flush=True)
        print(f"  misfit {J0:.6e}, directional derivative {G:.6e}")
        r1 = [r[1] for r in rows]
        assert 25 < r1[0] / r1[1] < 400            # the remainder falls as h^2 ...
        assert max(r1[2:]) < 1e-5 * J0             # ... to the FP32 rounding floor of the misfit
        assert min(r[2] for r in rows) < 1e-3      # the gradient is the derivative of the misfit
    finally:
        ctx.close()


def test_config4_downsampled_copy_vs_oracle(lib, large):
    g4, vel, start, shots = large
    q = 4
    g = Geometry(nx=g4.nx // q, nz=g4.nz // q, npml=20, dx=g4.dx * q, dz=g4.dz * q, dt=g4.dt * q)
    v = start[q // 2::q, q // 2::q].copy()
    nt = 500
    picks = [shots[2], shots[11]]
    sub = []
    for s in picks:
        irx = np.unique(s.irx // q).astype(np.int32)
        sub.append(Shot(nt=nt, isx=s.isx // q, isz=max(1, s.isz // q), wav=ricker(nt, g.dt, s.f0), irx=irx,
                        irz=np.full(len(irx), max(1, s.irz[0] // q), np.int32)))
    rng = np.random.default_rng(7)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in sub]
    ctx = lib.Context(g, max_shots=2, max_nt=nt, max_nr=max(s.nr for s in sub))
    ctx.set_model(v)
    ctx.set_shots(sub)
    syn = ctx.split(ctx.forward())
    ctx.set_residual(ctx.pack(res))
    grad = ctx.adjoint()[0]
    ctx.close()
    o32, o64 = Oracle(np.float32), Oracle(np.float64)
    g32 = np.zeros_like(grad)
    g64 = np.zeros_like(grad)
    for s, r, t in zip(sub, res, syn):
        f32 = o32.forward(g, v, s, want_strips=True)
        f64 = o64.forward(g, v.astype(np.float64), s, want_strips=True)
        assert rel(t, f32["syn"]) < 1e-5 and rel(t, f64["syn"]) < 1e-5
        g32 += o32.backward(g, v, s, r, f32)["grad"]
        g64 += o64.backward(g, v.astype(np.float64), s, r.astype(np.float64), f64)["grad"]
    print(f"downsampled {g.nx}x{g.nz}: gradient rel32 {rel(grad, g32):.2e} rel64 {rel(grad, g64):.2e}")
    assert rel(grad, g32) < 1e-4 and rel(grad, g64) < 1e-4
