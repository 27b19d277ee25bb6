"""The benchmark workloads on the GPU, with the default launch form, against
the oracle (needs an MI355X).

Default launch form: 2-D tile kernels, fused backward step, general strips
on a second stream beside the interior launches (interior workgroups capped
at 3 per CU), 8-shot imaging groups, 100-step graph blocks.

Tolerances (north star): seismograms rel-L2 <= 1e-5 and gradients rel-L2 <=
1e-4 against the FP32 oracle, both also checked against the FP64 oracle.

  * 1024x320, 64 shots x 1024 receivers (every column), nt cut to 400: the
    whole batch propagates; the traces of four shots (both x edges, and two
    in other 8-shot imaging groups) are checked, and the gradient of a
    residual that is nonzero for those four shots only (the other 60 shots
    back-propagate zero and image exactly zero).
  * the same workload at its full nt=4000: the default form equals the
    one-stream separate kernels and the 4-row strips bit for bit, the
    z-marching kernels up to the FP64 order of the image sum, and the batch
    equals the sum of its 64 single-shot runs.
  * 1536x400 shots with receiver apertures and uneven nt through the drop-in
    grad_all_shots with the host Butterworth band-pass.
  * 16-bit checkpointed snapshots at 1024x320 (2 shots) against FP64.
  * a source and receiver rows deep in the grid (general strips in the
    interior of the grid, the source tile outside the fused backward step):
    every launch form bit for bit, and the oracle.
"""
import ctypes

import numpy as np
import pytest

from oracle import Oracle
from fwihip.workloads import Geometry, Shot, ricker, config2, config3, triangle_smooth

pytestmark = pytest.mark.gpu

SYN_TOL = 1e-5
GRAD_TOL = 1e-4


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


@pytest.fixture(scope="module")
def o32():
    return Oracle(np.float32)


@pytest.fixture(scope="module")
def o64():
    return Oracle(np.float64)


def cut(shots, nt):
    return [Shot(nt=nt, isx=s.isx, isz=s.isz, wav=s.wav[:nt], irx=s.irx, irz=s.irz, f0=s.f0) for s in shots]


# ----------------------------------------------------------- 1024x320 x 64
PICK = (0, 20, 45, 63)   # x edges (isx 8 and 1016) and 8-shot groups 0, 2, 5, 7


def test_config2_batch_selected_shots_vs_oracle(lib, o32, o64):
    nt = 400
    g, vel, shots = config2(nshots=64, nt=4000)
    shots = cut(shots, nt)
    start = triangle_smooth(vel, 25)
    ctx = lib.Context(g, max_shots=64, max_nt=nt, max_nr=1024)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    obs = ctx.split(ctx.forward())
    ctx.set_model(start)
    syn = ctx.split(ctx.forward())
    res = [np.zeros_like(s) for s in syn]
    for j in PICK:
        res[j] = (syn[j] - obs[j]).astype(np.float32)
    ctx.set_residual(ctx.pack(res))
    grad, _ = ctx.adjoint()
    ctx.close()
    g32 = np.zeros_like(grad)
    g64 = np.zeros_like(grad)
    for j in PICK:
        s = shots[j]
        f32 = o32.forward(g, start, s, want_strips=True)
        f64 = o64.forward(g, start.astype(np.float64), s, want_strips=True)
        r32, r64 = rel(syn[j], f32["syn"]), rel(syn[j], f64["syn"])
        print(f"config2 shot {j} (isx {s.isx}): syn rel32 {r32:.2e} rel64 {r64:.2e}", flush=True)
        assert r32 < SYN_TOL and r64 < SYN_TOL
        g32 += o32.backward(g, start, s, res[j], f32)["grad"]
        g64 += o64.backward(g, start.astype(np.float64), s, res[j].astype(np.float64), f64)["grad"]
    print(f"config2 gradient of shots {PICK}: rel32 {rel(grad, g32):.2e} rel64 {rel(grad, g64):.2e}")
    assert rel(grad, g32) < GRAD_TOL and rel(grad, g64) < GRAD_TOL


def test_config2_full_nt_forms_and_single_shots(lib):
    g, vel, shots = config2(nshots=64, nt=4000)
    start = triangle_smooth(vel, 25)

    def gradient(ctx, obs_packed):
        ctx.set_model(start)
        tr = ctx.forward()
        ctx.set_observed(obs_packed)
        J = ctx.misfit_l2()
        return tr, J, ctx.adjoint()[0]

    ctx = lib.Context(g, max_shots=64, max_nt=4000, max_nr=1024)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    obs = ctx.forward()
    tr, J, grad = gradient(ctx, obs)
    print(f"config2 nt=4000 default form: misfit {J:.6e}", flush=True)
    assert np.abs(grad).max() > 0
    # one stream, separate adjoint and reconstruction kernels
    ctx.set_concurrent(False)
    ctx.set_bwd_fused(False)
    tr1, J1, g1 = gradient(ctx, obs)
    np.testing.assert_array_equal(tr1, tr)
    np.testing.assert_array_equal(g1, grad)
    assert J1 == J
    # the tile kernels with 4-row strips in both passes (the default form otherwise)
    ctx.set_concurrent(True)
    ctx.set_bwd_fused(True)
    ctx.set_pass_shape(0, 4, 3)
    ctx.set_pass_shape(1, 4, 3)
    tr4, J4, g4 = gradient(ctx, obs)
    np.testing.assert_array_equal(tr4, tr)
    np.testing.assert_array_equal(g4, grad)
    ctx.set_pass_shape(0, 8, 3)
    ctx.set_pass_shape(1, 8, 3)
    # z-marching kernels: the same traces, the image summed over shots in another order
    ctx.set_kernels(1)
    tr2, J2, g2 = gradient(ctx, obs)
    np.testing.assert_array_equal(tr2, tr)
    print(f"config2 nt=4000 march vs tile gradient rel {rel(g2, grad):.2e}", flush=True)
    assert rel(g2, grad) < 1e-12
    ctx.close()
    # the batch is the sum of its single-shot runs
    one = lib.Context(g, max_shots=1, max_nt=4000, max_nr=1024)
    gsum = np.zeros_like(grad)
    Jsum = 0.0
    o = 0
    for s in shots:
        n = s.nr * s.nt
        one.set_model(vel)
        one.set_shots([s])
        o1 = one.forward()
        np.testing.assert_array_equal(o1, obs[o:o + n])
        t1, j1, g1 = gradient(one, o1)
        np.testing.assert_array_equal(t1, tr[o:o + n])
        gsum += g1
        Jsum += j1
        o += n
    one.close()
    print(f"config2 nt=4000 batch vs 64 single shots: gradient rel {rel(grad, gsum):.2e}, "
          f"misfit {J:.9e} vs {Jsum:.9e}")
    assert rel(grad, gsum) < 1e-12
    assert abs(J - Jsum) < 1e-12 * J


# ----------------------------------------------------------- 1536x400 uneven, drop-in
def test_config3_uneven_apertures_bandpass_through_grad_all_shots(lib, o32, o64):
    from fwihip import plans as PC
    h = PC.lib()
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
        assert fn(y.ctypes.data_as(PF), y.shape[1], y.shape[0], g.dt, False, 4, 4, fhi, flo,
                  y.ctypes.data_as(PF)) == 0
        return y

    g, vel, all_shots = config3(nshots=96)
    # the smoothed model equals the true one in the thick top layer, which is
    # all that these short records see: 3 % slower as well
    start = (0.97 * triangle_smooth(vel, 25)).astype(np.float32)
    shots = []
    for j, nt in zip((0, 31, 62, 95), (281, 350, 317, 300)):   # uneven nt in one batch
        s = all_shots[j]
        shots.append(Shot(nt=nt, isx=s.isx, isz=s.isz, wav=ricker(nt, g.dt, s.f0), irx=s.irx, irz=s.irz, f0=s.f0))
    print("config3 subset: isx", [s.isx for s in shots], "receivers", [s.nr for s in shots],
          "nt", [s.nt for s in shots], flush=True)
    obs = [o32.forward(g, vel, s)["syn"] for s in shots]
    plans = [PC.make_plan(h, g, s, ob, start) for s, ob in zip(shots, obs)]
    arr = (ctypes.POINTER(PC.FwiPlan) * len(plans))(*plans)
    assert h.grad4_set_bandpass(1, 0, 4, 4, fhi, flo) == 0
    try:
        assert h.grad_all_shots(arr, len(plans), 0) == 0, h.grad4_last_error()
        body = PC.view(plans[0].contents.grad.contents.body, (g.nz, g.nx)).copy()
        syns = [PC.view(p.contents.msf.contents.syn[0], (s.nr, s.nt)).copy() for p, s in zip(plans, shots)]
        vals = [PC.view(p.contents.msf.contents.val, (2,)).copy() for p in plans]
    finally:
        h.grad4_set_bandpass(0, 0, 4, 4, fhi, flo)
        for p in plans:
            h.fwi_plan_free(p)
    g32 = np.zeros((g.nz, g.nx))
    g64 = np.zeros((g.nz, g.nx))
    for s, ob, sy, val in zip(shots, obs, syns, vals):
        f32 = o32.forward(g, start, s, want_strips=True)
        f64 = o64.forward(g, start.astype(np.float64), s, want_strips=True)
        assert rel(sy, f32["syn"]) < SYN_TOL and rel(sy, f64["syn"]) < SYN_TOL
        r = band(f32["syn"]) - band(ob)
        J = 0.5 * float(np.sum(r.astype(np.float64) ** 2))
        assert J > 0 and abs(val[0] - J) < 1e-4 * J, (val, J)
        res = band(r, adjoint=True)
        g32 += o32.backward(g, start, s, res, f32)["grad"]
        g64 += o64.backward(g, start.astype(np.float64), s, res.astype(np.float64), f64)["grad"]
    print(f"config3 subset gradient: rel32 {rel(body, g32):.2e} rel64 {rel(body, g64):.2e}")
    assert rel(body, g32) < GRAD_TOL and rel(body, g64) < GRAD_TOL


# ----------------------------------------------------------- 16-bit snapshots
def test_config5_snapshot_storage_vs_fp64(lib, o64):
    nt = 1000
    g, vel, shots = config2(nshots=64, nt=4000)
    shots = cut([shots[10], shots[50]], nt)
    start = triangle_smooth(vel, 25)
    ctx = lib.Context(g, max_shots=2, max_nt=nt, max_nr=1024)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    obs = ctx.split(ctx.forward())
    ctx.set_model(start)
    res = [(s - o).astype(np.float32) for s, o in zip(ctx.split(ctx.forward()), obs)]
    g64 = np.zeros((g.nz, g.nx))
    for s, r in zip(shots, res):
        f = o64.forward(g, start.astype(np.float64), s, want_strips=True)
        g64 += o64.backward(g, start.astype(np.float64), s, r.astype(np.float64), f)["grad"]
    errs = {}
    for every, storage, tol in ((0, "fp32", GRAD_TOL), (100, "fp32", GRAD_TOL), (100, "fp16", 1e-2),
                                (100, "bf16", 5e-2)):
        ctx.set_checkpointing(every, storage)
        ctx.forward(want_traces=False)
        ctx.set_residual(ctx.pack(res))
        grad, _ = ctx.adjoint()
        errs[(every, storage)] = rel(grad, g64)
        print(f"config5 snapshots every {every} in {storage}: gradient rel64 {errs[(every, storage)]:.2e}",
              flush=True)
        assert errs[(every, storage)] < tol
    ctx.close()
    assert errs[(100, "fp16")] < errs[(100, "bf16")]


# ----------------------------------------------------------- deep source and receivers
def deep_case():
    g = Geometry(nx=200, nz=64, npml=10, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(31)
    vel = np.clip(2300 + 250 * rng.standard_normal((g.nz, g.nx)), 1700, 3100).astype(np.float32)
    vel = triangle_smooth(vel, 3)
    nt = 350
    cols = np.arange(g.nx, dtype=np.int32)
    vsp = np.arange(8, 56, dtype=np.int32)                       # a vertical line of receivers
    shots = [
        Shot(nt=nt, isx=100, isz=30, wav=ricker(nt, g.dt, 18.0), irx=cols, irz=np.full(g.nx, 40, np.int32)),
        Shot(nt=nt - 40, isx=37, isz=2, wav=ricker(nt - 40, g.dt, 18.0), irx=cols, irz=np.full(g.nx, 2, np.int32)),
        Shot(nt=nt - 17, isx=160, isz=45, wav=ricker(nt - 17, g.dt, 15.0),
             irx=np.concatenate([np.full(len(vsp), 150, np.int32), cols[1::3]]),   # cells unique: 150 % 3 == 0
             irz=np.concatenate([vsp, np.full(len(cols[1::3]), 20, np.int32)])),
    ]
    return g, vel, shots


def test_deep_source_and_receivers_every_form_and_oracle(lib, o32, o64):
    g, vel, shots = deep_case()
    rng = np.random.default_rng(32)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = lib.Context(g, max_shots=3, max_nt=350, max_nr=max(s.nr for s in shots))
    ctx.set_model(vel)
    ctx.set_shots(shots)

    def run():
        tr = ctx.forward()
        ctx.set_residual(ctx.pack(res))
        return tr, ctx.adjoint()[1]

    ref = run()                                          # default form
    forms = {"one stream, fused": dict(concurrent=False, fused=True, rg=8),
             "one stream, separate": dict(concurrent=False, fused=False, rg=8),
             "two streams, separate": dict(concurrent=True, fused=False, rg=8),
             "two streams, fused, group 2": dict(concurrent=True, fused=True, rg=2)}
    try:
        for name, f in forms.items():
            ctx.set_concurrent(f["concurrent"])
            ctx.set_bwd_fused(f["fused"])
            ctx.set_rec_group(f["rg"])
            tr, gk = run()
            np.testing.assert_array_equal(tr, ref[0], err_msg=name)
            if f["rg"] == 8:
                np.testing.assert_array_equal(gk, ref[1], err_msg=name)
            else:                                        # another FP64 order of the sum over shots
                assert rel(gk, ref[1]) < 1e-12, name
        ctx.set_concurrent(True)
        ctx.set_bwd_fused(True)
        ctx.set_rec_group(8)
        ctx.set_kernels(1)
        tr, gk = run()
        np.testing.assert_array_equal(tr, ref[0])
        assert rel(gk, ref[1]) < 1e-12
        ctx.set_kernels(0)
        tr = ctx.split(ctx.forward())
        ctx.set_residual(ctx.pack(res))
        grad = ctx.adjoint()[0]
    finally:
        ctx.close()
    g32 = np.zeros_like(grad)
    g64 = np.zeros_like(grad)
    for s, t, r in zip(shots, tr, res):
        f32 = o32.forward(g, vel, s, want_strips=True)
        f64 = o64.forward(g, vel.astype(np.float64), s, want_strips=True)
        print(f"deep shot isz {s.isz}: syn rel32 {rel(t, f32['syn']):.2e} rel64 {rel(t, f64['syn']):.2e}")
        assert rel(t, f32["syn"]) < SYN_TOL and rel(t, f64["syn"]) < SYN_TOL
        g32 += o32.backward(g, vel, s, r, f32)["grad"]
        g64 += o64.backward(g, vel.astype(np.float64), s, r.astype(np.float64), f64)["grad"]
    print(f"deep case gradient rel32 {rel(grad, g32):.2e} rel64 {rel(grad, g64):.2e}")
    assert rel(grad, g32) < GRAD_TOL and rel(grad, g64) < GRAD_TOL
