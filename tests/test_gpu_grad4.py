"""The drop-in boundary: grad4_1shot / grad_all_shots on reference-layout
plans (include/fwi_plan.h), against the oracle."""
import ctypes

import numpy as np
import pytest

from oracle import Oracle
from fwihip import lib as L
from fwihip.workloads import Geometry, Shot, ricker
from fwihip import plans as PC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return PC.lib()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def case(nt=241, seed=4):
    g = Geometry(nx=120, nz=60, npml=12, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(seed)
    vel = np.clip(2300 + 250 * rng.standard_normal((g.nz, g.nx)), 1600, 3200).astype(np.float32)
    irx = np.arange(1, g.nx, 3, dtype=np.int32)
    shot = Shot(nt=nt, isx=57, isz=2, wav=ricker(nt, g.dt, 18.0), irx=irx, irz=np.full(len(irx), 2, np.int32))
    return g, vel, shot


@pytest.mark.parametrize("jump", [1, 2])
def test_grad4_1shot_matches_oracle(h, jump):
    g, vel, shot = case()
    o = Oracle(np.float32)
    true = o.forward(g, vel * 1.03, shot)["syn"]
    obs0 = np.ascontiguousarray(true[:, ::jump])
    p = PC.make_plan(h, g, shot, obs0, vel, jump=jump)
    h.grad4_1shot(p, 0)
    P = p.contents
    body = PC.view(P.grad.contents.body, (g.nz, g.nx)).copy()
    syn = PC.view(P.msf.contents.syn[0], obs0.shape).copy()
    val = PC.view(P.msf.contents.val, (2,)).copy()
    # oracle: same driver, residual on the sampled steps
    f = o.forward(g, vel, shot, want_strips=True)
    res = np.zeros_like(f["syn"])
    res[:, ::jump] = f["syn"][:, ::jump] - obs0
    b = o.backward(g, vel, shot, res, f)
    J = 0.5 * np.sum((f["syn"][:, ::jump].astype(np.float64) - obs0) ** 2)
    assert rel(syn, f["syn"][:, ::jump]) < 1e-5
    assert rel(body, b["grad"]) < 1e-4
    assert abs(val[0] - J) < 1e-5 * J and val[0] == val[1]
    h.fwi_plan_free(p)


def test_grad_all_shots_sums_shots(h):
    g, vel, shot = case(nt=201, seed=6)
    o = Oracle(np.float32)
    shots = [shot, Shot(nt=201, isx=20, isz=3, wav=shot.wav, irx=shot.irx, irz=shot.irz)]
    plans = []
    gsum = np.zeros((g.nz, g.nx))
    for s in shots:
        obs0 = o.forward(g, vel * 0.98, s)["syn"]
        plans.append(PC.make_plan(h, g, s, obs0, vel))
        gsum += o.gradient(g, vel, s, obs0)[2]
    arr = (ctypes.POINTER(PC.FwiPlan) * 2)(*plans)
    assert h.grad_all_shots(arr, 2, 0) == 0, h.grad4_last_error()
    body = PC.view(plans[0].contents.grad.contents.body, (g.nz, g.nx))
    assert rel(body, gsum) < 1e-4
    for p in plans:
        h.fwi_plan_free(p)


def test_grad_all_shots_repeats_with_kept_buffers(h):
    """The driver keeps its page-locked trace buffer and the misfit step's
    working traces across calls: 3 shots, then 2 (smaller buffers in use),
    then 3 again give the same misfits and gradients bit for bit, for L2 and
    NCC, and the 3-shot gradient matches the oracle's sum over the shots."""
    g, vel, shot = case(nt=201, seed=7)
    o = Oracle(np.float32)
    shots = [Shot(nt=201, isx=x, isz=2 + j, wav=shot.wav, irx=shot.irx, irz=shot.irz)
             for j, x in enumerate((20, 57, 95))]
    obs = [o.forward(g, vel * 0.98, s)["syn"] for s in shots]
    plans = [PC.make_plan(h, g, s, ob, vel) for s, ob in zip(shots, obs)]

    def run(n):
        arr = (ctypes.POINTER(PC.FwiPlan) * n)(*plans[:n])
        assert h.grad_all_shots(arr, n, 0) == 0, h.grad4_last_error()
        body = PC.view(plans[0].contents.grad.contents.body, (g.nz, g.nx)).copy()
        vals = [PC.view(p.contents.msf.contents.val, (2,)).copy() for p in plans[:n]]
        return body, vals

    try:
        for kind in (0, 1):
            assert h.grad4_set_misfit(kind, None, 1.0, 0.0) == 0
            b3, v3 = run(3)
            b2, v2 = run(2)
            b3b, v3b = run(3)
            b2b, v2b = run(2)
            np.testing.assert_array_equal(b3b, b3)
            np.testing.assert_array_equal(b2b, b2)
            for a, b in zip(v3 + v2, v3b + v2b):
                np.testing.assert_array_equal(a, b)
            assert rel(b2, b3) > 1e-3                # the third shot counts
            if kind == 0:
                gsum = sum(o.gradient(g, vel, s, ob)[2] for s, ob in zip(shots, obs))
                assert rel(b3, gsum) < 1e-4
    finally:
        h.grad4_set_misfit(0, None, 1.0, 0.0)
        for p in plans:
            h.fwi_plan_free(p)


def test_grad4_spline_residual_mode(h):
    """Residual mode 1: the data-rate residual is upsampled by the natural
    cubic spline of trcinp (spline_interp.c:19-38) before back-propagation;
    the oracle gets the same residual from scipy's natural spline."""
    from scipy.interpolate import CubicSpline
    jump = 4
    g, vel, shot = case(nt=241, seed=9)
    o = Oracle(np.float32)
    obs0 = np.ascontiguousarray(o.forward(g, vel * 1.02, shot)["syn"][:, ::jump])
    p = PC.make_plan(h, g, shot, obs0, vel, jump=jump)
    assert h.grad4_set_residual_mode(1) == 0
    try:
        h.grad4_1shot(p, 0)
    finally:
        h.grad4_set_residual_mode(0)
    body = PC.view(p.contents.grad.contents.body, (g.nz, g.nx)).copy()
    f = o.forward(g, vel, shot, want_strips=True)
    r0 = (f["syn"][:, ::jump] - obs0).astype(np.float64)
    dt = np.float32(g.dt)
    x0 = (np.arange(r0.shape[1], dtype=np.float32) * np.float32(dt * jump)).astype(np.float64)
    x1 = (np.arange(shot.nt, dtype=np.float32) * dt).astype(np.float64)
    res = CubicSpline(x0, r0, axis=1, bc_type="natural")(x1)
    res[:, 0], res[:, -1] = r0[:, 0], r0[:, -1]
    b = o.backward(g, vel, shot, res.astype(np.float32), f)
    print(f"spline-residual grad rel {rel(body, b['grad']):.2e}")
    assert rel(body, b["grad"]) < 1e-4
    h.fwi_plan_free(p)
