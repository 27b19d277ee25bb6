"""Reflection waveform inversion, difference form (grad4_1shot_rfl /
grad_all_shots_rfl, include/fwi_plan.h): the reflected data R(p_per - p_inc)
and dJ/dm0 = the gradient through p_per (model m0 + m1, adjoint source r)
plus the gradient through p_inc (model m0, adjoint source -r), against the
FP64 oracle, and dJ/dm0 against a central finite difference of the misfit
the driver reports."""
import ctypes

import numpy as np
import pytest

from oracle import Oracle
from fwihip import lib as L
from fwihip.workloads import Geometry, Shot, ricker
from fwihip import plans as PC

pytestmark = pytest.mark.gpu

SYN_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.fixture(scope="module")
def h():
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return PC.lib()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def case(seed=8, nt=320):
    """A smooth background, a reflecting layer and a lens below it (two-way
    times of about 190 and 270 steps), two shots."""
    g = Geometry(nx=128, nz=64, npml=12, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(seed)
    z = np.arange(g.nz)[:, None] * np.ones((1, g.nx))
    m0 = (2000.0 + 12.0 * z + 20.0 * rng.standard_normal((1, g.nx)).cumsum(axis=1) / g.nx).astype(np.float32)
    m1 = np.zeros((g.nz, g.nx), np.float32)
    m1[20:24] = 400.0                                   # a reflecting layer
    m1[30:33, 40:90] = -300.0                           # a lens below it
    irx = np.arange(1, g.nx, 2, dtype=np.int32)
    shots = [Shot(nt=nt, isx=int(x), isz=2, wav=ricker(nt, g.dt, 20.0), irx=irx, irz=np.full(len(irx), 2, np.int32))
             for x in (40, 90)]
    # observed reflection data: the same difference form in a slightly different background
    o = Oracle(np.float64)
    mt = (m0 * 1.02).astype(np.float32)
    obs = [o.forward(g, mt + m1, s)["syn"] - o.forward(g, mt, s)["syn"] for s in shots]
    return g, m0, m1, shots, obs


def run(h, g, m0, m1, shots, obs, want_m1=False):
    plans = [PC.make_rwi_plan(h, g, s, ob.astype(np.float32), m0, m1) for s, ob in zip(shots, obs)]
    arr = (ctypes.POINTER(PC.RwiPlan) * len(plans))(*plans)
    gm1 = np.zeros((g.nz, g.nx), np.float32)
    rc = h.grad_all_shots_rfl(arr, len(plans), 0, gm1.ctypes.data_as(PC.PF) if want_m1 else None)
    assert rc == 0, h.grad4_last_error()
    body = PC.view(plans[0].contents.grad.contents.body, (g.nz, g.nx)).copy()
    syn = [PC.view(p.contents.msf.contents.syn[0], (s.nr, s.nt)).copy() for p, s in zip(plans, shots)]
    J = sum(float(PC.view(p.contents.msf.contents.val, (2,))[1]) for p in plans)
    for p in plans:
        h.rwi_plan_free(p)
    return body, gm1, syn, J


def test_rwi_gradient_matches_oracle(h):
    g, m0, m1, shots, obs = case()
    body, gm1, syn, J = run(h, g, m0, m1, shots, obs, want_m1=True)
    o = Oracle(np.float64)
    g0 = np.zeros((g.nz, g.nx))
    g1 = np.zeros((g.nz, g.nx))
    J64 = 0.0
    for s, ob, sy in zip(shots, obs, syn):
        fi = o.forward(g, m0, s, want_strips=True)
        fp = o.forward(g, m0 + m1, s, want_strips=True)
        refl = fp["syn"] - fi["syn"]
        r = refl - ob.astype(np.float32)
        J64 += 0.5 * np.sum(r ** 2)
        g1 += o.backward(g, m0 + m1, s, r, fp)["grad"]
        g0 += o.backward(g, m0, s, -r, fi)["grad"]
        print(f"reflected data rel64 {rel(sy, refl):.2e}", flush=True)
        assert rel(sy, refl) < SYN_TOL * 10   # a difference of two fields: FP32 cancellation
    print(f"dJ/dm0 rel64 {rel(body, g0 + g1):.2e}, dJ/dm1 rel64 {rel(gm1, g1):.2e}, J {J:.6e} vs {J64:.6e}",
          flush=True)
    assert rel(body, g0 + g1) < GRAD_TOL
    assert rel(gm1, g1) < GRAD_TOL
    assert abs(J - J64) < 1e-4 * J64


def test_rwi_gradient_is_the_derivative_of_the_misfit(h):
    """<dJ/dm0, d> against (J(m0 + e d) - J(m0 - e d)) / 2e for a smooth d."""
    g, m0, m1, shots, obs = case(seed=9)
    body, _, _, _ = run(h, g, m0, m1, shots, obs)
    zz, xx = np.meshgrid(np.arange(g.nz), np.arange(g.nx), indexing="ij")
    d = np.exp(-(((zz - 25) / 8.0) ** 2 + ((xx - 64) / 20.0) ** 2)).astype(np.float32)
    e = 5.0
    Jp = run(h, g, (m0 + e * d).astype(np.float32), m1, shots, obs)[3]
    Jm = run(h, g, (m0 - e * d).astype(np.float32), m1, shots, obs)[3]
    fd = (Jp - Jm) / (2 * e)
    an = float(np.sum(body.astype(np.float64) * d))
    print(f"FD {fd:.6e} analytic {an:.6e} rel {abs(fd - an) / abs(an):.2e}", flush=True)
    assert abs(fd - an) < 5e-3 * abs(an)


def test_rwi_one_shot_entry(h):
    """grad4_1shot_rfl on one plan equals the batched form on that plan."""
    g, m0, m1, shots, obs = case(seed=10)
    p = PC.make_rwi_plan(h, g, shots[0], obs[0].astype(np.float32), m0, m1)
    h.grad4_1shot_rfl(p, 0)
    one = PC.view(p.contents.grad.contents.body, (g.nz, g.nx)).copy()
    h.rwi_plan_free(p)
    body = run(h, g, m0, m1, shots[:1], obs[:1])[0]
    np.testing.assert_array_equal(one, body)
