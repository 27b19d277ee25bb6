"""Fault injection through the C ABI on the GPU: every fault is refused on
the host before any kernel runs, with an error code and a message, and the
context (or the driver) works normally afterwards.
  * an HBM budget below what a context needs (creation refused, ENOMEM,
    whatever it had allocated released);
  * a model with a NaN, a non-positive velocity or a velocity above the
    stability limit (set_model refused, EINVAL; the previous model stays);
  * the same models through grad4_1shot (the reference's void driver): the
    misfit is poisoned with NaN and grad4_last_error names the cause."""
import math

import numpy as np
import pytest

from oracle import Oracle
from fwihip import lib as L
from fwihip import plans as PC
from fwihip.workloads import Geometry, Shot, ricker

pytestmark = pytest.mark.gpu

C0, C1, LZ_FS = 0.1129042e1, -0.4301412e-1, 6.0195791


@pytest.fixture(scope="module")
def lib():
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


def case():
    g = Geometry(nx=96, nz=48, npml=10, dx=10.0, dz=10.0, dt=1e-3)
    vel = np.full((g.nz, g.nx), 2200.0, np.float32)
    vel[20:, :] = 2800.0
    nt = 150
    irx = np.arange(0, g.nx, 2, dtype=np.int32)
    shot = Shot(nt=nt, isx=40, isz=2, wav=ricker(nt, g.dt, 20.0), irx=irx, irz=np.full(len(irx), 3, np.int32))
    return g, vel, shot


def vlimit(g):
    return 2.0 / (g.dt * math.sqrt(4 * (C0 - C1) ** 2 / g.dx ** 2 + LZ_FS / g.dz ** 2))


def bad_models(g, vel):
    out = []
    for name, cell, val in (("nan", 100, np.nan), ("zero", 5, 0.0), ("negative", 4000, -2000.0),
                            ("unstable", 777, 1.01 * vlimit(g))):
        v = vel.copy()
        v.flat[cell] = val
        out.append((name, cell, v))
    return out


def test_budget_below_the_context_is_refused(lib):
    g, vel, shot = case()
    print("step: context under a 16 KiB budget", flush=True)
    with pytest.raises(L.FwiError) as e:
        L.Context(g, max_shots=1, max_nt=shot.nt, max_nr=shot.nr, hbm_budget=16 << 10)
    assert e.value.code == L.ENOMEM
    print(f"refused: {e.value}")
    ctx = L.Context(g, max_shots=1, max_nt=shot.nt, max_nr=shot.nr)
    try:
        ctx.set_model(vel)
        ctx.set_shots([shot])
        syn = ctx.split(ctx.forward())[0]
        ref = Oracle(np.float32).forward(g, vel, shot)["syn"]
        assert np.linalg.norm(syn - ref) / np.linalg.norm(ref) < 1e-5
    finally:
        ctx.close()


def test_bad_models_refused_and_the_context_keeps_working(lib):
    g, vel, shot = case()
    ctx = L.Context(g, max_shots=1, max_nt=shot.nt, max_nr=shot.nr)
    try:
        ctx.set_model(vel)
        ctx.set_shots([shot])
        ref = ctx.split(ctx.forward())[0].copy()
        for name, cell, v in bad_models(g, vel):
            print(f"step: set_model with a {name} cell", flush=True)
            with pytest.raises(L.FwiError) as e:
                ctx.set_model(v)
            assert e.value.code == L.EINVAL and f"cell {cell}" in str(e.value), str(e.value)
            # the model of the last accepted call is still in place
            np.testing.assert_array_equal(ctx.split(ctx.forward())[0], ref)
    finally:
        ctx.close()


def test_bad_models_poison_the_driver_misfit(lib):
    g, vel, shot = case()
    obs0 = Oracle(np.float32).forward(g, vel * 1.01, shot)["syn"]
    h = PC.lib()
    assert h.grad4_set_misfit(0, None, 1.0, 0.0) == 0
    p = PC.make_plan(h, g, shot, obs0, vel)
    P = p.contents
    md = PC.view(P.model.contents.md[0], (g.nz, g.nx))
    val = PC.view(P.msf.contents.val, (2,))
    try:
        md[:] = vel
        h.grad4_1shot(p, 0)
        good = float(val[1])
        assert np.isfinite(good) and good > 0
        for name, cell, v in bad_models(g, vel):
            print(f"step: grad4_1shot with a {name} cell", flush=True)
            md[:] = v
            h.grad4_1shot(p, 0)
            assert np.isnan(val[0]) and np.isnan(val[1])
            msg = h.grad4_last_error().decode()
            assert f"cell {cell}" in msg, msg
        md[:] = vel
        h.grad4_1shot(p, 0)
        assert float(val[1]) == good
    finally:
        h.fwi_plan_free(p)
