"""fwi_invert: the L-BFGS loop (MBotOpt.c:65-221) over batched GPU gradients
(grad_all_shots) with grad_precond_new5, on a small two-shot problem: the
misfit falls monotonically, the model moves toward the true one, and the
JSON-lines log has one record per iteration."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import Oracle
from fwihip import lib as L
from fwihip.workloads import Geometry, Shot, ricker
from fwihip import plans as PC
from test_gradprec_lbfgs import LbfgsParam

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    lib = PC.lib()
    lib.fwi_invert.argtypes = [ctypes.POINTER(ctypes.POINTER(PC.FwiPlan)), ctypes.c_int, ctypes.POINTER(LbfgsParam),
                               ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_char_p,
                               ctypes.POINTER(ctypes.c_int)]
    lib.fwi_invert.restype = ctypes.c_int
    lib.fwi_lbfgs_last_error.restype = ctypes.c_char_p
    lib.fwi_plan_set_grad.argtypes = [ctypes.POINTER(PC.FwiPlan), ctypes.c_float, ctypes.c_float] + [ctypes.c_int] * 6
    lib.fwi_lbfgs_default_param.argtypes = [ctypes.POINTER(LbfgsParam)]
    lib.fwi_lbfgs_default_param.restype = None
    return lib


def test_fwi_invert_reduces_misfit_and_model_error(h, tmp_path):
    g = Geometry(nx=100, nz=50, npml=12, dx=10.0, dz=10.0, dt=1e-3)
    nt = 500
    true = np.full((g.nz, g.nx), 2000.0, np.float32)
    true[25:] = 2600.0
    zz, xx = np.meshgrid(np.arange(g.nz), np.arange(g.nx), indexing="ij")
    true += (250.0 * np.exp(-((zz - 18) ** 2 + (xx - 50) ** 2) / 60.0)).astype(np.float32)
    start = np.full_like(true, 2000.0)
    start[25:] = 2600.0
    irx = np.arange(2, g.nx - 2, 2, dtype=np.int32)
    o = Oracle(np.float32)
    plans = []
    for isx in (25, 75):
        shot = Shot(nt=nt, isx=isx, isz=2, wav=ricker(nt, g.dt, 15.0), irx=irx, irz=np.full(len(irx), 2, np.int32))
        obs = o.forward(g, true, shot)["syn"]
        p = PC.make_plan(h, g, shot, obs, start)
        plans.append(p)
    arr = (ctypes.POINTER(PC.FwiPlan) * 2)(*plans)
    # alpha_grad scales the gradient into model units, so that MBotOpt.c's
    # first trial step (0.8 d) moves the model by about 40 m/s
    assert h.grad_all_shots(arr, 2, 0) == 0, h.grad4_last_error()
    gmax = float(np.abs(PC.view(plans[0].contents.grad.contents.body, (g.nz, g.nx))).max())
    assert gmax > 0
    assert h.fwi_plan_set_grad(plans[0], 0.0, 50.0 / gmax, 2, 2, 1, 1, 1, 0) == 0
    ib = np.ctypeslib.as_array(plans[0].contents.model.contents.ibtry, (g.nx,))
    ib[:] = 2                                                    # no update in the two top rows
    par = LbfgsParam()
    h.fwi_lbfgs_default_param(ctypes.byref(par))
    par.m, par.max_iterations = 5, 4
    log = str(tmp_path / "inv.jsonl")
    st = ctypes.c_int(-1)
    rc = h.fwi_invert(arr, 2, ctypes.byref(par), 1, true.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                      log.encode(), ctypes.byref(st))
    assert rc == 0, h.fwi_lbfgs_last_error()
    recs = [json.loads(line) for line in open(log)]
    print(recs, flush=True)
    assert recs[0]["k"] == 0 and len(recs) >= 3
    mis = [r["misfit"] for r in recs]
    err = [r["model_rel_l2"] for r in recs]
    assert all(b < a for a, b in zip(mis[:-1], mis[1:]))
    assert mis[-1] < 0.7 * mis[0]
    assert err[-1] < err[0]
    md = PC.view(plans[0].contents.model.contents.md[0], (g.nz, g.nx))
    assert np.abs(md[:2] - start[:2]).max() == 0.0               # above the sea bottom: gradient zeroed
    for p in plans:
        h.fwi_plan_free(p)
