"""The host misfit step of grad4_1shot with each misfit of the pool (NCC,
XCORR1D, MF1D, entropic OT, dynamic warping, soft ranking, L2) and the trace conditioning (band-pass and
first-break mute, with their transposes on the residual): the gradient the
driver returns is the derivative of msf->val[1] with respect to the model,
checked by central differences of val[1] along a smooth direction."""
import ctypes

import numpy as np
import pytest

from oracle import Oracle
from fwihip import lib as L
import host_ctypes as HC
from fwihip import plans as PC
from test_gpu_grad4 import case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    lib = PC.lib()
    yield lib
    lib.grad4_set_misfit(0, None, 1.0, 0.0)             # L2 over every trace
    lib.grad4_set_bandpass(0, 0, 0, 0, 0.0, 0.0)
    lib.grad4_set_fbt_mute(0, 0.0, 0.0, 0)


def params(kind):
    if kind == 2:
        return HC.ParXcorr1d(0, 0, 3, b"middle")
    if kind == 3:
        return HC.Epsdf1(0, 0, 41, 0, 0, 300, b"middle", b"diff", 0.0, 0.3, 1.0)
    if kind == 4:
        return HC.EotPar(0, 0, 3000, None, None, None, 2e-4, 1e-9, 0.0)
    if kind == 5:
        return HC.DiwPar(11, 21, 1e-6, 5.0, 0.5, 100.0)
    return None


@pytest.mark.parametrize("kind,name", [(0, "L2"), (1, "NCC"), (2, "XCORR1D"), (3, "MF1D"), (4, "EOT"),
                                       (6, "SOFTRANK")])
@pytest.mark.parametrize("nt", [241, 240])
def test_driver_gradient_is_the_derivative_of_the_misfit(h, kind, name, nt):
    if kind >= 3 and nt == 240:
        pytest.skip("even nt only changes XCORR1D")
    g, vel, shot = case(nt=nt, seed=8)
    o = Oracle(np.float32)
    obs0 = o.forward(g, vel * 1.02, shot)["syn"]
    par = params(kind)
    if kind == 6:   # soft ranks of samples a fiftieth of the largest amplitude apart
        par = ctypes.c_float(float(np.abs(obs0).max()) / 50.0)
    # receivers beyond 400 m see no wave within nt steps: left out by offset, as
    # the reference's misfits do (xcorr_msf.c:69-78)
    assert h.grad4_set_misfit(kind, ctypes.byref(par) if par is not None else None, 0.0, 400.0) == 0
    assert h.grad4_set_bandpass(1, 0, 4, 4, 35.0, 3.0) == 0
    assert h.grad4_set_fbt_mute(1, 30.0, 0.02, 1) == 0
    p = PC.make_plan(h, g, shot, obs0, vel)
    P = p.contents
    fbt = PC.view(P.msf.contents.fbt, (shot.nr,))
    fbt[:] = np.abs(shot.irx - shot.isx) * g.dx / 2300.0
    xyz = PC.view(P.sou.contents.xyz_sr, (1 + shot.nr, 2))
    xyz[0] = (shot.isx * g.dx, (shot.isz + 0.5) * g.dz)
    xyz[1:, 0] = shot.irx * g.dx
    xyz[1:, 1] = (shot.irz + 0.5) * g.dz
    md = PC.view(P.model.contents.md[0], (g.nz, g.nx))
    val = PC.view(P.msf.contents.val, (2,))
    body = PC.view(P.grad.contents.body, (g.nz, g.nx))

    def run(v):
        md[:] = v
        h.grad4_1shot(p, 0)
        assert np.isfinite(val[1]), h.grad4_last_error()
        return float(val[1]), body.copy()

    J0, g0 = run(vel)
    assert J0 != 0.0   # entropic OT includes eps * sum P (log P - 1): its value may be negative
    rng = np.random.default_rng(3)
    zz, xx = np.meshgrid(np.arange(g.nz), np.arange(g.nx), indexing="ij")
    d = np.zeros((g.nz, g.nx))
    for _ in range(4):
        cz, cx = rng.uniform(10, g.nz - 5), rng.uniform(10, g.nx - 10)
        d += rng.standard_normal() * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 8.0 ** 2))
    d *= 30.0 / np.abs(d).max()                       # up to 30 m/s
    # central differences: O(e^2) error; MF1D's penalty bends faster (1.5 % at
    # e = 0.5, 0.08 % at 0.1), while EOT's small slope needs the larger step
    # against the FP32 rounding of val[1]
    e = 0.1 if kind == 3 else 0.5
    Jp, _ = run((vel + e * d).astype(np.float32))
    Jm, _ = run((vel - e * d).astype(np.float32))
    fd = (Jp - Jm) / (2 * e)
    an = float((g0.astype(np.float64) * d).sum())
    print(f"{name} nt={nt}: J={J0:.6e} fd={fd:.6e} adjoint={an:.6e}", flush=True)
    assert abs(fd - an) <= 2e-2 * abs(an), (fd, an)
    h.fwi_plan_free(p)


def test_driver_diw_chains_the_host_residual(h):
    """The warping misfit's residual is its derivative with the lag path held
    (pinned in syn on the CPU); the path itself moves with the model, so a
    difference quotient along the model does not pin it.  Here the driver must
    return lcc_diw_msf's value on the synthetics it computed, and the gradient
    of its residual: the FP64 oracle's adjoint of that residual."""
    g, vel, shot = case(nt=241, seed=8)
    o64 = Oracle(np.float64)
    obs0 = Oracle(np.float32).forward(g, vel * 1.02, shot)["syn"]
    par = params(5)
    assert h.grad4_set_misfit(5, ctypes.byref(par), 1.0, 0.0) == 0
    assert h.grad4_set_bandpass(0, 0, 0, 0, 0.0, 0.0) == 0     # no conditioning: the raw traces
    assert h.grad4_set_fbt_mute(0, 0.0, 0.0, 0) == 0
    p = PC.make_plan(h, g, shot, obs0, vel)
    P = p.contents
    PC.view(P.model.contents.md[0], (g.nz, g.nx))[:] = vel
    h.grad4_1shot(p, 0)
    val = PC.view(P.msf.contents.val, (2,)).copy()
    body = PC.view(P.grad.contents.body, (g.nz, g.nx)).copy()
    syn = PC.view(P.msf.contents.syn[0], (shot.nr, shot.nt)).copy()
    h.fwi_plan_free(p)
    hc = HC.lib()
    res = np.zeros_like(syn)
    mis = np.zeros(2, np.float32)
    rng = np.array([1.0, 0.0], np.float32)
    assert hc.lcc_diw_msf(HC.rows(obs0), HC.rows(syn), None, HC.fp(rng), par.eps, shot.nt, shot.nr, par.nl, par.nw,
                          par.alpha_win, par.sim_barrier, par.lag_barrier, HC.fp(mis), HC.rows(res)) == 0
    np.testing.assert_array_equal(val, mis)
    assert mis[1] > 0 and np.abs(res).max() > 0
    f = o64.forward(g, vel.astype(np.float64), shot, want_strips=True)
    g64 = o64.backward(g, vel.astype(np.float64), shot, res.astype(np.float64), f)["grad"]
    err = float(np.linalg.norm(body - g64) / np.linalg.norm(g64))
    print(f"DIW: misfit {mis}, gradient vs FP64 oracle adjoint of the host residual {err:.2e}")
    assert err < 1e-4


def test_xcorr1d_leading_refused(h):
    par = HC.ParXcorr1d(0, 0, 0, b"leading")
    assert h.grad4_set_misfit(2, ctypes.byref(par), 1.0, 0.0) == 1
    assert h.grad4_set_misfit(9, None, 1.0, 0.0) == 1
    assert h.grad4_set_misfit(5, ctypes.byref(HC.DiwPar(10, 21, 1e-6, 5.0, 0.5, 100.0)), 1.0, 0.0) == 1
    assert h.grad4_set_misfit(6, ctypes.byref(ctypes.c_float(0.0)), 1.0, 0.0) == 1
    assert h.grad4_set_misfit(0, None, 1.0, 0.0) == 0
