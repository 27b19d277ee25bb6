"""The large-grid workload (4096x1024, dx=dz=5 m, dt=0.4 ms, 16 shots x 4096
receivers, nt=10000): one shot on the full grid against the FP32 oracle at a
truncated nt, and the full-size batch through properties that hold at any
size.

At full size the two-stream fused backward pass (the default) and the
one-stream separate kernels give the same traces and gradient bit for bit.
The pass is linear in the wavelet and the gradient bilinear in wavelet and
residual, so doubling the wavelet and the observed data doubles the traces
and quadruples the gradient up to FP32 roundoff: not bit for bit, because
gradual underflow rounds the denormal values at the numerical wavefront
differently for x and 2x (measured from step 9 on), and each such difference
can flip later roundings, after which the two runs carry independent FP32
roundoff (measured with the L2 residual: traces 9.0e-6, misfit 1.5e-7,
gradient 6.7e-6 relative)."""
import numpy as np
import pytest

from oracle import Oracle
from fwihip.workloads import Shot, config4

pytestmark = pytest.mark.gpu

SYN_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.fixture(scope="module")
def lib():
    from fwihip import lib as L
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_large_grid_one_shot_matches_oracle(lib):
    nt = 300
    g, vel, shots = config4(nshots=16, nt=nt)
    s = shots[0]
    o = Oracle(np.float32)
    obs = o.forward(g, (vel * 1.02).astype(np.float32), s)["syn"]
    J, syn, grad = o.gradient(g, vel, s, obs)[:3]
    ctx = lib.Context(g, max_shots=1, max_nt=nt, max_nr=s.nr)
    try:
        ctx.set_model(vel)
        ctx.set_shots([s])
        tr = ctx.split(ctx.forward())[0]
        ctx.set_observed(obs.reshape(-1))
        Jg = ctx.misfit_l2()
        gg, _ = ctx.adjoint()
    finally:
        ctx.close()
    print(f"large grid, 1 shot, nt={nt}: syn rel32 {rel(tr, syn):.2e}, grad rel32 {rel(gg, grad):.2e}, "
          f"J {Jg:.6e} vs {J:.6e}", flush=True)
    assert rel(tr, syn) < SYN_TOL
    assert rel(gg, grad) < GRAD_TOL
    assert abs(Jg - J) < 1e-4 * abs(J)


def test_large_batch_full_size(lib):
    g, vel, shots = config4()
    nt = shots[0].nt
    start = (vel * np.float32(0.98)).astype(np.float32)
    ctx = lib.Context(g, max_shots=len(shots), max_nt=nt, max_nr=max(s.nr for s in shots))
    try:
        print(f"full size: {len(shots)} shots x {shots[0].nr} receivers, nt={nt}", flush=True)
        ctx.set_model(vel)
        ctx.set_shots(shots)
        obs = ctx.forward()

        def gradient(concurrent, fused, scale):
            ctx.set_concurrent(concurrent)
            ctx.set_bwd_fused(fused)
            ctx.set_model(start)
            if scale != 1:
                ctx.set_shots([Shot(nt=s.nt, isx=s.isx, isz=s.isz, wav=(s.wav * np.float32(scale)).astype(np.float32),
                                    irx=s.irx, irz=s.irz) for s in shots])
            else:
                ctx.set_shots(shots)
            ctx.set_observed(obs * np.float32(scale))
            tr = ctx.forward()
            J = ctx.misfit_l2()
            gr, _ = ctx.adjoint()
            return tr, J, gr

        ta, Ja, ga = gradient(True, True, 1)
        print(f"{ctx.bytes / 2**30:.1f} GiB on the device, J {Ja:.6e}, |grad| {np.abs(ga).max():.3e}", flush=True)
        tb, Jb, gb = gradient(False, False, 1)
        assert np.array_equal(ta, tb) and Ja == Jb
        assert np.array_equal(ga, gb)
        tc, Jc, gc = gradient(True, True, 2)
    finally:
        ctx.close()
    et = rel(tc, 2.0 * ta.astype(np.float64))
    eg = rel(gc, 4.0 * ga)
    print(f"wavelet and data x2: traces x2 rel {et:.2e}, misfit x4 rel {abs(Jc - 4 * Ja) / (4 * Ja):.2e}, "
          f"gradient x4 rel {eg:.2e}", flush=True)
    assert Ja > 0 and np.abs(ga).max() > 0
    assert et < 5e-5
    assert abs(Jc - 4 * Ja) < 1e-4 * 4 * Ja
    assert eg < 5e-3
