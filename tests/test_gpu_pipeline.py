"""Time-major trace hand-over beside the propagation (fwi_hip_forward_tm) and
the time-major residual uploaded beside the adjoint steps
(fwi_hip_set_residual_tm), against the whole-array entry points on the same
inputs: bit-identical traces and gradients, in every launch form, for a batch
of alike shots (one strided copy per block) and an uneven one (per-shot
copies, shots ending before the last block)."""
import numpy as np
import pytest

from fwihip import lib as L
from fwihip.workloads import Geometry, Shot, ricker, triangle_smooth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


def batch(uneven):
    g = Geometry(nx=160, nz=70, npml=10, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(5)
    vel = triangle_smooth(np.clip(2300 + 250 * rng.standard_normal((g.nz, g.nx)), 1600, 3200), 3).astype(np.float32)
    shots = []
    nts = (337, 251, 305) if uneven else (330,) * 4
    for j, nt in enumerate(nts):
        irx = np.arange(j % 3, g.nx, 2 + (j % 2 if uneven else 0), dtype=np.int32)
        shots.append(Shot(nt=nt, isx=20 + 37 * j, isz=2 + 5 * j, wav=ricker(nt, g.dt, 18.0), irx=irx,
                          irz=np.full(len(irx), 2 + 4 * j, np.int32)))
    res = [np.random.default_rng(60 + j).standard_normal((s.nr, s.nt)).astype(np.float32) for j, s in
           enumerate(shots)]
    return g, vel, shots, res


def make(lib, g, vel, shots):
    ctx = lib.Context(g, max_shots=len(shots), max_nt=max(s.nt for s in shots), max_nr=max(s.nr for s in shots))
    ctx.set_model(vel)
    ctx.set_shots(shots)
    return ctx


@pytest.mark.parametrize("uneven", [False, True])
def test_trace_blocks_arrive_in_order_with_their_final_values(lib, uneven):
    g, vel, shots, _ = batch(uneven)
    ctx = make(lib, g, vel, shots)
    try:
        ref = [t.T.copy() for t in ctx.split(ctx.forward())]
        kmax = max(s.nt for s in shots)
        for graph in (100, 0):
            print(f"step: forward_tm uneven={uneven} graph={graph}", flush=True)
            ctx.set_graph(graph)
            seen = []
            out = np.empty(ctx.ntrace(), np.float32)

            def on_block(k0, k1):
                for r, t in zip(ref, ctx.split_tm(out)):
                    kk = min(k1, t.shape[0])
                    if k0 < kk and not np.array_equal(t[k0:kk], r[k0:kk]):
                        return 7
                seen.append((k0, k1))
                return 0

            ctx.forward_tm(on_block, out)
            assert seen[0][0] == 0 and seen[-1][1] == kmax
            assert all(a[1] == b[0] for a, b in zip(seen, seen[1:]))
            if graph:
                assert len(seen) == (kmax + graph - 1) // graph
            for r, t in zip(ref, ctx.split_tm(out)):
                np.testing.assert_array_equal(t, r)
        ctx.set_graph(100)
        calls = []

        def stop(k0, k1):
            calls.append(k0)
            return 5 if len(calls) == 2 else 0

        with pytest.raises(L.FwiError) as e:
            ctx.forward_tm(stop)
        assert e.value.code == 5 and len(calls) == 2
        np.testing.assert_array_equal(ctx.split(ctx.forward())[0].T, ref[0])   # the context still works
    finally:
        ctx.close()


FORMS = {
    "graph, fused": dict(),
    "per-step launches": dict(graph=0),
    "two shot chunks": dict(chunk=2),
    "binomial checkpoints": dict(revolve=3),
}


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("uneven", [False, True])
def test_time_major_residual_gives_the_same_gradient(lib, uneven, form):
    g, vel, shots, res = batch(uneven)
    opt = FORMS[form]
    ctx = make(lib, g, vel, shots)
    try:
        if "graph" in opt:
            ctx.set_graph(opt["graph"])
        if "chunk" in opt:
            ctx.set_bwd_chunk(opt["chunk"])
        if "revolve" in opt:
            ctx.set_revolve(opt["revolve"])
        print(f"step: residual forms uneven={uneven} {form}", flush=True)
        ctx.forward(want_traces=False)
        ctx.set_residual(ctx.pack(res))
        g_rm = ctx.adjoint()[0]
        ctx.forward_tm()
        ctx.set_residual_tm(ctx.pack([r.T for r in res]))
        g_tm = ctx.adjoint()[0]
        np.testing.assert_array_equal(g_tm, g_rm)
        assert np.abs(g_rm).max() > 0
    finally:
        ctx.close()
