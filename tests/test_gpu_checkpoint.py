"""Checkpointed backward pass (fwi_hip_set_checkpointing): snapshots of p
and the CPML memory every K steps in FP32, FP16 or BF16 (per-shot max-abs
scaling), segment recompute storing X, adjoint imaging from that store.

The forward traces are those of the boundary-saving forward bit for bit (the
snapshots only feed the backward pass).  The gradient with FP32 snapshots is
the exact discrete gradient recomputed in FP32 arithmetic: it meets the FP64
oracle within the gradient tolerance for any K and uneven nt.  16-bit
snapshots restart each segment from a rounded state: the gradient error
against FP64 is bounded by the storage precision (FP16 2^-11, BF16 2^-8 of the
snapshot's largest value), and is printed per dtype."""
import numpy as np
import pytest

from oracle import Oracle
from test_gpu_parity import small_case, rel, GRAD_TOL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from fwihip import lib as L
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


@pytest.fixture(scope="module")
def case():
    g, vel, shots = small_case(nshots=3, nt=240, seed=31)
    rng = np.random.default_rng(32)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    o64 = Oracle(np.float64)
    g64 = np.zeros((g.nz, g.nx))
    for s, r in zip(shots, res):
        f = o64.forward(g, vel, s, want_strips=True)
        g64 += o64.backward(g, vel, s, r.astype(np.float64), f)["grad"]
    return g, vel, shots, res, g64


def run(lib, g, vel, shots, res, every=0, storage="fp32", rec_group=2):
    ctx = lib.Context(g, max_shots=len(shots), max_nt=max(s.nt for s in shots), max_nr=100)
    ctx.set_rec_group(rec_group)
    ctx.set_checkpointing(every, storage)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    tr = ctx.forward()
    ctx.set_residual(ctx.pack(res))
    grad, _ = ctx.adjoint()
    ctx.close()
    return tr, grad


@pytest.mark.parametrize("every", [1, 7, 64, 500])
def test_fp32_snapshots_give_the_gradient(lib, case, every):
    g, vel, shots, res, g64 = case
    tr0, gb = run(lib, g, vel, shots, res)
    tr, gc = run(lib, g, vel, shots, res, every=every)
    np.testing.assert_array_equal(tr, tr0)
    print(f"K={every}: checkpointed grad rel64 {rel(gc, g64):.2e} (boundary-saving {rel(gb, g64):.2e}), "
          f"vs boundary-saving {rel(gc, gb):.2e}", flush=True)
    assert rel(gc, g64) < GRAD_TOL
    assert rel(gc, gb) < GRAD_TOL


@pytest.mark.parametrize("storage,bound", [("fp16", 5e-3), ("bf16", 4e-2)])
def test_16bit_snapshots_error_bounded_by_storage_precision(lib, case, storage, bound):
    g, vel, shots, res, g64 = case
    tr0, _ = run(lib, g, vel, shots, res)
    tr, gc = run(lib, g, vel, shots, res, every=20, storage=storage)
    np.testing.assert_array_equal(tr, tr0)
    e = rel(gc, g64)
    print(f"{storage} snapshots every 20 steps: grad rel64 {e:.2e}", flush=True)
    assert 0 < e < bound


def test_checkpointing_options(lib, case):
    g, vel, shots, res, _ = case
    ctx = lib.Context(g, max_shots=3, max_nt=240, max_nr=100)
    with pytest.raises(lib.FwiError):
        ctx.set_checkpointing(5, "fp8")
    with pytest.raises(lib.FwiError):
        ctx.set_checkpointing(-1)
    ctx.set_kernels(1)
    ctx.set_checkpointing(10)
    ctx.set_model(vel)
    ctx.set_shots(shots)
    with pytest.raises(lib.FwiError):
        ctx.forward()                       # the z-marching kernels have no checkpointed form
    ctx.close()
