"""The launch-parameter tuner (fwihip/autotune.py): it returns one of its
candidates, caches the choice per shape, and the chosen setting gives the
gradient of the default setting (every candidate computes the same
gradient: the kernel forms agree to the FP64 order of the image sums)."""
import json

import numpy as np
import pytest

from fwihip import autotune
from test_gpu_parity import small_case, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from fwihip import lib as L
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


def gradient(ctx, vel, shots, res):
    ctx.set_model(vel)
    ctx.set_shots(shots)
    ctx.forward(want_traces=False)
    ctx.set_residual(ctx.pack(res))
    return ctx.adjoint()[0]


def test_autotune_picks_a_candidate_and_caches_it(lib, tmp_path):
    g, vel, shots = small_case(nshots=3, nt=200, seed=21)
    cache = str(tmp_path / "tune.json")
    p, t, hit = autotune.tune(g, vel, shots, nt_probe=120, reps=1, cache_path=cache, verbose=True)
    assert not hit and t > 0 and p in autotune.candidates(len(shots))
    p2, t2, hit2 = autotune.tune(g, vel, shots, cache_path=cache)
    assert hit2 and p2 == p and t2 == t
    assert autotune.shape_key(g, shots) in json.load(open(cache))
    rng = np.random.default_rng(22)
    res = [rng.standard_normal((s.nr, s.nt)).astype(np.float32) for s in shots]
    ctx = autotune.make_context(g, shots, p, 200)
    gt = gradient(ctx, vel, shots, res)
    ctx.close()
    ctx = autotune.make_context(g, shots, dict(autotune.DEFAULT), 200)
    gd = gradient(ctx, vel, shots, res)
    ctx.close()
    assert rel(gt, gd) < 1e-12
