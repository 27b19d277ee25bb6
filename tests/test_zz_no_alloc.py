"""fwi_hip_op_pml_noalloc (the allocation-free a_lap_opt_2d_pml step) against
fwi_hip_op_pml, through the C ABI (needs an MI355X).

Both entry points run the same reference-structured kernels; they differ only
in where the CPML tables live (a fresh allocation per call against the
library's static device array), so every output and every CPML memory array
must agree bit for bit.  Two calls in a row, the second on a smaller grid
with a thinner CPML, so that the second call also runs over the longer table
the first one left in the array.
"""
import numpy as np
import pytest

from fwihip.workloads import Geometry

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "ux", "vz", "u0", "u1", "v1", "ux0", "ux1", "vz1")


def shapes(g):
    pnx, pnz, n = g.pnx, g.pnz, g.npml
    return dict(u=(pnz, pnx + 1), v=(pnz + 1, pnx), ux=(pnz, pnx), vz=(pnz, pnx), u0=(pnz, n + 1),
                u1=(pnz, n + 1), v1=(n + 1, pnx), ux0=(pnz, n), ux1=(pnz, n), vz1=(n, pnx))


@pytest.fixture(scope="module")
def lib():
    from fwihip import lib as L
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    return L


def run_both(L, g, seed, steps):
    rng = np.random.default_rng(seed)
    sh = shapes(g)
    # the same (nonzero) starting state for both entry points
    init = {k: rng.standard_normal(s).astype(np.float32) for k, s in sh.items()}
    a = {k: L.DevArray(s, init[k]) for k, s in sh.items()}
    b = {k: L.DevArray(s, init[k]) for k, s in sh.items()}
    pa = L.DevArray((g.pnz, g.pnx))
    for _ in range(steps):
        pa.upload(rng.standard_normal((g.pnz, g.pnx)).astype(np.float32))
        L.op_pml(g, pa, *[a[k] for k in NAMES])
        L.op_pml_noalloc(g, pa, *[b[k] for k in NAMES])
    for k in NAMES:
        np.testing.assert_array_equal(b[k].numpy(), a[k].numpy(), err_msg=f"{k} on {g}")
        assert np.abs(a[k].numpy()).max() > 0


def test_noalloc_matches_op_pml_twice_second_smaller(lib):
    print("noalloc: call 1, 70x40 npml 9, 3 steps", flush=True)
    run_both(lib, Geometry(nx=70, nz=40, npml=9, dx=10.0, dz=8.0, dt=1e-3), seed=21, steps=3)
    print("noalloc: call 2, 33x20 npml 4, 2 steps", flush=True)
    run_both(lib, Geometry(nx=33, nz=20, npml=4, dx=12.5, dz=10.0, dt=1e-3), seed=22, steps=2)


def test_noalloc_refuses_npml_beyond_the_static_array(lib):
    g = Geometry(nx=40, nz=20, npml=65, dx=10.0, dz=10.0, dt=1e-3)
    dummy = [lib.DevArray((4,)) for _ in range(11)]
    with pytest.raises(lib.FwiError) as e:
        lib.op_pml_noalloc(g, *dummy)
    assert e.value.code == lib.EINVAL
