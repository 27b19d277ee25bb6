"""The no-free-surface helpers with bodies in the reference: model_ext_fs_no
(fwi_2d.c:29-65, host) against numpy's edge padding, and write_bdr_fs_no
(fwi_2d.c:962-997, device) against the strips cut from the field by slicing
(its physical grid sits npml rows down)."""
import numpy as np
import pytest

from fwihip import lib as L
from fwihip.workloads import Geometry


@pytest.mark.parametrize("n2,n1,en2,en1", [(7, 9, 13, 15), (1, 1, 5, 3), (4, 6, 4, 6), (5, 3, 25, 43)])
def test_model_ext_fs_no_is_edge_padding(n2, n1, en2, en1):
    m = np.random.default_rng(n1 * 10 + n2).standard_normal((n2, n1)).astype(np.float32)
    out = L.model_ext_fs_no(m, en1, en2)
    g2, g1 = (en2 - n2) // 2, (en1 - n1) // 2
    np.testing.assert_array_equal(out, np.pad(m, ((g2, g2), (g1, g1)), mode="edge"))


def test_model_ext_fs_no_refuses_uneven_margins():
    m = np.ones((4, 4), np.float32)
    for en1, en2 in ((7, 6), (6, 7), (3, 4)):
        with pytest.raises(L.FwiError) as e:
            L.model_ext_fs_no(m, en1, en2)
        assert e.value.code == L.EINVAL


@pytest.mark.gpu
def test_write_bdr_fs_no_strips():
    L.load()
    if L.device_count() < 1:
        pytest.skip("no GPU")
    g = Geometry(nx=37, nz=23, npml=6, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(9)
    pf = rng.standard_normal((g.nz + 2 * g.npml, g.nx + 2 * g.npml)).astype(np.float32)
    ud, lr = L.DevArray((6, g.nx)), L.DevArray((g.nz, 6))
    L.write_bdr_fs_no(g, ud, lr, L.DevArray(pf.shape, pf))
    b = pf[g.npml:g.npml + g.nz, g.npml:g.npml + g.nx]
    want_ud = np.concatenate([b[:3], b[::-1][:3]])
    want_lr = np.concatenate([b[:, :3], b[:, ::-1][:, :3]], axis=1)
    np.testing.assert_array_equal(ud.numpy(), want_ud)
    np.testing.assert_array_equal(lr.numpy(), want_lr)
    # with the free surface the same call on the top-aligned field gives the rows at 0
    pfs = pf[g.npml:].copy()
    L.write_bdr(g, ud, lr, L.DevArray(pfs.shape, pfs))
    np.testing.assert_array_equal(ud.numpy(), want_ud)
    np.testing.assert_array_equal(lr.numpy(), want_lr)
