"""Host input layer (fwi_io.c): the fwi_init file reads and index arithmetic
(fwi_2d.c:320-371, fwi_files_path.h:7-18, read_write.c:4-56) and the
trcinp natural-cubic-spline resampling (spline_interp.c:19-38).  CPU only."""
import ctypes
import math
import os

import numpy as np
import pytest

from fwihip import plans as PC
from fwihip import lib as L


def c_round(x):
    """C round(): half away from zero"""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def write(path, a):
    np.asarray(a, np.float32).tofile(path)


@pytest.fixture()
def h():
    return PC.lib()


def make_inputs(tmp, nx, nz, nt, nt0, nr, ishot, rng, dx, dz):
    d = os.path.join(tmp, "input")
    os.makedirs(d, exist_ok=True)
    obs0 = rng.standard_normal((nr, nt0)).astype(np.float32)
    # source then receivers (x, z) in metres, with z values on the .5 rounding edges
    xyz = np.zeros((nr + 1) * 2, np.float32)
    xyz[0::2] = rng.uniform(0, (nx - 1) * dx, nr + 1)
    xyz[1::2] = rng.uniform(0, (nz - 1) * dz, nr + 1)
    xyz[1] = 2.0 * dz   # z/dz - 0.5 = 1.5 -> C round gives 2 (numpy would give 2 too; 2.5 below gives 3)
    xyz[3] = 3.0 * dz   # 2.5 -> 3 (numpy's round-half-even would give 2)
    xyz[2] = 2.5 * dx   # 2.5 -> 3
    md = rng.uniform(1500, 4500, (nz, nx)).astype(np.float32)
    btry = rng.uniform(0, nz * dz, nx).astype(np.float32)
    wav = rng.standard_normal(nt).astype(np.float32)
    fbt = rng.uniform(0, 0.5, nr).astype(np.float32)
    write(os.path.join(d, f"shot2ms_{ishot}.bin"), obs0)
    write(os.path.join(d, f"sr_location_{ishot}.bin"), xyz)
    write(os.path.join(d, f"fbt_{ishot}.bin"), fbt)
    write(os.path.join(d, "initial_model.bin"), md)
    write(os.path.join(d, "bottom.bin"), btry)
    write(os.path.join(d, "wavelet_1ms.bin"), wav)
    return obs0, xyz, md, btry, wav, fbt


def test_plan_read_inputs_matches_fwi_init(h, tmp_path):
    nx, nz, npml, nt, jump, nr, dx, dz, dt = 30, 20, 5, 41, 2, 7, 12.5, 10.0, 1e-3
    nt0 = (nt - 1) // jump + 1
    rng = np.random.default_rng(11)
    obs0, xyz, md, btry, wav, fbt = make_inputs(str(tmp_path), nx, nz, nt, nt0, nr, 3, rng, dx, dz)
    p = h.fwi_plan_alloc(nx, nz, npml, nt, jump, nr, dx, dz, dt)
    common = (str(tmp_path) + "/").encode()
    assert h.fwi_plan_read_inputs(p, common, 3, None, 1, 1, 2.5) == 0, h.fwi_io_last_error()
    P = p.contents
    np.testing.assert_array_equal(PC.view(P.msf.contents.obs0[0], (nr, nt0)), obs0)
    np.testing.assert_array_equal(PC.view(P.model.contents.md[0], (nz, nx)), md)
    np.testing.assert_array_equal(np.ctypeslib.as_array(P.sou.contents.wav0, (nt,)), wav * np.float32(2.5))
    ix = np.ctypeslib.as_array(P.sou.contents.ixyz_sr, ((nr + 1) * 2,))
    f32 = np.float32
    for i in range(nr + 1):
        xq = float(f32(xyz[2 * i]) / f32(dx))                  # float quotient (fwi_2d.c:363-371)
        zq = float(f32(xyz[2 * i + 1]) / f32(dz)) - 0.5        # then the double literal
        assert ix[2 * i] == c_round(xq)
        assert ix[2 * i + 1] == c_round(zq)
    assert (ix[1], ix[3], ix[2]) == (2, 3, 3)
    ib = np.ctypeslib.as_array(P.model.contents.ibtry, (nx,))
    np.testing.assert_array_equal(ib, [int(float(f32(b) / f32(dz)) - 0.5) for b in btry])
    ifb = np.ctypeslib.as_array(P.msf.contents.ifbt, (nr,))
    np.testing.assert_array_equal(ifb, [int(f32(f) / f32(dt * jump)) for f in fbt])
    h.fwi_plan_free(p)


def test_rwi_plan_read_inputs_matches_rwi_init(h, tmp_path):
    """rwi_init (fwi_2d.c:637-688) reads what fwi_init reads except the
    initial model, plus the two models m0, m1: the RWI reader leaves every
    shared field as the FWI reader does, and m0, m1 as written."""
    nx, nz, npml, nt, jump, nr, dx, dz, dt = 30, 20, 5, 41, 2, 7, 12.5, 10.0, 1e-3
    nt0 = (nt - 1) // jump + 1
    rng = np.random.default_rng(12)
    obs0, xyz, md, btry, wav, fbt = make_inputs(str(tmp_path), nx, nz, nt, nt0, nr, 4, rng, dx, dz)
    m0 = rng.uniform(1500, 4500, (nz, nx)).astype(np.float32)
    m1 = (0.05 * rng.standard_normal((nz, nx))).astype(np.float32)
    f0, f1 = str(tmp_path / "m0.bin"), str(tmp_path / "m1.bin")
    write(f0, m0)
    write(f1, m1)
    common = (str(tmp_path) + "/").encode()
    pf = h.fwi_plan_alloc(nx, nz, npml, nt, jump, nr, dx, dz, dt)
    pr = h.rwi_plan_alloc(nx, nz, npml, nt, jump, nr, dx, dz, dt)
    try:
        assert h.fwi_plan_read_inputs(pf, common, 4, None, 1, 1, 1.5) == 0, h.fwi_io_last_error()
        assert h.rwi_plan_read_inputs(pr, common, 4, f0.encode(), f1.encode(), None, 1, 1, 1.5) == 0, \
            h.fwi_io_last_error()
        F, R = pf.contents, pr.contents
        np.testing.assert_array_equal(PC.view(R.model.contents.m0[0], (nz, nx)), m0)
        np.testing.assert_array_equal(PC.view(R.model.contents.m1[0], (nz, nx)), m1)
        np.testing.assert_array_equal(PC.view(R.msf.contents.obs0[0], (nr, nt0)), obs0)
        for name, n in (("xyz_sr", 2 * (nr + 1)), ("ixyz_sr", 2 * (nr + 1)), ("wav0", nt)):
            np.testing.assert_array_equal(np.ctypeslib.as_array(getattr(R.sou.contents, name), (n,)),
                                          np.ctypeslib.as_array(getattr(F.sou.contents, name), (n,)))
        for name, n in (("btry", nx), ("ibtry", nx)):
            np.testing.assert_array_equal(np.ctypeslib.as_array(getattr(R.model.contents, name), (n,)),
                                          np.ctypeslib.as_array(getattr(F.model.contents, name), (n,)))
        for name in ("fbt", "ifbt"):
            np.testing.assert_array_equal(np.ctypeslib.as_array(getattr(R.msf.contents, name), (nr,)),
                                          np.ctypeslib.as_array(getattr(F.msf.contents, name), (nr,)))
        assert h.rwi_plan_read_inputs(pr, common, 4, None, f1.encode(), None, 1, 1, 1.5) == L.EINVAL
        assert "m0" in h.fwi_io_last_error().decode()
        os.remove(f1)
        assert h.rwi_plan_read_inputs(pr, common, 4, f0.encode(), f1.encode(), None, 1, 1, 1.5) == L.EINVAL
        assert "m1.bin" in h.fwi_io_last_error().decode()
    finally:
        h.fwi_plan_free(pf)
        h.rwi_plan_free(pr)


@pytest.mark.parametrize("case", ["missing", "long", "short"])
def test_read_errors_return_codes(h, tmp_path, case):
    nx, nz, npml, nt, jump, nr = 12, 10, 4, 21, 2, 3
    rng = np.random.default_rng(2)
    make_inputs(str(tmp_path), nx, nz, nt, (nt - 1) // jump + 1, nr, 0, rng, 10.0, 10.0)
    mdl = os.path.join(str(tmp_path), "input", "initial_model.bin")
    if case == "missing":
        os.remove(mdl)
    elif case == "long":
        write(mdl, np.zeros(nx * nz + 1))
    else:
        write(mdl, np.zeros(nx * nz - 1))
    p = h.fwi_plan_alloc(nx, nz, npml, nt, jump, nr, 10.0, 10.0, 1e-3)
    rc = h.fwi_plan_read_inputs(p, (str(tmp_path) + "/").encode(), 0, None, 1, 0, 1.0)
    assert rc == L.EINVAL
    msg = h.fwi_io_last_error().decode()
    assert "initial_model.bin" in msg
    assert {"missing": "opening", "long": "left", "short": "end of file"}[case] in msg
    h.fwi_plan_free(p)


def test_write_read_round_trip(h, tmp_path):
    a = np.random.default_rng(0).standard_normal(257).astype(np.float32)
    f = str(tmp_path / "a.bin").encode()
    assert h.fwi_write_f32(f, a.ctypes.data_as(PC.PF), a.size) == 0
    b = np.zeros_like(a)
    assert h.fwi_read_f32(f, b.ctypes.data_as(PC.PF), b.size, 1) == 0
    np.testing.assert_array_equal(a, b)


def trcinp(h, obs0, dt0, nt, dt):
    obs0 = np.ascontiguousarray(obs0, np.float32)
    out = np.zeros((obs0.shape[0], nt), np.float32)
    assert h.fwi_trcinp(PC.rows(obs0), obs0.shape[1], dt0, PC.rows(out), nt, dt, obs0.shape[0]) == 0
    return out


@pytest.mark.parametrize("jump", [2, 3, 4])
def test_trcinp_is_the_natural_cubic_spline(h, jump):
    from scipy.interpolate import CubicSpline
    nt0, dt = 101, np.float32(1e-3)
    nt = (nt0 - 1) * jump + 1
    dt0 = np.float32(dt * jump)
    rng = np.random.default_rng(jump)
    y = rng.standard_normal((5, nt0)).astype(np.float32)
    out = trcinp(h, y, dt0, nt, dt)
    x0 = (np.arange(nt0, dtype=np.float32) * dt0).astype(np.float64)   # float products, as spline_interp.c:21-25
    x1 = (np.arange(nt, dtype=np.float32) * dt).astype(np.float64)
    for r in range(y.shape[0]):
        ref = CubicSpline(x0, y[r].astype(np.float64), bc_type="natural")(x1)
        ref[0], ref[-1] = y[r, 0], y[r, -1]
        np.testing.assert_allclose(out[r], ref.astype(np.float32), rtol=0, atol=2e-6 * np.abs(y[r]).max())
        same = x1[::jump] == x0   # knots reproduced where the float abscissae coincide
        assert same.sum() > nt0 // 4
        np.testing.assert_array_equal(out[r, ::jump][same], y[r][same])


def test_trcinp_exact_on_lines_and_natural_ends(h):
    nt0, jump, dt = 41, 5, np.float32(2e-3)
    dt0 = np.float32(dt * jump)
    nt = (nt0 - 1) * jump + 1
    t0 = np.arange(nt0) * float(dt0)
    y = np.stack([3.0 - 2.0 * t0 / t0[-1], np.full(nt0, 0.75)]).astype(np.float32)
    out = trcinp(h, y, dt0, nt, dt)
    t1 = np.arange(nt) * float(dt)
    np.testing.assert_allclose(out[0], 3.0 - 2.0 * t1 / t0[-1], atol=2e-6)
    np.testing.assert_allclose(out[1], 0.75, atol=1e-7)
    # natural ends: zero curvature at the first knot, so a parabola is not
    # reproduced there (a not-a-knot spline of the same samples is, exactly)
    from scipy.interpolate import CubicSpline
    yq = (t0 / t0[-1]) ** 2
    oq = trcinp(h, yq[None, :].astype(np.float32), dt0, nt, dt)[0].astype(np.float64)
    nak = CubicSpline(t0, yq, bc_type="not-a-knot")(t1)
    curv = lambda a, i: a[i + 1] - 2 * a[i] + a[i - 1]
    mid = curv(nak, nt // 2)
    assert abs(curv(nak, 1) - mid) < 1e-3 * abs(mid)
    assert abs(curv(oq, 1)) < 0.35 * abs(mid)


def test_trcinp_rejects_bad_arguments(h):
    a = np.zeros((1, 4), np.float32)
    b = np.zeros((1, 7), np.float32)
    assert h.fwi_trcinp(PC.rows(a), 1, 1e-3, PC.rows(b), 7, 5e-4, 1) == L.EINVAL
    assert h.fwi_trcinp(PC.rows(a), 4, 0.0, PC.rows(b), 7, 5e-4, 1) == L.EINVAL


def test_residual_mode_setter(h):
    assert h.grad4_set_residual_mode(2) == L.EINVAL
    assert h.grad4_set_residual_mode(1) == 0
    assert h.grad4_set_residual_mode(0) == 0
