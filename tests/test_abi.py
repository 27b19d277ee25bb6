"""C ABI checks that run without a GPU: the library loads, exports every
symbol include/fwi_hip.h declares, and rejects bad arguments before any
device call."""
import ctypes
import os
import re

import numpy as np
import pytest

from fwihip import lib as L
from fwihip.workloads import Geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|size_t|const char \*)\s*\*?\s*(\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared("fwi_hip.h")
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(L.EXPORTS) <= set(names)


def test_plan_header_symbols_exported():
    lib = L.load()
    for hdr in ("fwi_plan.h", "fwi_io.h", "fwi_sigproc.h", "fwi_msf.h", "fwi_gradprec.h", "fwi_lbfgs.h"):
        names = declared(hdr)
        assert names, hdr
        for n in names:
            assert hasattr(lib, n), n


def _desc(**kw):
    g = Geometry(nx=kw.pop("nx", 32), nz=kw.pop("nz", 24), npml=kw.pop("npml", 6), dx=10.0, dz=10.0, dt=1e-3)
    d = L.Desc(L.make_grid(g), kw.pop("max_shots", 1), kw.pop("max_nt", 10), kw.pop("max_nr", 4),
               kw.pop("tile_z", 0), kw.pop("shots_per_group", 0), kw.pop("hbm_budget", 0), None)
    return d


@pytest.mark.parametrize("kw", [dict(nx=4), dict(nz=5), dict(npml=1), dict(max_shots=0), dict(max_nt=0),
                                dict(tile_z=5), dict(shots_per_group=9)])
def test_create_rejects_bad_descriptors(kw):
    lib = L.load()
    h = ctypes.c_void_p()
    rc = lib.fwi_hip_create(ctypes.byref(_desc(**kw)), ctypes.byref(h))
    assert rc == L.EINVAL
    assert not h.value
    assert lib.fwi_hip_last_error(None)


def test_null_arguments():
    lib = L.load()
    assert lib.fwi_hip_create(None, None) == L.EINVAL
    assert lib.fwi_hip_set_model(None, None) == L.EINVAL
    assert lib.fwi_hip_forward(None, None) == L.EINVAL
    assert lib.fwi_hip_adjoint(None, None, None) == L.EINVAL


def test_cpml_coe_host_matches_oracle():
    from oracle import Oracle
    o = Oracle(np.float32)
    a, b = L.cpml_coe(20, 10.0, 1e-3)
    ao, bo = o.cpml_coe(20, 10.0, 1e-3)
    np.testing.assert_array_equal(a, ao)
    np.testing.assert_array_equal(b, bo)


def test_plan_layout_and_sampling_check():
    """fwi_plan_alloc builds the reference's shapes (fwi_2d.c:148-165), the
    ctypes mirror of header.h reads them back, and grad_all_shots rejects
    (nt0-1)*jump != nt-1 (fwi_2d.c:195) before touching the device."""
    from fwihip import plans as PC
    h = PC.lib()
    p = h.fwi_plan_alloc(30, 20, 5, 41, 2, 7, 10.0, 10.0, 1e-3)
    P = p.contents
    f, b = P.grid.contents.f.contents, P.grid.contents.b.contents
    assert (f.pnx, f.pnz, f.unx, f.unz, f.vnx, f.vnz) == (40, 25, 41, 25, 40, 26)
    assert (b.pnx, b.pnz, b.unx, b.vnz) == (30, 20, 31, 21)
    msf = P.msf.contents
    assert (msf.nt, msf.nr, msf.jump) == (21, 7, 2)
    assert abs(P.sou.contents.iarea - 0.01) < 1e-9
    base = P.grid.contents.base.contents
    assert (base.nx, base.nz, base.npml, base.nt, base.free_surface) == (30, 20, 5, 41, 1)
    off = PC.FwiMsf.nt.offset
    ctypes.c_size_t.from_address(ctypes.addressof(msf) + off).value = 20
    arr = (ctypes.POINTER(PC.FwiPlan) * 1)(p)
    assert h.grad_all_shots(arr, 1, 0) == L.EINVAL
    assert b"nt0" in h.grad4_last_error()
    ctypes.c_size_t.from_address(ctypes.addressof(msf) + off).value = 21
    # the data rate must be jump x the propagation rate, within 1e-3 s (fwi_2d.c:195)
    assert abs(msf.dt - 2e-3) < 1e-9
    msf.dt = 4e-3                            # 2 ms off, beyond the reference's 1e-3 s tolerance
    assert h.grad_all_shots(arr, 1, 0) == L.EINVAL
    assert b"msf.dt" in h.grad4_last_error()
    msf.dt = 2e-3
    h.fwi_plan_free(p)
    assert not h.fwi_plan_alloc(30, 20, 5, 40, 2, 7, 10.0, 10.0, 1e-3)  # (nt-1) % jump != 0
