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
    for n in declared("fwi_plan.h"):
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
