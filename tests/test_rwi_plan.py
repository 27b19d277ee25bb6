"""RWI plan layout (include/fwi_plan.h rwi_plan = header.h:187-220): the C
allocator's fields read back through the ctypes mirror at the offsets the
reference's layout puts them, and the driver refuses bad input without a
device call."""
import ctypes

import numpy as np

from fwihip import plans as PC


def test_rwi_plan_alloc_fields():
    h = PC.lib()
    nx, nz, npml, nt, jump, nr = 37, 21, 5, 41, 4, 9
    p = h.rwi_plan_alloc(nx, nz, npml, nt, jump, nr, 10.0, 12.5, 1e-3)
    assert p
    P = p.contents
    b = P.grid.contents.base.contents
    assert (b.nx, b.nz, b.npml, b.nt, b.free_surface) == (nx, nz, npml, nt, 1)
    f = P.grid.contents.f.contents
    assert (f.pnx, f.pnz) == (nx + 2 * npml, nz + npml)       # fwi_2d.c:154-157
    m = P.msf.contents
    assert (m.nt, m.nr, m.jump) == ((nt - 1) // jump + 1, nr, jump)
    assert abs(m.dt - 4e-3) < 1e-9
    assert abs(P.sou.contents.iarea - 1.0 / (10.0 * 12.5)) < 1e-9
    M = P.model.contents
    # row-pointer arrays over one slab (alloc.c:3-16), background and perturbation distinct
    m0 = PC.view(M.m0[0], (nz, nx))
    m1 = PC.view(M.m1[0], (nz, nx))
    m0[:] = 1.0
    m1[:] = 2.0
    assert ctypes.addressof(M.m0[1].contents) - ctypes.addressof(M.m0[0].contents) == 4 * nx
    assert float(PC.view(M.m0[0], (nz, nx)).sum()) == nz * nx
    ek = PC.view(M.ekp_total[0], (nz + npml, nx + 2 * npml))
    assert ek.shape == (nz + npml, nx + 2 * npml)
    assert P.grad.contents.jump == 1
    assert PC.view(P.grad.contents.body, (nz, nx)).shape == (nz, nx)
    assert not P.fw_inc and not P.pml_per and not P.path
    h.rwi_plan_free(p)


def test_rwi_plan_alloc_rejects_bad_shapes():
    h = PC.lib()
    assert not h.rwi_plan_alloc(0, 10, 5, 41, 1, 3, 10.0, 10.0, 1e-3)
    assert not h.rwi_plan_alloc(10, 10, 5, 41, 3, 3, 10.0, 10.0, 1e-3)   # (nt - 1) % jump != 0


def test_grad_all_shots_rfl_refuses_empty_input():
    h = PC.lib()
    assert h.grad_all_shots_rfl(None, 0, 0, None) != 0
    assert b"no plans" in h.grad4_last_error()
