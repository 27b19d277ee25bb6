"""Host-side fault injection (no GPU): models the propagator must refuse
(non-finite, non-positive, above the stability limit), with the limit itself
pinned by running the FP64 oracle just below and just above it, and HBM
budgets the memory planner cannot meet."""
import math

import numpy as np
import pytest

from oracle import Oracle
from fwihip import lib as L
from fwihip.workloads import Geometry, Shot, ricker

C0, C1 = 0.1129042e1, -0.4301412e-1          # the staggered pair of fwi_fdcoe (fwi_2d.c:302-316)


def z_operator(pnz):
    """-d/dz (vz) d/dz (v) on pnz rows with the free-surface rows on top
    (fwi_2d.c:805-807, 848) and the two-point rows at the bottom, dz = 1"""
    vnz = pnz + 1
    dv = np.zeros((vnz, pnz))
    for iz in range(2, vnz - 2):
        dv[iz, [iz - 2, iz - 1, iz, iz + 1]] = [-C1, -C0, C0, C1]
    dv[0, :4] = np.array([3675, -1225, 441, -75]) / 840
    dv[1, :4] = np.array([-1085, 1015, -63, 5]) / 840
    dv[vnz - 2, [pnz - 2, pnz - 1]] = [-1, 1]
    dz = np.zeros((pnz, vnz))
    for iz in range(1, pnz - 1):
        dz[iz, [iz - 1, iz, iz + 1, iz + 2]] = [-C1, -C0, C0, C1]
    dz[0, :5] = np.array([-22, 17, 9, -5, 1]) / 24
    dz[pnz - 1, [vnz - 2, vnz - 1]] = [-1, 1]
    return -dz @ dv


LZ_FS = 6.0195791                              # the library's bound for the z operator


def vlimit(g):
    return 2.0 / (g.dt * math.sqrt(4 * (C0 - C1) ** 2 / g.dx ** 2 + LZ_FS / g.dz ** 2))


def test_free_surface_bound_is_the_z_operators_largest_eigenvalue():
    for n in (30, 120):
        ev = np.linalg.eigvals(z_operator(n))
        assert np.abs(ev.imag).max() < 1e-9 and ev.real.min() > -1e-9   # real, non-negative: leapfrog applies
        assert abs(ev.real.max() - LZ_FS) < 1e-6
        assert ev.real.max() > 4 * (C0 - C1) ** 2                        # above the interior Nyquist value


def test_model_values_refused_with_the_cell():
    g = Geometry(nx=30, nz=20, npml=10, dx=10.0, dz=10.0, dt=1e-3)
    vel = np.full((g.nz, g.nx), 2000.0, np.float32)
    L.check_model(g, vel)
    for bad, cell in ((np.nan, 37), (np.inf, 5), (0.0, 599), (-1500.0, 0)):
        v = vel.copy()
        v.flat[cell] = bad
        with pytest.raises(L.FwiError) as e:
            L.check_model(g, v)
        assert e.value.code == L.EINVAL and f"cell {cell}" in str(e.value), str(e.value)
    with pytest.raises(ValueError):
        L.check_model(g, vel[:, :-1])


@pytest.mark.parametrize("dx,dz,dt", [(10.0, 10.0, 1e-3), (12.5, 5.0, 4e-4), (20.0, 20.0, 2e-3)])
def test_stability_limit_refused(dx, dz, dt):
    g = Geometry(nx=40, nz=30, npml=10, dx=dx, dz=dz, dt=dt)
    vm = vlimit(g)
    vel = np.full((g.nz, g.nx), 0.5 * vm, np.float32)
    vel[7, 11] = 0.999 * vm
    L.check_model(g, vel)
    vel[20, 3] = 1.001 * vm
    with pytest.raises(L.FwiError) as e:
        L.check_model(g, vel)
    assert "stability limit" in str(e.value) and f"cell {20 * g.nx + 3}" in str(e.value)


@pytest.mark.parametrize("dx,dz", [(10.0, 10.0), (10.0, 4.0), (5.0, 12.5)])
def test_stability_limit_is_where_the_scheme_blows_up(dx, dz):
    """0.5 % above the limit the unstable mode grows by about 1.2 per step
    from rounding and overflows; 0.5 % below it the field decays."""
    g = Geometry(nx=48, nz=40, npml=10, dx=dx, dz=dz, dt=1e-4 * min(dx, dz))
    nt = 900
    irx = np.arange(0, g.nx, 3, dtype=np.int32)
    shot = Shot(nt=nt, isx=g.nx // 2, isz=g.nz // 2, wav=ricker(nt, g.dt, 0.25 / g.dt / 10), irx=irx,
                irz=np.full(len(irx), 5, np.int32))
    o = Oracle(np.float64)
    peak = {}
    for f in (0.995, 1.005):
        with np.errstate(all="ignore"):
            syn = o.forward(g, np.full((g.nz, g.nx), f * vlimit(g)), shot)["syn"]
        a = np.abs(syn)
        peak[f] = (float(a[:, :nt // 3].max()), float(a[:, -nt // 6:].max()) if np.isfinite(a).all() else math.inf)
    print(f"dx {dx} dz {dz}: early/late peak at 0.995 of the limit {peak[0.995]}, at 1.005 {peak[1.005]}")
    assert peak[0.995][1] < peak[0.995][0]
    assert peak[1.005][1] > 1e20 * peak[1.005][0]


def test_hbm_budget_below_the_context_is_refused():
    g = Geometry(nx=1024, nz=320, npml=20, dx=10.0, dz=10.0, dt=0.8e-3)
    _, _, need = L.plan_memory(g, 64, 4000, 1024, 0)
    for budget in (1, 1 << 20, 1 << 30):
        with pytest.raises(L.FwiError) as e:
            L.plan_memory(g, 64, 4000, 1024, budget)
        assert e.value.code == L.ENOMEM
    m, sn, b = L.plan_memory(g, 64, 4000, 1024, need // 2)
    assert m == 2 and b <= need // 2
