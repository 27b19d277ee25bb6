"""Pins the CPU oracle (oracle/fwi_oracle.c) to mathematics.

The reference ships no tests or data and does not build here, so each oracle
function is checked against something a plausible mistake (dropped term,
wrong sign or index, swapped operand) cannot satisfy: exact rational weights,
closed forms of the stencils on polynomials, the geometric-sum solution of the
CPML recursion, the leapfrog's cosine solution, brute-force index maps, the
adjoint dot-product identity, and a second-order Taylor remainder.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import Oracle, MAX_F, MAX_V, COEF_R
from fwihip.workloads import Geometry, Shot, ricker, config1, cell_x, cell_z

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "closed_forms.json")))


@pytest.fixture(scope="module")
def o64():
    return Oracle(np.float64)


@pytest.fixture(scope="module")
def o32():
    return Oracle(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


# --------------------------------------------------------------------- a5
def test_cpml_closed_form_config1(o64, o32):
    g = GOLD["cpml_config1"]
    for o in (o64, o32):
        a, b = o.cpml_coe(20, 10.0, 1e-3)
        tol = 1e-12 if o.dtype == np.float64 else 2e-6
        assert abs(b[0] - g["b_outer"]) < tol * 10
        assert abs(a[0] - g["a_outer"]) < tol * 10
        assert a[40] == 0.0                      # d = 0 on the inner edge
        assert abs(b[40] - g["b_inner"]) < tol * 10
        assert np.all(a[:40] < 0) and np.all((b > 0) & (b <= 1))
    # d0 and b0[0] for npml = 20, dx = 10 m, dt = 1 ms, to the digits quoted for this geometry
    assert abs(g["d0"] - 310.85) < 0.01 and abs(g["b_outer"] - 0.73282) < 1e-5


def test_cpml_mirror_tables(o64):
    tab = o64.cpml_tables(12, 12.5, 10.0, 1e-3)
    ax, bx = o64.cpml_coe(12, 12.5, 1e-3)
    az, bz = o64.cpml_coe(12, 10.0, 1e-3)
    np.testing.assert_array_equal(tab[0], ax)
    np.testing.assert_array_equal(tab[1], bx)
    np.testing.assert_array_equal(tab[2], ax[::-1])
    np.testing.assert_array_equal(tab[3], bx[::-1])
    np.testing.assert_array_equal(tab[4], az[::-1])
    np.testing.assert_array_equal(tab[5], bz[::-1])


# --------------------------------------------------------------------- a6
def test_fd_coefficients(o64):
    c = o64.fdcoe(12.5, 10.0)
    assert abs(c[0] * 12.5 - 1.129042) < 1e-12 and abs(c[1] * 12.5 + 0.04301412) < 1e-12
    assert abs(c[3] * 10.0 - 1.129042) < 1e-12
    assert abs((c[0] + 3 * c[1]) * 12.5 - float(Fraction(GOLD["c0_plus_3c1_times_dx"]))) < 1e-12
    assert abs(c[6] - 0.1 / 840) < 1e-15 and abs(c[7] - 0.1 / 24) < 1e-15


# --------------------------------------------------------------------- a9
def test_model_ext_is_edge_padding(o64):
    m = np.random.default_rng(1).standard_normal((7, 9))
    em = o64.model_ext(m, 4)
    np.testing.assert_array_equal(em, np.pad(m, ((0, 4), (4, 4)), mode="edge"))


# --------------------------------------------------------------------- a1
def _quadratic(pnz, pnx, a, b, c, d, e):
    iz, ix = np.mgrid[0:pnz, 0:pnx].astype(np.float64)
    return a * ix ** 2 + b * iz ** 2 + c * ix * iz + d * ix + e * iz


def test_op_interior_on_quadratics(o64):
    """u, v, ux, vz of a quadratic are s*(exact derivative) with
    s = (c0 + 3 c1)*dx, at every cell the interior formulas cover."""
    dx = dz = 1.0
    coe = o64.fdcoe(dx, dz)
    s = float(Fraction(GOLD["c0_plus_3c1_times_dx"]))
    pnz, pnx = 20, 24
    a, b, c, d, e = 0.3, -0.7, 0.25, 1.5, -2.0
    p = _quadratic(pnz, pnx, a, b, c, d, e)
    u, v, ux, vz = o64.op_interior(p, coe)
    iz, ix = np.mgrid[0:pnz, 0:pnx + 1].astype(np.float64)
    u_ex = s * (2 * a * (ix - 0.5) + c * iz + d)
    np.testing.assert_allclose(u[2:pnz - 2, 2:pnx - 1], u_ex[2:pnz - 2, 2:pnx - 1], rtol=0, atol=1e-11)
    iz, ix = np.mgrid[0:pnz + 1, 0:pnx].astype(np.float64)
    v_ex = s * (2 * b * (iz - 0.5) + c * ix + e)
    np.testing.assert_allclose(v[2:pnz - 1, 2:pnx - 2], v_ex[2:pnz - 1, 2:pnx - 2], rtol=0, atol=1e-11)
    np.testing.assert_allclose(ux[3:-3, 3:-3], s * s * 2 * a, atol=1e-11)
    np.testing.assert_allclose(vz[3:-3, 3:-3], s * s * 2 * b, atol=1e-11)


def test_free_surface_rows_are_lagrange_derivatives(o64):
    """v[0], v[1] and vz[0] of a_lap_opt_2d_pml (fwi_2d.c:805-807,848) are
    the exact derivatives of the interpolants; weights from exact arithmetic."""
    w0 = [float(Fraction(x)) / 840 for x in GOLD["v_row0_times_840"]]
    w1 = [float(Fraction(x)) / 840 for x in GOLD["v_row1_times_840"]]
    wz = [float(Fraction(x)) / 24 for x in GOLD["vz_row0_times_24"]]
    npml, pnx, pnz = 4, 16, 14
    coe = o64.fdcoe(1.0, 1.0)
    tab = np.zeros((6, 2 * npml + 1))
    st = o64.new_pml_state(pnx, pnz, npml)
    p = np.random.default_rng(3).standard_normal((pnz, pnx))
    o64.op_pml(p, st, coe, tab, npml)
    v = st["v"]
    np.testing.assert_allclose(v[0], sum(w * p[k] for k, w in enumerate(w0)), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(v[1], sum(w * p[k] for k, w in enumerate(w1)), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(st["vz"][0], sum(w * v[k] for k, w in enumerate(wz)), rtol=1e-13, atol=1e-13)
    # a quartic vanishing at the surface z = -1/2 is differentiated exactly
    z = np.arange(pnz, dtype=np.float64)[:, None]
    q = (z + 0.5) * (z - 1.3) * (z + 2.1) * (z - 0.4) * np.ones((1, pnx))
    dq = np.polyder(np.poly1d(np.poly([-0.5, 1.3, -2.1, 0.4])))
    st = o64.new_pml_state(pnx, pnz, npml)
    o64.op_pml(q, st, coe, tab, npml)
    np.testing.assert_allclose(st["v"][0], dq(-0.5), rtol=1e-12)
    np.testing.assert_allclose(st["v"][1], dq(0.5), rtol=1e-12)


def test_op_pml_edges_and_never_written(o64):
    """Edge rows are 2nd order and u[:,0], u[:,unx-1], v[vnz-1] stay zero."""
    npml, pnx, pnz = 5, 20, 16
    coe = o64.fdcoe(2.0, 3.0)
    tab = o64.cpml_tables(npml, 2.0, 3.0, 1e-3)
    st = o64.new_pml_state(pnx, pnz, npml)
    rng = np.random.default_rng(4)
    for _ in range(5):
        p = rng.standard_normal((pnz, pnx))
        o64.op_pml(p, st, coe, tab, npml)
    assert np.all(st["u"][:, 0] == 0) and np.all(st["u"][:, pnx] == 0) and np.all(st["v"][pnz] == 0)
    # the inner-edge CPML nodes have a = 0, so their memory never moves
    assert np.all(st["u0"][:, npml] == 0) and np.all(st["u1"][:, 0] == 0) and np.all(st["v1"][0] == 0)
    # with zero tables the strips vanish: edge rows are plain 2nd-order
    st = o64.new_pml_state(pnx, pnz, npml)
    o64.op_pml(p, st, coe, np.zeros_like(tab), npml)
    np.testing.assert_allclose(st["u"][:, 1], (p[:, 1] - p[:, 0]) / 2.0, rtol=1e-14)
    np.testing.assert_allclose(st["u"][:, pnx - 1], (p[:, pnx - 1] - p[:, pnx - 2]) / 2.0, rtol=1e-14)
    np.testing.assert_allclose(st["v"][pnz - 1], (p[pnz - 1] - p[pnz - 2]) / 3.0, rtol=1e-14)


def test_cpml_memory_geometric_sum(o64):
    """Constant input x: psi_n = a x (1 - b^n)/(1 - b), output x + psi_n."""
    npml, pnx, pnz = 6, 24, 18
    dx = 5.0
    coe = o64.fdcoe(dx, dx)
    tab = o64.cpml_tables(npml, dx, dx, 2e-3)
    st = o64.new_pml_state(pnx, pnz, npml)
    ix = np.arange(pnx, dtype=np.float64)
    p = 0.7 * ix[None, :] * np.ones((pnz, 1))  # u_raw = 0.7*s/dx on all 4th-order nodes
    s = (coe[0] + 3 * coe[1]) * dx
    nsteps = 9
    for _ in range(nsteps):
        o64.op_pml(p, st, coe, tab, npml)
    k = 3  # a left-strip u node inside the 4th-order range
    a, b = tab[0, 2 * k], tab[1, 2 * k]
    x = 0.7 * s / dx
    psi = a * x * (1 - b ** nsteps) / (1 - b)
    np.testing.assert_allclose(st["u0"][:, k], psi, rtol=1e-12)
    np.testing.assert_allclose(st["u"][:, k], x + psi, rtol=1e-12)


def test_time_integral_is_leapfrog_cosine(o64):
    """p^{n+1} = 2p^n - p^{n-1} - k p^n has the solution cos(n theta),
    cos(theta) = 1 - k/2 (fwi_2d.c:878 with ekp*(pxx+pzz) = -k p)."""
    k = 0.3
    theta = math.acos(1 - k / 2)
    pold = np.array([[math.cos(-theta)]])
    pnow = np.array([[1.0]])
    ekp = np.array([[1.0]])
    for n in range(1, 200):
        pnew = o64.time_integral(pnow, pold, -k * pnow, 0 * pnow, ekp)
        pold, pnow = pnow, pnew
        assert abs(pnow[0, 0] - math.cos(n * theta)) < 1e-10


# --------------------------------------------------------------------- a4
def test_boundary_strips_round_trip(o64):
    nx, nz, npml = 11, 9, 3
    rng = np.random.default_rng(5)
    pf = rng.standard_normal((nz + npml, nx + 2 * npml))
    ud, lr = o64.write_bdr(pf, nx, nz, npml)
    b = pf[:nz, npml:npml + nx]
    for i in range(3):
        np.testing.assert_array_equal(ud[i], b[i])
        np.testing.assert_array_equal(ud[3 + i], b[nz - 1 - i])
        np.testing.assert_array_equal(lr[:, i], b[:, i])
        np.testing.assert_array_equal(lr[:, 3 + i], b[:, nx - 1 - i])
    out = o64.read_bdr(ud, lr, np.zeros((nz, nx)))
    mask = np.zeros((nz, nx), bool)
    mask[:3], mask[-3:], mask[:, :3], mask[:, -3:] = True, True, True, True
    np.testing.assert_array_equal(out[mask], b[mask])
    assert np.all(out[~mask] == 0)


# --------------------------------------------------------------------- a1 adjoint
def test_op_pml_adjoint_dot_product(o64):
    """<A_n(p_0..p_n), w_n> summed over steps equals the backward sweep of
    op_pml_adj: exact transpose including the CPML memory."""
    npml, pnx, pnz = 5, 23, 17
    coe = o64.fdcoe(7.0, 5.0)
    tab = o64.cpml_tables(npml, 7.0, 5.0, 1e-3)
    rng = np.random.default_rng(6)
    T = 6
    ps = [rng.standard_normal((pnz, pnx)) for _ in range(T)]
    ws = [rng.standard_normal((pnz, pnx)) for _ in range(T)]
    st = o64.new_pml_state(pnx, pnz, npml)
    lhs = 0.0
    for t in range(T):
        o64.op_pml(ps[t], st, coe, tab, npml)
        lhs += np.sum((st["ux"] + st["vz"]) * ws[t])
    sa = o64.new_adj_state(pnx, pnz, npml)
    rhs = 0.0
    for t in reversed(range(T)):
        rhs += np.sum(ps[t] * o64.op_pml_adj(ws[t], sa, coe, tab, npml))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


# --------------------------------------------------------------------- a10
def _small_case(seed=0, nt=240):
    g = Geometry(nx=48, nz=32, npml=8, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(seed)
    vel = (2000 + 250 * rng.standard_normal((g.nz, g.nx))).astype(np.float32)
    irx = np.arange(0, g.nx, 2, dtype=np.int32)
    shot = Shot(nt=nt, isx=21, isz=2, wav=ricker(nt, g.dt, 25.0), irx=irx,
                irz=np.full(len(irx), 2, np.int32))
    return g, vel, shot


def test_driver_dot_product(o64, o32):
    """<F f, d> = <f, F^T d>: forward (source -> traces) against backward
    (residual -> adjoint at the source), whole time loop."""
    g, vel, shot = _small_case()
    iarea = 1.0 / (g.dx * g.dz)
    d = np.random.default_rng(7).standard_normal((shot.nr, shot.nt))
    for o, tol in ((o64, 1e-12), (o32, 1e-5)):
        f = o.forward(g, vel, shot, want_strips=True)
        b = o.backward(g, vel, shot, d, f)
        lhs = np.sum(f["syn"].astype(np.float64) * d.astype(o.dtype))
        rhs = np.sum(b["musrc"].astype(np.float64) * (shot.wav.astype(o.dtype) * o.dtype.type(iarea)))
        scale = np.linalg.norm(f["syn"].astype(np.float64)) * np.linalg.norm(d)
        assert abs(lhs - rhs) <= tol * scale


def test_reconstruction_matches_stored_wavefield(o64):
    g, vel, shot = _small_case(seed=1)
    f = o64.forward(g, vel, shot, want_strips=True, want_xall=True)
    d = np.random.default_rng(8).standard_normal((shot.nr, shot.nt))
    b1 = o64.backward(g, vel, shot, d, f)
    b2 = o64.backward(g, vel, shot, d, f, use_xall=True)
    assert rel(b1["gk"], b2["gk"]) < 1e-11


def test_taylor_second_order(o64):
    """|J(v+h dv) - J(v) - h <g, dv>| = O(h^2) for a smooth dv that vanishes
    on the cells model_ext replicates into the PML."""
    g, vel, shot = _small_case(seed=2, nt=300)
    obs = o64.forward(g, vel, shot)["syn"]
    v0 = vel.astype(np.float64).copy()
    v0[8:24, 12:36] *= 1.04
    J0, _, grad, _ = o64.gradient(g, v0, shot, obs)
    zz, xx = np.mgrid[0:g.nz, 0:g.nx]
    dv = 80.0 * np.sin(np.pi * xx / (g.nx - 1)) * np.sin(np.pi * (zz + 1) / g.nz)
    dv[:, 0] = dv[:, -1] = dv[-1, :] = 0.0
    gd = float(np.sum(grad * dv))
    rems = []
    for h in (1e-1, 1e-2, 1e-3):
        Jh = o64.gradient(g, v0 + h * dv, shot, obs)[0]
        rems.append(abs(Jh - J0 - h * gd))
    assert rems[1] < rems[0] / 50 and rems[2] < rems[1] / 50


def test_fp32_vs_fp64_config1(o64, o32):
    g, vel, start, shots = config1()
    s = shots[0]
    obs = o64.forward(g, vel, s)["syn"]
    J64, syn64, gr64, _ = o64.gradient(g, start.astype(np.float64), s, obs)
    J32, syn32, gr32, _ = o32.gradient(g, start, s, obs.astype(np.float32))
    assert rel(syn32, syn64) < 1e-5
    assert rel(gr32, gr64) < 1e-4
    assert abs(J32 - J64) < 1e-5 * J64


def test_homogeneous_green_function(o64):
    """Far from boundaries the trace is c^2 (f * G2D), G2D = H(ct-r) /
    (2 pi c sqrt(c^2 t^2 - r^2)); checks source scaling (ekp * wav * iarea,
    fwi_2d.c:216) and the receiver time index (syn[it] = p^{it+1})."""
    c, dx, dt, nt = 2000.0, 5.0, 5e-4, 700
    g = Geometry(nx=240, nz=240, npml=30, dx=dx, dz=dx, dt=dt)
    vel = np.full((g.nz, g.nx), c, np.float32)
    f0 = 15.0
    wav = ricker(nt, dt, f0)
    shot = Shot(nt=nt, isx=120, isz=120, wav=wav, irx=np.array([180], np.int32), irz=np.array([120], np.int32))
    syn = o64.forward(g, vel, shot)["syn"][0]
    r = 60 * dx
    t = (np.arange(nt) + 1) * dt
    # G sampled finely and convolved with the wavelet (integrable singularity)
    sub = 40
    tf = np.arange(nt * sub) * (dt / sub)
    G = np.zeros_like(tf)
    m = c * tf > r * (1 + 1e-9)
    G[m] = 1.0 / (2 * np.pi * c * np.sqrt(c * c * tf[m] ** 2 - r * r))
    wf = np.interp(tf, np.arange(nt) * dt, wav)
    ana = c * c * np.convolve(wf, G)[: nt * sub] * (dt / sub)
    ana = np.interp(t, tf, ana)
    win = t < 0.2  # before anything returns from the PML
    assert rel(syn[win], ana[win]) < 0.05
