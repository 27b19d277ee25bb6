"""Host misfit pool (csrc/fwi_msf.c): L2norm, NCCnorm, XCORR1D
(xcorr_msf.c:4-181), MF1D (lsq_decon_cg.c:82-392) and entropic OT
(entropy_ot.h:18-30).

XCORR1D and MF1D need FFTW in the reference, which is absent, so they are
pinned by a direct-convolution numpy restatement (oracle/msf_oracle.py) that
is itself pinned here: its adjoint sources are finite-difference gradients of
its misfits (where the reference's are), and its MF1D filter and adjoint
source equal dense normal-equation solves.  NCC and entropic OT have no body
in the reference; they are pinned by finite differences and closed forms
(translation and small-eps limits of optimal transport).  CPU only."""
import ctypes

import numpy as np
import pytest

import host_ctypes as H
import msf_oracle as MO


@pytest.fixture(scope="module")
def h():
    return H.lib()


def wavelets(nr, nt, shift, scale=1.0, seed=0):
    rng = np.random.default_rng(seed)
    t = np.arange(nt)
    c = nt * 0.4 + rng.uniform(-3, 3, nr)
    obs = np.stack([np.exp(-((t - c[r]) / 4.0) ** 2) * np.cos(0.5 * (t - c[r])) for r in range(nr)])
    syn = np.stack([np.exp(-((t - c[r] - shift) / 4.0) ** 2) * np.cos(0.5 * (t - c[r] - shift)) for r in range(nr)])
    return (scale * obs).astype(np.float32), (scale * syn).astype(np.float32)


def call_xcorr(h, obs, syn, mode, sr=None, rmin=0.0, rmax=0.0, npad=7):
    nr, nt = obs.shape
    res = np.zeros_like(obs)
    mis = np.zeros(2, np.float32)
    par = H.ParXcorr1d(nr, nt, npad, mode.encode())
    rc = h.XCORR1D(H.rows(obs), H.rows(syn), H.fp(sr) if sr is not None else None, rmin, rmax, ctypes.byref(par),
                   H.fp(mis), H.rows(res))
    return rc, mis, res


def fd_grad(f, s, e):
    """central differences of the scalar f at every entry of s (FP64)"""
    g = np.zeros_like(s)
    for idx in np.ndindex(*s.shape):
        sp, sm = s.copy(), s.copy()
        sp[idx] += e
        sm[idx] -= e
        g[idx] = (f(sp) - f(sm)) / (2 * e)
    return g


# ---------------------------------------------------------------- L2 / NCC

def test_l2norm_closed_form(h):
    obs, syn = wavelets(4, 50, 2)
    res = np.zeros_like(obs)
    val = np.zeros(2, np.float32)
    assert h.L2norm(H.rows(obs), H.rows(syn), 50, 4, None, 0.0, 0.0, H.fp(val), H.rows(res)) == 0
    d = syn.astype(np.float64) - obs
    assert np.array_equal(res, (syn - obs))
    assert val[0] == val[1] == pytest.approx(0.5 * (d * d).sum(), rel=1e-6)


def test_offset_selection(h):
    """Traces whose source-receiver distance is outside [min, max] get res = 0
    and no misfit (xcorr_msf.c:69-78)."""
    nr, nt = 5, 60
    obs, syn = wavelets(nr, nt, 3)
    sr = np.zeros(2 + 2 * nr, np.float32)
    sr[0], sr[1] = 100.0, 0.0
    sr[2::2] = [100.0, 150.0, 400.0, 900.0, 2000.0]           # distances 0, 50, 300, 800, 1900
    keep = np.array([False, True, True, True, False])         # [40, 1000]
    for fn in ("L2norm", "NCCnorm"):
        res = np.zeros_like(obs)
        val = np.zeros(2, np.float32)
        assert getattr(h, fn)(H.rows(obs), H.rows(syn), nt, nr, H.fp(sr), 40.0, 1000.0, H.fp(val), H.rows(res)) == 0
        assert not res[~keep].any() and res[keep].any()
    rc, mis, res = call_xcorr(h, obs, syn, "middle", sr, 40.0, 1000.0)
    assert rc == 0 and not res[~keep].any() and res[keep].any()
    m_all, _ = MO.xcorr1d(obs, syn, "middle")
    m_sel, _ = MO.xcorr1d(obs, syn, "middle", sr, 40.0, 1000.0)
    assert mis[1] == pytest.approx(m_sel[1], rel=1e-5) and m_sel[1] < m_all[1]


def ncc_value(obs, syn):
    o, s = obs.astype(np.float64), syn.astype(np.float64)
    return sum(1 - (s[r] @ o[r]) / np.sqrt((s[r] @ s[r]) * (o[r] @ o[r])) for r in range(len(o)))


def test_ncc_closed_forms_and_gradient(h):
    nr, nt = 3, 40
    obs, syn = wavelets(nr, nt, 2.5, seed=3)
    res = np.zeros_like(obs)
    val = np.zeros(2, np.float32)
    assert h.NCCnorm(H.rows(obs), H.rows(syn), nt, nr, None, 0.0, 0.0, H.fp(val), H.rows(res)) == 0
    assert val[1] == pytest.approx(ncc_value(obs, syn), rel=1e-6)
    assert val[0] == pytest.approx(((syn.astype(np.float64) - obs) ** 2).sum(), rel=1e-6)
    g = fd_grad(lambda s: ncc_value(obs, s), syn.astype(np.float64), 1e-6)
    assert np.abs(res - g).max() < 1e-5 * np.abs(g).max()
    # identical traces: zero misfit; negated: 2 per trace; scale invariance
    for syn2, want in ((obs, 0.0), (-obs, 2.0 * nr), (3.0 * obs, 0.0)):
        syn2 = np.ascontiguousarray(syn2, np.float32)
        assert h.NCCnorm(H.rows(obs), H.rows(syn2), nt, nr, None, 0.0, 0.0, H.fp(val), H.rows(res)) == 0
        assert val[1] == pytest.approx(want, abs=1e-6)
    # a trace with no usable waveform (energy below 1e-8 of the largest):
    # skipped, instead of an adjoint source of order 1/|s|
    weak = syn.copy()
    weak[1] *= 1e-5
    assert h.NCCnorm(H.rows(obs), H.rows(weak), nt, nr, None, 0.0, 0.0, H.fp(val), H.rows(res)) == 0
    assert not res[1].any() and res[0].any()
    assert val[1] == pytest.approx(ncc_value(obs[[0, 2]], weak[[0, 2]]), rel=1e-6)
    # zero synthetic: skipped, no NaN
    z = np.zeros_like(obs)
    assert h.NCCnorm(H.rows(obs), H.rows(z), nt, nr, None, 0.0, 0.0, H.fp(val), H.rows(res)) == 0
    assert val[1] == 0.0 and not res.any()


# ---------------------------------------------------------------- XCORR1D

@pytest.mark.parametrize("nt", [63, 64])
@pytest.mark.parametrize("mode", ["middle", "leading"])
def test_xcorr1d_vs_restatement(h, nt, mode):
    obs, syn = wavelets(4, nt, 3.0, seed=nt)
    rc, mis, res = call_xcorr(h, obs, syn, mode)
    assert rc == 0
    m, ro = MO.xcorr1d(obs, syn, mode)
    assert mis[0] == pytest.approx(m[0], rel=1e-6) and mis[1] == pytest.approx(m[1], rel=1e-6)
    assert np.abs(res - ro).max() < 1e-6 * np.abs(ro).max()
    # the FFT length does not matter
    rc2, mis2, res2 = call_xcorr(h, obs, syn, mode, npad=0)
    assert rc2 == 0 and np.array_equal(res2, res) and np.array_equal(mis2, mis)


@pytest.mark.parametrize("nt", [63, 64])
def test_xcorr1d_adjoint_source_is_the_gradient(nt):
    """Middle slicing, odd nt: res = d misfit[1] / d syn.  Even nt: the
    reference's window is centred one lag off, so res[i] = d/d syn[i+1]."""
    obs, syn = wavelets(2, nt, 3.0, seed=7)
    s = syn.astype(np.float64)
    _, res = MO.xcorr1d(obs, s, "middle")
    g = fd_grad(lambda x: MO.xcorr1d(obs, x, "middle")[0][1], s, 1e-6)
    if nt % 2:
        assert np.abs(res - g).max() < 1e-7 * np.abs(g).max()
    else:
        assert np.abs(res[:, :-1] - g[:, 1:]).max() < 1e-7 * np.abs(g).max()
        assert np.abs(res - g).max() > 0.1 * np.abs(g).max()


def test_xcorr1d_penalty_closed_form():
    """Middle slicing, odd nt: x is the correlation at lags -nt/2..nt/2, so a
    synthetic equal to the observed trace delayed by k samples gives a
    correlation peak at |lag| = k."""
    nt = 101
    obs, _ = wavelets(1, nt, 0.0, seed=2)
    for k in (0, 4, 9):
        syn = np.roll(obs, k, axis=1)
        x = np.convolve(obs[0, ::-1].astype(np.float64), syn[0])[nt // 2: nt // 2 + nt]
        assert abs(int(np.argmax(x)) - nt // 2) == k
        m, _ = MO.xcorr1d(obs, syn, "middle")
        T = np.abs(np.arange(nt) - nt // 2)
        assert m[1] == pytest.approx(((T * x) ** 2).sum() / (x * x).sum())


def test_xcorr1d_skips_and_errors(h):
    nr, nt = 3, 40
    obs, syn = wavelets(nr, nt, 2.0)
    syn[1] *= 20.0                                   # L1 ratio > 10: skipped (xcorr_msf.c:77)
    syn[2] = 0.0
    obs[2] = 0.0                                     # both empty: skipped here (0/0 in the reference)
    rc, mis, res = call_xcorr(h, obs, syn, "middle")
    assert rc == 0 and np.isfinite(mis).all() and not res[1:].any() and res[0].any()
    assert call_xcorr(h, obs, syn, "sideways")[0] == 1
    assert b"conv_how_to_slice" in h.fwi_msf_last_error()
    assert call_xcorr(h, obs, syn, "middle", npad=-1)[0] == 1


# ---------------------------------------------------------------- MF1D

def call_mf1d(h, obs, syn, nm, niter, mode, reg, tor0=0.0, eps0=0.5, Td=1.0):
    nr, nt = obs.shape
    res = np.zeros_like(obs)
    mis = np.zeros(2, np.float32)
    par = H.Epsdf1(nr, nt, nm, 5, 3, niter, mode.encode(), reg.encode(), tor0, eps0, Td)
    rc = h.MF1D(H.rows(obs), H.rows(syn), ctypes.byref(par), H.fp(mis), H.rows(res))
    return rc, mis, res


MF_CASES = [(m, r) for m in ("middle", "leading") for r in ("diff", "ide")]


@pytest.mark.parametrize("mode,reg", MF_CASES)
def test_mf1d_vs_restatement(h, mode, reg):
    nt, nm = 48, 21
    obs, syn = wavelets(3, nt, 3.0, scale=0.01, seed=5)
    rc, mis, res = call_mf1d(h, obs, syn, nm, nm, mode, reg)
    assert rc == 0
    m, ro = MO.mf1d(obs, syn, nm, nm, mode, reg, 0.0, 0.5, 1.0)
    assert mis[0] == pytest.approx(m[0], rel=1e-6) and mis[1] == pytest.approx(m[1], rel=1e-5)
    assert np.abs(res - ro).max() < 1e-5 * np.abs(ro).max()
    # the reference's stopping rule (|rnorm1 - rnorm0| < tor0) stops both the same way
    rc, mis, res = call_mf1d(h, obs, syn, nm, 200, mode, reg, tor0=1e-7)
    m, ro = MO.mf1d(obs, syn, nm, 200, mode, reg, 1e-7, 0.5, 1.0)
    assert rc == 0 and mis[1] == pytest.approx(m[1], rel=1e-5)
    assert np.abs(res - ro).max() < 1e-4 * np.abs(ro).max()


@pytest.mark.parametrize("mode,reg", MF_CASES)
def test_mf1d_restatement_is_the_dense_solution_and_gradient(mode, reg):
    """Converged (niter = nmodel CG steps): the filter w solves
    (A^T A + D^T D) w = A^T b, the adjoint source is A (A^T A + D^T D)^-1 c,
    and 100 x that is the finite-difference gradient of the penalty in syn."""
    nt, nm = 48, 21
    obs, syn = wavelets(1, nt, 3.0, scale=0.01, seed=6)
    obs, syn = obs.astype(np.float64), syn.astype(np.float64)
    m, ro, ws = MO.mf1d(obs, syn, nm, nm, mode, reg, 0.0, 0.5, 1.0, return_w=True)
    Am, Dm = MO.mf_dense(100 * obs[0], nt, nm, mode == "middle", reg == "diff", 0.5)
    N = Am.T @ Am + Dm.T @ Dm
    T = (np.abs(np.arange(nm) - nm // 2) if mode == "middle" else np.arange(nm)).astype(np.float64)

    def penalty(s):
        w = np.linalg.solve(N, Am.T @ (100 * s))
        return (T * w) @ (T * w) / (w @ w), w

    al, w = penalty(syn[0])
    assert np.abs(ws[0] - w).max() < 1e-7 * np.abs(w).max()
    assert m[1] == pytest.approx(al, rel=1e-8)
    c = 2 * (T * T * w - al * w) / (w @ w)
    y0 = Am @ np.linalg.solve(N, c)
    assert np.abs(ro[0] - y0).max() < 1e-7 * np.abs(y0).max()
    g = fd_grad(lambda s: penalty(s)[0], syn[0], 1e-8)
    assert np.abs(100 * y0 - g).max() < 1e-6 * np.abs(g).max()


def test_mf1d_errors(h):
    obs, syn = wavelets(1, 30, 2.0)
    assert call_mf1d(h, obs, syn, 10, 10, "middle", "diff")[0] == 1      # even nmodel, middle
    assert call_mf1d(h, obs, syn, 11, 10, "middle", "tv")[0] == 1
    assert call_mf1d(h, obs, syn, 11, 10, "centre", "diff")[0] == 1
    assert call_mf1d(h, obs, syn, 11, 10, "leading", "ide")[0] == 0


# ---------------------------------------------------------------- entropic OT

def eot_par(h, nt, eps, dt, max_iter=5000, tor=1e-12, nr=1):
    par = H.EotPar(nt, nr, max_iter, None, None, None, eps, tor, dt)
    assert h.eot_msf_init(ctypes.byref(par)) == 0
    return par


def call_eot(h, p, q, par, log=False):
    msf = np.zeros(2, np.float32)
    res = np.zeros_like(p)
    if log:
        exd = H.EotLogPar()
        assert h.eot_log_msf_init(ctypes.byref(par), ctypes.byref(exd)) == 0
        rc = h.entropy_ot_log(H.fp(p), H.fp(q), ctypes.byref(par), ctypes.byref(exd), H.fp(msf), H.fp(res))
    else:
        rc = h.entropy_ot(H.fp(p), H.fp(q), ctypes.byref(par), H.fp(msf), H.fp(res))
    return rc, float(msf[0]), res


def bump(nt, c, w, floor=1e-9):
    t = np.arange(nt)
    return (np.exp(-((t - c) / w) ** 2) + floor).astype(np.float32)


def w2_exact(p, q, dt):
    """1-D optimal transport with cost ((i-j) dt)^2: the monotone coupling of
    the normalised histograms (north-west corner on the CDFs)."""
    p = p.astype(np.float64) / p.sum()
    q = q.astype(np.float64) / q.sum()
    i = j = 0
    a, b = p[0], q[0]
    cost = 0.0
    while i < len(p) and j < len(q):
        m = min(a, b)
        cost += m * ((i - j) * dt) ** 2
        a -= m
        b -= m
        if a <= 1e-18:
            i += 1
            a = p[i] if i < len(p) else 0.0
        if b <= 1e-18:
            j += 1
            b = q[j] if j < len(q) else 0.0
    return cost


def test_eot_translation_identity(h):
    """Quadratic cost: W(p, p shifted by k samples) = W(p, p) + (k dt)^2 (the
    coupling shifts with the trace; its mean displacement is zero)."""
    nt, dt, eps = 200, 0.004, 4e-5
    par = eot_par(h, nt, eps, dt)
    p = bump(nt, 80, 6)
    rc, w0, _ = call_eot(h, p, p, par)
    assert rc == 0
    for k in (5, 17):
        rc, wk, _ = call_eot(h, p, bump(nt, 80 + k, 6), par)
        assert rc == 0 and wk - w0 == pytest.approx((k * dt) ** 2, rel=2e-4)
    h.eot_msf_destroy(ctypes.byref(par))


def test_eot_small_eps_limit_and_forms_agree(h):
    """eps -> 0: W -> the exact 1-D transport cost, within the entropic bound
    eps (1 + 2 log nt) (the scaling form underflows there and falls back to
    the log form); at moderate eps the scaling and log forms agree."""
    nt, dt = 120, 0.004
    p = bump(nt, 40, 5) + 0.5 * bump(nt, 70, 3)
    q = bump(nt, 60, 8)
    exact = w2_exact(p, q, dt)
    for eps in (2e-5, 5e-6):
        par = eot_par(h, nt, eps, dt)
        rc, w, g = call_eot(h, p, q, par)
        rcl, wl, gl = call_eot(h, p, q, par, log=True)
        assert rc == 0 and rcl == 0
        assert abs(wl - exact) <= eps * (1 + 2 * np.log(nt)), (wl, exact)
        assert w == pytest.approx(wl, rel=1e-6) and np.array_equal(g, gl)
        h.eot_msf_destroy(ctypes.byref(par))
    for eps in (1e-3, 2e-4):
        par = eot_par(h, nt, eps, dt)
        rc, w, g = call_eot(h, p, q, par)
        rcl, wl, gl = call_eot(h, p, q, par, log=True)
        assert rc == 0 and rcl == 0
        assert wl == pytest.approx(w, rel=1e-5)
        assert np.abs(gl - g).max() < 1e-4 * np.abs(g).max()
        h.eot_msf_destroy(ctypes.byref(par))
    # symmetric in its arguments
    par = eot_par(h, nt, 2e-5, dt)
    assert call_eot(h, q, p, par)[1] == pytest.approx(call_eot(h, p, q, par)[1], rel=1e-5)
    h.eot_msf_destroy(ctypes.byref(par))


def test_eot_gradient(h):
    """res = dW/dp through the normalisation, checked along random directions
    (FP32 outputs: directional differences with a large step)."""
    nt, dt, eps = 80, 0.004, 1e-4
    par = eot_par(h, nt, eps, dt)
    p = bump(nt, 30, 5, floor=1e-2)
    q = bump(nt, 45, 7, floor=1e-2)
    rc, w, g = call_eot(h, p, q, par)
    assert rc == 0
    rng = np.random.default_rng(9)
    for _ in range(3):
        d = (rng.standard_normal(nt) * 0.05 * p).astype(np.float32)    # relative: p stays positive
        e = 0.1
        rp, wp, _ = call_eot(h, (p + e * d).astype(np.float32), q, par)
        rm, wm, _ = call_eot(h, (p - e * d).astype(np.float32), q, par)
        assert rp == 0 and rm == 0
        fd = (wp - wm) / (2 * e)
        assert fd == pytest.approx(float(g @ d), rel=2e-3, abs=1e-9)
    h.eot_msf_destroy(ctypes.byref(par))


def test_eot_trace_by_trace(h):
    """p = syn^2 + eta, q = obs^2 + eta per trace (eta from obs only): misfit
    is the sum over traces and res = 2 syn dW/dp."""
    nr, nt, dt, eps = 3, 70, 0.004, 1e-4
    obs, syn = wavelets(nr, nt, 4.0, seed=4)
    par = eot_par(h, nt, eps, dt, nr=nr)
    msf = np.zeros(2, np.float32)
    res = np.zeros_like(obs)
    assert h.entropy_ot_2d_trace_by_trace(H.rows(obs), H.rows(syn), ctypes.byref(par), H.fp(msf), H.rows(res)) == 0
    total = 0.0
    for r in range(nr):
        eta = 1e-3 * float((obs[r].astype(np.float64) ** 2).max())
        p = (syn[r].astype(np.float64) ** 2 + eta).astype(np.float32)
        q = (obs[r].astype(np.float64) ** 2 + eta).astype(np.float32)
        rc, w, g = call_eot(h, p, q, par)
        assert rc == 0
        total += w
        assert np.abs(res[r] - 2 * syn[r] * g).max() < 1e-4 * np.abs(res[r]).max()
    assert msf[1] == pytest.approx(total, rel=1e-5)
    assert msf[0] == pytest.approx(((syn.astype(np.float64) - obs) ** 2).sum(), rel=1e-5)
    # directional derivative in syn
    rng = np.random.default_rng(10)
    d = (rng.standard_normal((nr, nt)) * 0.02).astype(np.float32)
    e = 0.05
    vals = []
    for sgn in (1, -1):
        s2 = np.ascontiguousarray(syn + sgn * e * d, np.float32)
        r2 = np.zeros_like(obs)
        assert h.entropy_ot_2d_trace_by_trace(H.rows(obs), H.rows(s2), ctypes.byref(par), H.fp(msf),
                                              H.rows(r2)) == 0
        vals.append(float(msf[1]))
    fd = (vals[0] - vals[1]) / (2 * e)
    assert fd == pytest.approx(float((res * d).sum()), rel=5e-3)
    h.eot_msf_destroy(ctypes.byref(par))


def test_eot_errors(h):
    nt = 30
    par = eot_par(h, nt, 1e-4, 0.004)
    p = bump(nt, 10, 3)
    q = p.copy()
    q[5] = -1.0
    assert call_eot(h, p, q, par)[0] == 1                      # negative mass
    par.eps = 0.0
    assert call_eot(h, p, p, par)[0] == 1
    h.eot_msf_destroy(ctypes.byref(par))
