"""Host trace conditioning (csrc/fwi_sigproc.c): the Butterworth band-pass
(butterworth_filter.c:127-200) and the mutes/tapers (mute_fun.c:5-190).

Pinned three ways:
  * against the reference's own butterworth_filter.c / mute_fun.c, compiled
    from its sources under oracle/_ref (FP32 arithmetic there, FP64 here:
    agreement within the reference's own FP32 rounding, shown by measuring
    that rounding against an FP64 filter);
  * the Butterworth against scipy.signal.butter + sosfilt in FP64 (the same
    bilinear-transform Butterworth; zero phase = forward, reverse, forward,
    reverse with the pole count halved);
  * closed forms of the tapers at their breakpoints.
CPU only: these are host functions."""
import numpy as np
import pytest
import scipy.signal as ss

import host_ctypes as H


@pytest.fixture(scope="module")
def h():
    return H.lib()


@pytest.fixture(scope="module")
def ref():
    r = H.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref/libref_sigproc.so not built (reference sources absent)")
    return r


def scipy_bandpass(x, dt, phase, nphi, nplo, fhi, flo):
    y = x.astype(np.float64)
    lo, hi = flo * dt, fhi * dt
    nplo = max(nplo, 1)
    nphi = max(nphi, 1)
    if nplo > 1 and not phase:
        nplo //= 2
    if nphi > 1 and not phase:
        nphi //= 2

    def run(sos, y):
        y = ss.sosfilt(sos, y, axis=1)
        if not phase:
            y = ss.sosfilt(sos, y[:, ::-1], axis=1)[:, ::-1]
        return y

    if lo > 1e-4:
        y = run(ss.butter(nplo, 2 * lo, "high", output="sos"), y)
    if hi < 0.5 - 1e-4:
        y = run(ss.butter(nphi, 2 * hi, "low", output="sos"), y)
    return y


CASES = [(p, nplo, nphi, flo, fhi) for p in (True, False)
         for (nplo, nphi) in ((1, 1), (2, 2), (3, 4), (4, 6), (5, 3))
         for (flo, fhi) in ((3.0, 20.0), (0.0, 40.0), (8.0, 260.0), (1.0, 10.0))]


@pytest.mark.parametrize("phase,nplo,nphi,flo,fhi", CASES)
def test_bandpass_vs_scipy_fp64(h, phase, nplo, nphi, flo, fhi):
    rng = np.random.default_rng(11)
    nt, nr, dt = 1200, 4, 0.002
    x = rng.standard_normal((nr, nt)).astype(np.float32)
    out = np.zeros_like(x)
    assert h.sfbandpass(H.fp(x), nt, nr, dt, phase, nphi, nplo, fhi, flo, H.fp(out)) == 0
    want = scipy_bandpass(x, dt, phase, nphi, nplo, fhi, flo)
    err = np.abs(out - want).max() / np.abs(want).max()
    assert err < 1e-6, err   # FP64 filter, FP32 output rounding


@pytest.mark.parametrize("phase,nplo,nphi,flo,fhi", CASES)
def test_bandpass_vs_reference(h, ref, phase, nplo, nphi, flo, fhi):
    rng = np.random.default_rng(12)
    nt, nr, dt = 1200, 4, 0.002
    x = rng.standard_normal((nr, nt)).astype(np.float32)
    out = np.zeros_like(x)
    got_ref = np.zeros_like(x)
    assert h.sfbandpass(H.fp(x), nt, nr, dt, phase, nphi, nplo, fhi, flo, H.fp(out)) == 0
    ref.sfbandpass(H.fp(x), nt, nr, dt, phase, nphi, nplo, fhi, flo, H.fp(got_ref))
    m = np.abs(got_ref).max()
    err = np.abs(out - got_ref).max() / m
    # the reference's own FP32 recursion error, measured against an FP64 filter
    ref_own = np.abs(got_ref - scipy_bandpass(x, dt, phase, nphi, nplo, fhi, flo)).max() / m
    assert err < 1e-6 + 1.5 * ref_own, (err, ref_own)
    assert err < 1e-3, err


def test_bandpass_in_place_and_errors(h):
    rng = np.random.default_rng(13)
    x = rng.standard_normal((3, 500)).astype(np.float32)
    out = np.zeros_like(x)
    assert h.sfbandpass(H.fp(x), 500, 3, 0.004, False, 4, 4, 20.0, 2.0, H.fp(out)) == 0
    y = x.copy()
    assert h.sfbandpass(H.fp(y), 500, 3, 0.004, False, 4, 4, 20.0, 2.0, H.fp(y)) == 0
    assert np.array_equal(y, out)
    # negative flo: the reference prints and exits; here an error code
    assert h.sfbandpass(H.fp(x), 500, 3, 0.004, False, 4, 4, 20.0, -1.0, H.fp(out)) == 1
    assert b"flo" in h.fwi_sigproc_last_error()


def test_bandpass_passes_band_and_stops_outside(h):
    """A sine inside the band keeps its amplitude with no shift (zero phase);
    one outside is scaled by the squared Butterworth magnitude of the
    prewarped cut-off, 1 / (1 + (tan(pi f dt) / tan(pi fc dt))^(2n)), n = 8/2."""
    nt, dt = 4000, 0.001
    t = np.arange(nt) * dt
    inside = np.sin(2 * np.pi * 10 * t).astype(np.float32)[None, :].copy()
    outside = np.sin(2 * np.pi * 120 * t).astype(np.float32)[None, :].copy()
    a, b = np.zeros_like(inside), np.zeros_like(outside)
    assert h.sfbandpass(H.fp(inside), nt, 1, dt, False, 8, 8, 40.0, 2.0, H.fp(a)) == 0
    # high cut only: a 2 Hz low cut would ring for seconds after the sine switches on
    assert h.sfbandpass(H.fp(outside), nt, 1, dt, False, 8, 8, 40.0, 0.0, H.fp(b)) == 0
    mid = slice(1000, 3000)
    assert np.abs(a[0, mid] - inside[0, mid]).max() < 2e-3
    ratio = np.tan(np.pi * 120 * dt) / np.tan(np.pi * 40 * dt)
    gain = 1.0 / (1.0 + ratio ** 8)
    assert np.abs(b[0, mid]).max() == pytest.approx(gain, rel=0.02)


def _geometry(rng, nr):
    sr = np.zeros(2 + 2 * nr, np.float32)
    sr[0], sr[1] = 1000.0, 20.0
    sr[2::2] = rng.uniform(0, 4000, nr)
    sr[3::2] = 20.0
    return sr


TAPERS = ["linear_upper_off", "linear_lower_off", "quadratic_upper_off", "quadratic_lower_off"]


@pytest.mark.parametrize("name", TAPERS)
def test_cos_tapers_vs_reference(h, ref, name):
    rng = np.random.default_rng(21)
    nr, nt, dt = 12, 900, 0.002
    sr = _geometry(rng, nr)
    x = rng.standard_normal((nr, nt)).astype(np.float32)
    a, b = x.copy(), x.copy()
    assert getattr(h, name)(H.rows(a), H.fp(sr), nr, nt, dt, 0.1, 2000.0, 25) == 0
    getattr(ref, name)(H.rows(b), H.fp(sr), nr, nt, dt, 0.1, 2000.0, 25)
    assert np.abs(a - b).max() < 2e-5 * np.abs(x).max()
    # zeroed / untouched regions identical, ramp in between
    assert np.array_equal(a == 0, b == 0)


@pytest.mark.parametrize("upper", [True, False])
def test_linear_taper_breakpoints(h, upper):
    """At t1 = t0 + d/v0 the factor is 1, N samples away it reaches 0, and the
    other side is untouched."""
    nr, nt, dt, t0, v0, N = 1, 400, 0.01, 0.5, 1000.0, 20
    sr = np.array([0.0, 0.0, 1000.0, 0.0], np.float32)      # d = 1000 m: t1 = 1.5 s = sample 150
    x = np.ones((nr, nt), np.float32)
    fn = h.linear_upper_off if upper else h.linear_lower_off
    assert fn(H.rows(x), H.fp(sr), nr, nt, dt, t0, v0, N) == 0
    y = x[0]
    # samples 130/150/170 sit on the breakpoints, where FP32 dt rounding decides the side
    if upper:
        assert np.all(y[:130] == 0) and np.all(y[151:] == 1)
        k = np.arange(131, 150)
        assert np.allclose(y[k], 0.5 * (1 - np.cos(np.pi * (k - 130) / 20)), atol=1e-6)
    else:
        assert np.all(y[:151] == 1) and np.all(y[171:] == 0)
        k = np.arange(151, 170)
        assert np.allclose(y[k], 0.5 * (1 - np.cos(np.pi * (k - 170) / 20)), atol=1e-6)


@pytest.mark.parametrize("name", ["upper_off_srcood", "lower_off_srcood"])
def test_tanh_tapers_vs_reference(h, ref, name):
    rng = np.random.default_rng(22)
    nr, nt, dt = 10, 700, 0.002
    sr = _geometry(rng, nr)
    x = rng.standard_normal((nr, nt)).astype(np.float32)
    a, b = x.copy(), x.copy()
    dummy = np.zeros_like(x)
    assert getattr(h, name)(H.rows(a), H.fp(sr), nr, nt, dt, 0.05, 1800.0, 30.0, H.rows(dummy)) == 0
    getattr(ref, name)(H.rows(b), H.fp(sr), nr, nt, dt, 0.05, 1800.0, 30.0, H.rows(dummy))
    assert np.abs(a - b).max() < 1e-5 * np.abs(x).max()
    assert not dummy.any()   # the out argument is unused, as in the reference


def test_tanh_taper_midpoint(h):
    sr = np.array([0.0, 0.0, 0.0, 0.0], np.float32)        # d = 0: t1 = t0
    x = np.ones((1, 101), np.float32)
    y = x.copy()
    d = np.zeros_like(x)
    assert h.upper_off_srcood(H.rows(x), H.fp(sr), 1, 101, 0.01, 0.5, 1500.0, 5.0, H.rows(d)) == 0
    assert h.lower_off_srcood(H.rows(y), H.fp(sr), 1, 101, 0.01, 0.5, 1500.0, 5.0, H.rows(d)) == 0
    assert x[0, 50] == pytest.approx(0.5) and y[0, 50] == pytest.approx(1.0)   # lower: no 1/2 (sic)
    assert x[0, 0] < 0.01 and x[0, 100] > 0.99


@pytest.mark.parametrize("early_gone", [True, False])
def test_fbt_muter_vs_reference(h, ref, early_gone):
    rng = np.random.default_rng(23)
    nr, nt, dt = 9, 800, 0.002
    fbt = rng.uniform(0.2, 1.2, nr).astype(np.float32)
    x = rng.standard_normal((nr, nt)).astype(np.float32)
    a, b = x.copy(), x.copy()
    assert h.fbt_muter(H.fp(fbt), H.fp(a), nt, nr, 8.0, dt, 0.05, early_gone) == 0
    ref.fbt_muter(H.fp(fbt), H.fp(b), nt, nr, 8.0, dt, 0.05, early_gone)
    assert np.abs(a - b).max() < 1e-5 * np.abs(x).max()
    # closed form: samples on the kept side of fbt + tune are untouched
    t = np.arange(nt) * dt
    for r in range(nr):
        kept = t > fbt[r] + 0.05 + 1e-4 if early_gone else t < fbt[r] + 0.05 - 1e-4
        assert np.array_equal(a[r, kept], x[r, kept])


def test_win_apply_vs_reference(h, ref):
    rng = np.random.default_rng(24)
    nr, nt, dt, alpha = 7, 600, 0.002, 20.0
    win = np.zeros(2 * nr, np.float32)
    win[0::2] = rng.uniform(0.1, 0.4, nr)
    win[1::2] = win[0::2] + rng.uniform(0.2, 0.6, nr)
    x = rng.standard_normal((nr, nt)).astype(np.float32)
    a, b = np.zeros_like(x), np.zeros_like(x)
    assert h.win_apply(H.fp(win), H.fp(x), nt, nr, dt, alpha, H.fp(a)) == 0
    ref.win_apply(H.fp(win), H.fp(x), nt, nr, dt, alpha, H.fp(b))
    # the reference accumulates ss += dt in FP32 (mute_fun.c:37,47): its window
    # edges drift by about nt * ulp(t); the exact products here do not
    drift = nt * np.spacing(np.float32(nt * dt)) * alpha
    assert np.abs(a - b).max() <= (2 * drift + 1e-5) * np.abs(x).max()
    # inside the window untouched
    t = (np.arange(nt) + 1) * dt
    for r in range(nr):
        inside = (t > win[2 * r] + 1e-4) & (t < win[2 * r + 1] - 1e-4)
        assert np.array_equal(a[r, inside], x[r, inside])


@pytest.mark.parametrize("phase", [True, False])
@pytest.mark.parametrize("flo,fhi", [(3.0, 30.0), (0.0, 30.0), (3.0, 400.0)])
def test_bandpass_adjoint_dot_product(h, phase, flo, fhi):
    """<F x, y> = <x, F^T y> for the band-pass and sfbandpass_adjoint."""
    rng = np.random.default_rng(31)
    nt, nr, dt = 700, 3, 0.002
    x = rng.standard_normal((nr, nt)).astype(np.float32)
    y = rng.standard_normal((nr, nt)).astype(np.float32)
    fx, fty = np.zeros_like(x), np.zeros_like(y)
    assert h.sfbandpass(H.fp(x), nt, nr, dt, phase, 5, 4, fhi, flo, H.fp(fx)) == 0
    assert h.sfbandpass_adjoint(H.fp(y), nt, nr, dt, phase, 5, 4, fhi, flo, H.fp(fty)) == 0
    a = float((fx.astype(np.float64) * y).sum())
    b = float((x.astype(np.float64) * fty).sum())
    assert a == pytest.approx(b, rel=1e-5, abs=1e-5 * np.sqrt((fx.astype(np.float64) ** 2).sum() * nt * nr))
    if flo > 0 and fhi < 250:
        # not self-adjoint in general: the transpose differs from the filter itself
        fy = np.zeros_like(y)
        assert h.sfbandpass(H.fp(y), nt, nr, dt, phase, 5, 4, fhi, flo, H.fp(fy)) == 0
        assert not np.allclose(fy, fty, atol=1e-4)
