"""Dynamic warping by local cross-correlation (lcc_diw.c, lcc_dtw.h) and soft
ranking (soft_ranking.c, soft_ranking.h) in the host misfit pool
(csrc/fwi_msf_rank.c).  CPU only.

lcc_diw.c includes header.h, which needs GSL and FFTW headers absent here, so
the warping is pinned by mathematics on tiny inputs:
  * the accumulations equal the best sums over every lag path of the
    constraint, enumerated by brute force (3-point: |step| <= 1; 5-point: no
    two consecutive lag changes), forwards and backwards;
  * each backtracked path obeys its constraint and its summed similarity is
    the optimum;
  * the correlation cube obeys Cauchy-Schwarz, scale invariance, odd
    symmetry in obs, and peaks at 1 on the lag of an exactly shifted copy;
  * the misfit recovers a constant integer shift exactly (zero residual,
    1/2 k^2 per sample) and its residual is the central difference of
    misfit[1] with respect to syn.
soft_ranking.c builds on its own: its argsort and isotonic regression are
compared with the reference compiled under oracle/_ref (when present), with
numpy's stable argsort and with scikit-learn's isotonic regression.  The soft
ranks (declared only in the reference) are pinned as the projection onto the
permutahedron: membership (Rado's majorisation) and the variational
inequality against all n! vertices, the reg -> 0 / reg -> inf limits, and
the residual as the transposed Jacobian by finite differences."""
import ctypes
import itertools

import numpy as np
import pytest

import host_ctypes as H
from fwihip import lib as L

BAR = -1e30


@pytest.fixture(scope="module")
def h():
    return H.lib()


def rows(a):
    return H.rows(a)


def cube(a):
    """float *** over a C-contiguous float32 [nr, nt, m] array"""
    keep = [rows(a[i]) for i in range(a.shape[0])]
    c = (H.PPF * a.shape[0])(*[ctypes.cast(k, H.PPF) for k in keep])
    return c, keep


# ------------------------------------------------------------ brute force

def valid(lags, five):
    d = np.diff(lags)
    if np.any(np.abs(d) > 1):
        return False
    if five and np.any((d[1:] != 0) & (d[:-1] != 0)):
        return False
    return True


def brute(err, nl, five):
    """best prefix (ending at i, j) and suffix (starting at i, j) sums over every valid lag path"""
    nt = err.shape[0]
    fw = np.full((nt, nl + 2), -np.inf)
    bw = np.full((nt, nl + 2), -np.inf)
    best = -np.inf
    for start in range(1, nl + 1):
        for steps in itertools.product((-1, 0, 1), repeat=nt - 1):
            lags = start + np.concatenate(([0], np.cumsum(steps)))
            if lags.min() < 1 or lags.max() > nl or not valid(lags, five):
                continue
            vals = err[np.arange(nt), lags].astype(np.float64)
            pre, suf = np.cumsum(vals), np.cumsum(vals[::-1])[::-1]
            for i in range(nt):
                fw[i, lags[i]] = max(fw[i, lags[i]], pre[i])
                bw[i, lags[i]] = max(bw[i, lags[i]], suf[i])
            best = max(best, pre[-1])
    return fw, bw, best


def rand_err(nt, nl, seed):
    e = np.random.default_rng(seed).uniform(-1, 1, (nt, nl + 2)).astype(np.float32)
    e[:, 0] = e[:, -1] = BAR
    return e


@pytest.mark.parametrize("five", [False, True])
@pytest.mark.parametrize("nt,nl,seed", [(6, 5, 1), (7, 3, 2), (2, 3, 3), (5, 1, 4)])
def test_accumulations_are_the_best_path_sums(h, five, nt, nl, seed):
    err = rand_err(nt, nl, seed)
    fw_ref, bw_ref, best = brute(err, nl, five)
    for name, ref in (("fw", fw_ref), ("bw", bw_ref)):
        d = np.zeros_like(err)
        getattr(h, f"ac_{'5' if five else '3'}pts_{name}")(rows(err), rows(d), nl, nt)
        np.testing.assert_allclose(d[:, 1:nl + 1], ref[:, 1:nl + 1], rtol=0, atol=1e-5)
        assert np.all(d[:, 0] == np.float32(BAR)) and np.all(d[:, nl + 1] == np.float32(BAR))
    assert abs(fw_ref[nt - 1].max() - best) < 1e-9 and abs(bw_ref[0].max() - best) < 1e-9


def cells(path, fw):
    """the similarity cells a backtracked path visits: a half lag sits in the
    middle of a diagonal step, on the column of its far end"""
    c = np.array(path, np.float64)
    out = c.copy()
    for i in range(len(c)):
        if c[i] != np.floor(c[i]):
            out[i] = c[i + 1] if fw else c[i - 1]
    return out.astype(int)


@pytest.mark.parametrize("seed", range(6))
def test_backtracked_paths_are_optimal(h, seed):
    nt, nl = 9, 5
    err = rand_err(nt, nl, 100 + seed)
    for five, acc, bt, fw in ((False, "ac_3pts_bw", "bt_fw_3pts", True), (True, "ac_5pts_bw", "bt_fw_5pts", True),
                              (True, "ac_5pts_fw", "bt_bw_5pts", False)):
        d = np.zeros_like(err)
        getattr(h, acc)(rows(err), rows(d), nl, nt)
        path = np.zeros(nt, np.float32)
        if bt == "bt_fw_3pts":
            h.bt_fw_3pts(rows(d), nl, nt, H.fp(path))
        else:
            getattr(h, bt)(rows(d), rows(err), nl, nt, H.fp(path))
        steps = np.diff(path)
        assert set(np.round(steps * 2).astype(int)) <= {-2, -1, 0, 1, 2}
        c = cells(path, fw)
        assert c.min() >= 1 and c.max() <= nl and valid(c, five), (bt, path)
        total = float(err[np.arange(nt), c].astype(np.float64).sum())
        _, _, best = brute(err, nl, five)
        assert abs(total - best) < 1e-5, (bt, path, total, best)
        # half lags only in the middle of a two-sample diagonal
        half = np.nonzero(path != np.floor(path))[0]
        for i in half:
            assert 0 < i < nt - 1 and path[i - 1] != path[i + 1]


def test_gaussian_window(h):
    for n, delta in ((1, 1.0), (7, 2.0), (31, 6.5)):
        w = np.zeros(n, np.float32)
        h.gaussian_win(n, delta, H.fp(w))
        i = np.arange(n) - n // 2
        np.testing.assert_allclose(w, np.exp(-0.5 * (i / delta) ** 2), rtol=1e-6)
        assert w[n // 2] == 1.0 and np.allclose(w, w[::-1])


def pulse(nt, c, width=5.0, f=0.35):
    t = np.arange(nt)
    return np.exp(-((t - c) / width) ** 2) * np.cos(f * (t - c))


def call_lcc(h, obs, syn, nl, nw, alpha, eps, sr=None, rng=None):
    nr, nt = obs.shape
    err = np.full((nr, nt, nl + 2), 7.0, np.float32)
    c, keep = cube(err)
    rngv = np.array(rng if rng is not None else [0, 0], np.float32)
    rc = h.lcc_err(rows(obs), rows(syn), H.fp(sr) if sr is not None else None, H.fp(rngv) if rng is not None else None,
                   eps, nt, nr, nl, nw, alpha, c)
    return rc, err


def test_correlation_cube_properties(h):
    nt, nl, nw, k = 160, 9, 21, 3
    syn = np.stack([pulse(nt, 70), 0.5 * pulse(nt, 80, 7.0)]).astype(np.float32)
    obs = np.roll(syn, k, axis=1)                      # obs(t) = syn(t - k): a delay of k samples
    rc, err = call_lcc(h, obs, syn, nl, nw, 6.0, 1e-3)
    assert rc == 0
    assert np.all(err[:, :, 0] == np.float32(BAR)) and np.all(err[:, :, nl + 1] == np.float32(BAR))
    e = err[:, :, 1:nl + 1]
    assert np.all(np.abs(e) <= 1 + 1e-5)                                # Cauchy-Schwarz
    for r in range(2):
        live = np.nonzero(syn[r] > 1e-3)[0]
        it0, it1 = live[0], live[-1]
        assert np.all(e[r, :it0] == 0) and np.all(e[r, it1:] == 0)      # the span [it0, it1)
        span = e[r, it0:it1]
        assert np.allclose(span[:, k + nl // 2], 1.0, atol=1e-5)       # lag il = 1 + nl/2 + k
        assert np.all(np.argmax(span, axis=1) == k + nl // 2)
    # scale invariance and odd symmetry in obs
    _, e2 = call_lcc(h, 3 * obs, (0.25 * syn).astype(np.float32), nl, nw, 6.0, 0.25e-3)
    _, e3 = call_lcc(h, -obs, syn, nl, nw, 6.0, 1e-3)
    np.testing.assert_allclose(e2, err, atol=2e-6)
    np.testing.assert_allclose(e3[:, :, 1:nl + 1], -e, atol=2e-6)
    # one entry by definition: tapered windows, zero outside the trace
    r, it, il = 0, 68, 2
    w = np.exp(-0.5 * ((np.arange(nw) - nw // 2) / 6.0) ** 2)
    ts = it - nw // 2 + np.arange(nw)
    to = ts + il - 1 - nl // 2
    x = np.where((ts >= 0) & (ts < nt), syn[r, np.clip(ts, 0, nt - 1)], 0) * w
    y = np.where((to >= 0) & (to < nt), obs[r, np.clip(to, 0, nt - 1)], 0) * w
    assert abs(err[r, it, il] - x @ y / (np.sqrt((x @ x) * (y @ y)) + 1e-12)) < 1e-6
    assert call_lcc(h, obs, syn, 8, nw, 6.0, 1e-3)[0] == L.EINVAL
    assert call_lcc(h, obs, syn, nl, 20, 6.0, 1e-3)[0] == L.EINVAL


def call_diw(h, obs, syn, nl=11, nw=21, alpha=6.0, eps=1e-3, sim=0.9, lag=100.0, sr=None, rng=None):
    nr, nt = obs.shape
    res = np.full_like(obs, 5.0)
    mis = np.zeros(2, np.float32)
    rngv = np.array(rng if rng is not None else [0, 1e9], np.float32)
    rc = h.lcc_diw_msf(rows(obs), rows(syn), H.fp(sr) if sr is not None else None, H.fp(rngv), eps, nt, nr, nl, nw,
                       alpha, sim, lag, H.fp(mis), rows(res))
    return rc, mis, res


@pytest.mark.parametrize("k", [-4, -1, 0, 2, 5])
def test_warping_recovers_an_integer_shift(h, k):
    nt = 200
    syn = np.stack([pulse(nt, 90), pulse(nt, 100, 7.0, 0.25)]).astype(np.float32)
    obs = np.roll(syn, k, axis=1)
    rc, mis, res = call_diw(h, obs, syn)
    assert rc == 0
    spans = 0
    for r in range(2):
        live = np.nonzero(syn[r] > 1e-3)[0]
        spans += live[-1] - live[0]
        # obs(t + k) = syn(t) on every counted sample but the span's last: past
        # the span every lag ties at 0, the backtracking starts from the first
        # column and its diagonal back to lag k may take that sample (half lag)
        assert np.all(np.delete(res[r], live[-1] - 1) == 0)
    assert mis[1] == pytest.approx(0.5 * float(np.sum(res.astype(np.float64) ** 2)), rel=1e-6)
    if k:   # 1/2 k^2 per counted sample, 1/2 (|k| +- 1/2)^2 on the half lags
        assert 0.5 * k ** 2 * (0.8 * spans - 2) <= mis[0] <= 0.5 * (abs(k) + 0.5) ** 2 * spans
    else:   # at most the two half lags
        assert mis[0] <= 0.5 * 0.25 * 2


def warped_case(nt=150):
    t = np.arange(nt)
    syn = 0.5 * (pulse(nt, 60, 6.0, 0.3) - 0.7 * pulse(nt, 95, 5.0, 0.4))
    tau = 2.3 + 1.5 * np.sin(t / 40.0)                # a smooth, fractional warp
    obs = 0.5 * (pulse(nt, 60 + np.interp(60, t, tau), 6.0, 0.3) - 0.7 * pulse(nt, 95 + np.interp(95, t, tau), 5.0, 0.4))
    return obs[None].astype(np.float32), syn[None].astype(np.float32)


def test_warping_residual_is_the_gradient(h):
    obs, syn = warped_case()
    e = 2.0 ** -10
    eps = 0.01
    assert np.all(np.abs(syn - eps) > 2 * e)          # the span [it0, it1) stays put under the steps
    kw = dict(nl=11, nw=25, alpha=6.0, eps=eps, sim=0.3, lag=100.0)
    rc, mis, res = call_diw(h, obs, syn, **kw)
    assert rc == 0 and mis[1] > 0
    fd = np.zeros(syn.shape[1])
    for t in range(syn.shape[1]):
        sp, sm = syn.copy(), syn.copy()
        sp[0, t] += e
        sm[0, t] -= e
        fd[t] = (call_diw(h, obs, sp, **kw)[1][1].astype(np.float64) -
                 call_diw(h, obs, sm, **kw)[1][1].astype(np.float64)) / (2 * e)
    err = np.abs(fd - res[0])
    scale = np.abs(res).max()
    print(f"max |res| {scale:.3e}, max |fd - res| {err.max():.3e}, median {np.median(err):.1e}, "
          f"{np.mean(err < 1e-3 * scale):.3f} within 1e-3")
    # misfit[1] is quadratic in syn with the path held; a step may still move
    # the path (or the similarity test) on a few samples
    assert np.mean(err < 1e-3 * scale) > 0.95


def test_warping_barriers_offsets_and_errors(h):
    obs, syn = warped_case()
    obs2 = np.concatenate([obs, obs]).astype(np.float32)
    syn2 = np.concatenate([syn, syn]).astype(np.float32)
    rc, m_all, res_all = call_diw(h, obs2, syn2, sim=-2.0)
    assert rc == 0 and np.array_equal(res_all[0], res_all[1])
    rc, m_none, res_none = call_diw(h, obs2, syn2, sim=2.0)           # no similarity reaches 2
    assert rc == 0 and np.all(m_none == 0) and np.all(res_none == 0)
    rc, m_lag0, res_lag0 = call_diw(h, obs2, syn2, sim=-2.0, lag=0.0)  # only zero lags count
    assert rc == 0 and m_lag0[0] == 0
    sr = np.array([0, 0, 100, 0, 300, 0], np.float32)                # offsets 100 and 300
    rc, m_one, res_one = call_diw(h, obs2, syn2, sim=-2.0, sr=sr, rng=[50, 200])
    assert rc == 0 and np.all(res_one[1] == 0) and np.array_equal(res_one[0], res_all[0])
    assert m_one[1] == pytest.approx(m_all[1] / 2, rel=1e-6)
    assert call_diw(h, obs, syn, nl=10)[0] == L.EINVAL
    assert "odd" in h.fwi_msf_rank_last_error().decode()


# ------------------------------------------------------------ soft ranking

def argsort(h, a, order):
    n = len(a)
    perm, rank = np.zeros(n, np.int32), np.zeros(n, np.int32)
    srt = np.zeros(n, np.float32)
    rc = h.sorting_flt_argsort(n, H.fp(a), order, perm.ctypes.data_as(H.PI), rank.ctypes.data_as(H.PI), H.fp(srt))
    return rc, perm, rank, srt


def test_argsort(h):
    ref = H.ref_rank_lib()
    rng = np.random.default_rng(0)
    for n in (1, 2, 9, 64):
        a = rng.standard_normal(n).astype(np.float32)
        for order, key in ((0, a), (1, -a)):
            rc, perm, rank, srt = argsort(h, a, order)
            p = np.argsort(key, kind="stable")
            assert rc == 0 and np.array_equal(perm, p + 1) and np.array_equal(srt, a[p])
            assert np.array_equal(rank[p], np.arange(1, n + 1))
            if ref is not None:                       # distinct values: the same order as the reference
                _, rp, rr, rs = (None, *[np.zeros(n, t) for t in (np.int32, np.int32, np.float32)])
                ref.sorting_flt_argsort(n, H.fp(a), order, rp.ctypes.data_as(H.PI), rr.ctypes.data_as(H.PI), H.fp(rs))
                assert np.array_equal(rp, perm) and np.array_equal(rr, rank) and np.array_equal(rs, srt)
    a = np.array([2, 1, 2, 1, 2], np.float32)          # ties keep their index order
    assert np.array_equal(argsort(h, a, 0)[1], [2, 4, 1, 3, 5])
    assert np.array_equal(argsort(h, a, 1)[1], [1, 3, 5, 2, 4])


def isotonic(lib, y):
    n = len(y)
    sol, tgt = np.zeros(n, np.float32), np.zeros(n, np.int32)
    rc = lib.isotonic_l2(n, H.fp(y), H.fp(sol), tgt.ctypes.data_as(H.PI))
    return rc, sol, tgt


def test_isotonic_regression(h):
    from sklearn.isotonic import IsotonicRegression
    ref = H.ref_rank_lib()
    rng = np.random.default_rng(1)
    cases = [rng.standard_normal(n) for n in (1, 2, 3, 10, 100)]
    cases += [np.arange(20.0), -np.arange(20.0), np.sin(np.arange(50) / 3.0) + 0.02 * np.arange(50),
              np.array([1.0, 1.0, 1.0]), np.array([3.0, 1.0, 2.0, 2.0, 0.0, 5.0])]
    for y in cases:
        y = y.astype(np.float32)
        n = len(y)
        rc, sol, tgt = isotonic(h, y)
        assert rc == 0
        want = IsotonicRegression(increasing=False).fit_transform(np.arange(n), y.astype(np.float64))
        np.testing.assert_allclose(sol, want, atol=2e-6 * (1 + np.abs(y).max()))
        # blocks: target[first] = last, target[last] = first, constant inside, decreasing across
        i = 0
        while i < n:
            e = tgt[i]
            assert e >= i and tgt[e] == i and np.all(sol[i:e + 1] == sol[i])
            if e + 1 < n:
                assert sol[e + 1] < sol[i]
            i = e + 1
        if ref is not None:
            rsol, rtgt = np.zeros(n, np.float32), np.zeros(n, np.int32)
            ref.isotonic_l2(n, H.fp(y), H.fp(rsol), rtgt.ctypes.data_as(H.PI))
            np.testing.assert_allclose(sol, rsol, atol=2e-6 * (1 + np.abs(y).max()))


class Ranker:
    def __init__(self, h, n):
        self.h, self.n = h, n
        libc = ctypes.CDLL(None)
        libc.calloc.restype = ctypes.c_void_p
        libc.calloc.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
        self.p = ctypes.cast(libc.calloc(1, ctypes.sizeof(H.SoftRankingVars)), ctypes.POINTER(H.SoftRankingVars))
        h.soft_ranking_init(n, self.p)

    def rank(self, a, reg):
        rc = self.h.soft_ranking(self.n, H.fp(a), reg, self.p)
        assert rc == 0
        v = self.p.contents
        soft = np.ctypeslib.as_array(v.soft_rank, (self.n,)).copy()
        hard = np.ctypeslib.as_array(v.hard_rank, (self.n,)).copy()
        return soft, hard

    def close(self):
        self.h.soft_ranking_free(self.p)


def in_permutahedron(x, tol):
    """Rado: x lies in the hull of the permutations of (1..n) iff its sorted
    partial sums are dominated by those of (n, n-1, ...) with equal totals"""
    n = len(x)
    xs = np.cumsum(np.sort(x)[::-1])
    ws = np.cumsum(np.arange(n, 0, -1, dtype=np.float64))
    return abs(xs[-1] - ws[-1]) < tol and np.all(xs <= ws + tol)


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_soft_ranks_are_the_permutahedron_projection(h, n):
    rng = np.random.default_rng(n)
    R = Ranker(h, n)
    try:
        verts = np.array(list(itertools.permutations(range(1, n + 1))), np.float64)
        for reg in (0.05, 0.7, 3.0, 40.0):
            for _ in range(5):
                a = rng.standard_normal(n).astype(np.float32)
                soft, hard = R.rank(a, reg)
                z = a.astype(np.float64) / reg
                x = soft.astype(np.float64)
                assert in_permutahedron(x, 1e-4)
                # the projection: (z - x) . (v - x) <= 0 for every vertex v
                assert np.max((verts - x) @ (z - x)) <= 1e-4 * (1 + np.abs(z).max())
                assert np.array_equal(hard, np.argsort(np.argsort(a)) + 1)
        a = rng.standard_normal(n).astype(np.float32)
        np.testing.assert_allclose(R.rank(a, 1e-4)[0], np.argsort(np.argsort(a)) + 1, atol=1e-4)
        np.testing.assert_allclose(R.rank(a, 1e6)[0], (n + 1) / 2, atol=1e-4)
    finally:
        R.close()


def rank_msf(h, obs, syn, reg, Ro, Rs):
    res = np.zeros_like(syn)
    mis = np.zeros(1, np.float32)
    rc = h.soft_ranking_msf(H.fp(obs), H.fp(syn), len(obs), reg, Ro.p, Rs.p, H.fp(mis), H.fp(res))
    return rc, float(mis[0]), res


def test_soft_rank_residual_is_the_transposed_jacobian(h):
    n, reg = 24, 0.3
    rng = np.random.default_rng(9)
    obs = rng.standard_normal(n).astype(np.float32)
    syn = (obs + 0.4 * rng.standard_normal(n)).astype(np.float32)
    Ro, Rs, Rj = Ranker(h, n), Ranker(h, n), Ranker(h, n)
    try:
        rc, J, res = rank_msf(h, obs, syn, reg, Ro, Rs)
        so, ss = Ro.rank(obs, reg)[0].astype(np.float64), Rs.rank(syn, reg)[0].astype(np.float64)
        assert rc == 0 and J == pytest.approx(0.5 * np.sum((ss - so) ** 2), rel=1e-6)
        e = 2.0 ** -9
        jac = np.zeros((n, n))
        for j in range(n):
            sp, sm = syn.copy(), syn.copy()
            sp[j] += e
            sm[j] -= e
            jac[:, j] = (Rj.rank(sp, reg)[0].astype(np.float64) - Rj.rank(sm, reg)[0].astype(np.float64)) / (2 * e)
        want = jac.T @ (ss - so)
        print(f"max |res| {np.abs(res).max():.3e}, max |J^T g - res| {np.abs(want - res).max():.3e}")
        np.testing.assert_allclose(res, want, atol=2e-3 * np.abs(want).max())
        # the Jacobian of a projection onto the permutahedron: symmetric, rows summing to zero
        assert np.allclose(jac, jac.T, atol=5e-3 / reg) and np.allclose(jac.sum(axis=1), 0, atol=5e-3 / reg)
    finally:
        for R in (Ro, Rs, Rj):
            R.close()


def test_soft_rank_trace_by_trace_and_errors(h):
    nt, nr, reg = 40, 3, 0.5
    rng = np.random.default_rng(4)
    obs = rng.standard_normal((nr, nt)).astype(np.float32)
    syn = (obs + 0.3 * rng.standard_normal((nr, nt))).astype(np.float32)
    Ro, Rs = Ranker(h, nt), Ranker(h, nt)
    try:
        par = (ctypes.POINTER(H.SoftRankingVars) * 2)(Ro.p, Rs.p)
        sr = np.array([0, 0, 100, 0, 200, 0, 300, 0], np.float32)
        for lo, hi, kept in ((1.0, 0.0, [0, 1, 2]), (150.0, 350.0, [1, 2]), (0.0, 10.0, [])):
            res = np.full_like(syn, 9.0)
            mis = np.zeros(2, np.float32)
            rc = h.soft_ranking_2d_trace_by_trace(rows(obs), rows(syn), H.fp(sr), lo, hi, par, reg, nt, nr,
                                                  H.fp(mis), rows(res))
            assert rc == 0
            tot = 0.0
            for r in range(nr):
                if r in kept:
                    _, J, rr = rank_msf(h, obs[r], syn[r], reg, Ro, Rs)
                    tot += J
                    np.testing.assert_array_equal(res[r], rr)
                else:
                    assert np.all(res[r] == 0)
            d = (obs[kept].astype(np.float64) - syn[kept]) ** 2
            assert mis[1] == pytest.approx(tot, rel=1e-6) and mis[0] == pytest.approx(d.sum(), rel=1e-6)
        mis = np.zeros(1, np.float32)
        assert h.soft_ranking(nt, H.fp(obs[0]), 0.0, Ro.p) == L.EINVAL
        assert "reg" in h.fwi_msf_rank_last_error().decode()
    finally:
        Ro.close()
        Rs.close()
