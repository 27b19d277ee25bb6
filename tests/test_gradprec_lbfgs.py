"""Gradient post-processing (csrc/fwi_gradprec.c: triangle smoothing,
triangle_smooth.c:5-110; grad_precond_new5, grad_prec.c:5-52) and the L-BFGS
loop (csrc/fwi_lbfgs.c: MBotOpt.c:65-221, lbfgs.h).  Neither reference file
builds here (triangle_smooth.h is absent; grad_prec.c needs header.h's GSL),
so they are pinned by closed forms: the triangle filter is the convolution
with the 1..r..1 / r^2 kernel over the evenly reflected input; the L-BFGS
direction is the dense BFGS inverse-Hessian update applied to -g, every
accepted step meets the strong Wolfe conditions, and the step schedule and
exit codes are MBotOpt.c's.  CPU only."""
import ctypes

import numpy as np
import pytest

from fwihip import plans as PC
from fwihip import lib as L

c_int, c_float = ctypes.c_int, ctypes.c_float
PF = ctypes.POINTER(c_float)
PD = ctypes.POINTER(ctypes.c_double)


class LbfgsParam(ctypes.Structure):
    _fields_ = [("m", c_int), ("epsilon", c_float), ("max_iterations", c_int), ("max_linesearch", c_int),
                ("ftol", c_float), ("wolfe", c_float)]


EVAL = ctypes.CFUNCTYPE(c_int, ctypes.c_void_p, PF, PF, PF, c_int)
PROG = ctypes.CFUNCTYPE(None, ctypes.c_void_p, c_int, ctypes.c_double, ctypes.c_double, c_int, PF, PF, PD, c_int)


@pytest.fixture(scope="module")
def h():
    lib = PC.lib()
    lib.triangle.argtypes = [c_int, c_int, PF, PF]
    lib.triangle.restype = c_int
    lib.triangle2.argtypes = [c_int, c_int, c_int, c_int, c_int, PF, PF]
    lib.triangle2.restype = c_int
    lib.grad_precond_new5.argtypes = [ctypes.POINTER(PC.FwiGrad), ctypes.POINTER(PC.FwiShape),
                                      ctypes.POINTER(c_int)]
    lib.grad_precond_new5.restype = c_int
    lib.fwi_plan_set_grad.argtypes = [ctypes.POINTER(PC.FwiPlan), c_float, c_float] + [c_int] * 6
    lib.fwi_plan_set_grad.restype = c_int
    lib.fwi_lbfgs_default_param.argtypes = [ctypes.POINTER(LbfgsParam)]
    lib.fwi_lbfgs_default_param.restype = None
    lib.fwi_lbfgs.argtypes = [c_int, PF, EVAL, PROG, ctypes.c_void_p, ctypes.POINTER(LbfgsParam),
                              ctypes.POINTER(c_int)]
    lib.fwi_lbfgs.restype = c_int
    return lib


def fp(a):
    return a.ctypes.data_as(PF)


def tri_ref(x, r):
    """convolution with the triangle 1..r..1 / r^2 over the evenly reflected input"""
    k = np.concatenate([np.arange(1, r + 1), np.arange(r - 1, 0, -1)]).astype(np.float64) / (r * r)
    xe = np.pad(x.astype(np.float64), r - 1, mode="symmetric")
    return np.convolve(xe, k, mode="valid")


@pytest.mark.parametrize("n,r", [(50, 1), (50, 3), (50, 8), (7, 7), (7, 8), (1, 1), (1, 2), (64, 20)])
def test_triangle_closed_form(h, n, r):
    rng = np.random.default_rng(n * 100 + r)
    x = rng.standard_normal(n).astype(np.float32)
    y = np.zeros_like(x)
    assert h.triangle(r, n, fp(x), fp(y)) == 0
    assert np.abs(y - tri_ref(x, r)).max() < 1e-6 * max(1.0, np.abs(x).max())
    c = np.full(n, 2.5, np.float32)                       # constants are kept
    assert h.triangle(r, n, fp(c), fp(y)) == 0
    assert np.allclose(y, 2.5, rtol=1e-6)


def test_triangle_rejects_too_wide(h):
    x = np.ones(5, np.float32)
    assert h.triangle(7, 5, fp(x), fp(x)) == 1            # the reference writes out of bounds there
    assert h.triangle(0, 5, fp(x), fp(x)) == 1


def test_triangle2_separable_repeat(h):
    rng = np.random.default_rng(3)
    n1, n2 = 30, 17
    u = rng.standard_normal((n2, n1)).astype(np.float32)
    v = np.zeros_like(u)
    assert h.triangle2(4, 3, 2, n1, n2, fp(u), fp(v)) == 0
    w = u.astype(np.float64)
    for _ in range(2):
        w = np.stack([tri_ref(row, 4) for row in w])
    for _ in range(2):
        w = np.stack([tri_ref(col, 3) for col in w.T]).T
    assert np.abs(v - w).max() < 1e-5
    uu = u.copy()
    assert h.triangle2(4, 3, 2, n1, n2, fp(uu), fp(uu)) == 0   # in place
    assert np.array_equal(uu, v)


def test_grad_precond_new5(h):
    nx, nz = 40, 25
    p = h.fwi_plan_alloc(nx, nz, 10, 11, 1, 1, 10.0, 10.0, 1e-3)
    rng = np.random.default_rng(5)
    gr = rng.standard_normal((nz, nx)).astype(np.float32)
    gr[10, 10] = 40.0                                       # an outlier the clip removes
    PC.view(p.contents.grad.contents.body, (nz, nx))[:] = gr
    bottom = (3 + (np.arange(nx) % 5)).astype(np.int32)
    assert h.fwi_plan_set_grad(p, 0.6, 2.0, 3, 9, 2, 2, 9, 1) == 0
    assert h.grad_precond_new5(p.contents.grad, p.contents.grid.contents.b,
                               bottom.ctypes.data_as(ctypes.POINTER(c_int))) == 0
    got = PC.view(p.contents.grad.contents.body, (nz, nx)).copy()
    w = gr.astype(np.float64)
    w = np.clip(w, w.min() * 0.6, w.max() * 0.6)

    def smooth(a, r, rep):        # rectx on both axes (grad_prec.c:37,44)
        for _ in range(rep):
            a = np.stack([tri_ref(row, r) for row in a])
        for _ in range(rep):
            a = np.stack([tri_ref(col, r) for col in a.T]).T
        return a

    w = smooth(w, 3, 2)
    w[np.arange(nz)[:, None] < bottom[None, :]] = 0.0
    w = smooth(w, 2, 1) * 2.0
    assert np.abs(got - w).max() < 1e-5 * np.abs(w).max()
    h.fwi_plan_free(p)


# ---------------------------------------------------------------- L-BFGS

class Quadratic:
    """f = 1/2 x^T A x - b^T x + c, recorded per evaluation."""

    def __init__(self, n, cond, seed, flip=False):
        rng = np.random.default_rng(seed)
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        self.A = q @ np.diag(np.geomspace(1.0, cond, n)) @ q.T
        self.b = rng.standard_normal(n)
        self.flip = flip
        self.trials = []
        self.accepted = []

    def f(self, x):
        return 0.5 * x @ self.A @ x - self.b @ x + 10.0

    def grad(self, x):
        return self.A @ x - self.b

    def callbacks(self, n):
        def ev(user, x, g, fx, k):
            xv = np.ctypeslib.as_array(x, (n,)).astype(np.float64)
            gv = self.grad(xv) * (-1.0 if self.flip else 1.0)
            np.ctypeslib.as_array(g, (n,))[:] = gv
            fx[0] = self.f(xv)
            self.trials.append((k, xv.copy(), float(np.float32(self.f(xv))), gv.astype(np.float32)))
            return 0

        def pr(user, k, fx, step, trials, x, g, d, nn):
            self.accepted.append(dict(k=k, fx=fx, step=step, trials=trials,
                                      x=np.ctypeslib.as_array(x, (n,)).astype(np.float64),
                                      g=np.ctypeslib.as_array(g, (n,)).astype(np.float64),
                                      d=None if not d else np.ctypeslib.as_array(d, (n,)).copy()))

        return EVAL(ev), PROG(pr)


def run_lbfgs(h, prob, n, x0, **kw):
    par = LbfgsParam()
    h.fwi_lbfgs_default_param(ctypes.byref(par))
    for k, v in kw.items():
        setattr(par, k, v)
    ev, pr = prob.callbacks(n)
    x = x0.astype(np.float32).copy()
    st = c_int(-1)
    rc = h.fwi_lbfgs(n, fp(x), ev, pr, None, ctypes.byref(par), ctypes.byref(st))
    return rc, st.value, x, par


def test_lbfgs_defaults(h):
    par = LbfgsParam()
    h.fwi_lbfgs_default_param(ctypes.byref(par))
    assert (par.m, par.max_iterations, par.max_linesearch) == (12, 8, 6)
    assert par.epsilon == pytest.approx(1e-9) and par.ftol == pytest.approx(1e-7) and par.wolfe == pytest.approx(0.98)


def test_lbfgs_wolfe_steps_and_bfgs_direction(h):
    n, m = 12, 4
    prob = Quadratic(n, 30.0, 1)
    x0 = np.zeros(n) + 0.3
    rc, st, x, par = run_lbfgs(h, prob, n, x0, m=m, max_iterations=25, max_linesearch=20, wolfe=0.9)
    assert rc == 0
    acc = prob.accepted
    assert acc[0]["k"] == 0 and len(acc) >= 6
    # every accepted step satisfies the Armijo and strong Wolfe tests of MBotOpt.c:104-130
    for a0, a1 in zip(acc[:-1], acc[1:]):
        d = a1["d"]
        dginit = a0["g"] @ d
        assert dginit < 0
        assert a1["fx"] <= a0["fx"] + a1["step"] * par.ftol * dginit + 1e-6 * abs(a0["fx"])
        dg = a1["g"] @ d
        assert par.wolfe * dginit - 1e-6 <= dg <= -par.wolfe * dginit + 1e-6
    # the search direction is the BFGS inverse-Hessian update of the last m pairs applied to -g
    xs = [a["x"] for a in acc]
    gs = [a["g"] for a in acc]
    for k in range(1, len(acc) - 1):
        pairs = [(xs[j + 1] - xs[j], gs[j + 1] - gs[j]) for j in range(max(0, k - m), k)]
        s, y = pairs[-1]
        H = (s @ y) / (y @ y) * np.eye(n)
        for s, y in pairs:
            rho = 1.0 / (y @ s)
            V = np.eye(n) - rho * np.outer(y, s)
            H = V.T @ H @ V + rho * np.outer(s, s)
        want = H @ (-gs[k])
        got = acc[k + 1]["d"]
        assert np.abs(got - want).max() < 1e-4 * np.abs(want).max(), k
    # monotone descent; MBotOpt.c's schedule keeps the first trial at the
    # accepted 0.2 rather than the unit step, so progress is steady, not quadratic
    fx = [a["fx"] for a in acc]
    assert all(b < a for a, b in zip(fx[:-1], fx[1:]))
    assert np.linalg.norm(gs[-1]) < 1e-2 * np.linalg.norm(gs[0])


def test_lbfgs_step_schedule(h):
    """First trial step 0.8; after iterations 1-2 the first trial is 0.2,
    from the fourth on the mean of the last two accepted steps (MBotOpt.c:71,214)."""
    n = 6
    prob = Quadratic(n, 5.0, 2)
    rc, st, x, par = run_lbfgs(h, prob, n, np.ones(n), m=3, max_iterations=7, max_linesearch=20)
    assert rc == 0 and st == 2                              # LBFGSERR_MAXIMUMITERATION
    acc = prob.accepted
    steps = {a["k"]: a["step"] for a in acc}
    xs = {a["k"]: a["x"] for a in acc}
    ds = {a["k"]: a["d"] for a in acc}
    for k in range(1, max(steps) + 1):
        first = [t for t in prob.trials if t[0] == k][0]
        trial_step = (first[1] - xs[k - 1]) @ ds[k] / (ds[k] @ ds[k])
        want = 0.8 if k == 1 else 0.2 if k < 4 else 0.5 * (steps[k - 1] + steps[k - 2])
        assert trial_step == pytest.approx(want, rel=1e-4), k
    assert max(steps) == 7


def test_lbfgs_line_search_exhausted(h):
    """A gradient of the wrong sign makes every trial fail the Armijo test:
    max_linesearch trials, then LBFGSERR_MAXIMUMLINESEARCH (MBotOpt.c:140-145)."""
    n = 5
    prob = Quadratic(n, 3.0, 3, flip=True)
    rc, st, x, par = run_lbfgs(h, prob, n, np.ones(n), max_linesearch=6)
    assert rc == 0 and st == 1
    assert len(prob.trials) == 1 + 6
    steps = [np.linalg.norm(t[1] - np.ones(n)) for t in prob.trials[1:]]
    assert np.allclose(np.array(steps[1:]) / np.array(steps[:-1]), 0.5, rtol=1e-4)   # halved each time
