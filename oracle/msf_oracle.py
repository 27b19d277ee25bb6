"""Test infrastructure only: numpy restatements of the reference's FFT-based
trace misfits, in FP64 with direct (np.convolve) linear convolutions in place
of the reference's FFTW circular ones.  The product never imports this file.

    xcorr1d   xcorr_msf.c:4-181      (XCORR1D)
    mf1d      lsq_decon_cg.c:82-392  (MF1D, with fun_Ax/Ax_slicing/Atr_slicing/
                                      fun_Dx/fun_Dtr at lsq_decon_cg.c:3-79)

The reference computes on FFT lengths s0 >= 2nt-1 (xcorr_msf.c:10) and
s0 >= nt+nmodel-1, s1 >= 2nt-1 (lsq_decon_cg.c:88-89), on which its circular
convolutions equal the linear ones used here.  Pinned in tests/test_msf.py by
finite-difference gradients and dense normal-equation solves, not by the
reference itself (it needs FFTW, which is absent).
"""
import numpy as np


def _offsets_keep(sr, rmin, rmax, r):
    if sr is None:
        return True
    d = np.hypot(float(sr[2 + 2 * r]) - float(sr[0]), float(sr[3 + 2 * r]) - float(sr[1]))
    return not (d < rmin or d > rmax)


def xcorr1d(obs, syn, mode="middle", sr=None, rmin=-np.inf, rmax=np.inf):
    """XCORR1D, xcorr_msf.c:4-181.  obs, syn: (nr, nt).  Returns (misfit[2], res)."""
    obs = np.asarray(obs, np.float64)
    syn = np.asarray(syn, np.float64)
    nr, nt = obs.shape
    h = nt // 2
    if mode == "middle":                                  # xcorr_msf.c:39-54
        T = np.abs(np.arange(nt) - h).astype(np.float64)
        off = h
    else:
        T = np.arange(nt, dtype=np.float64)
        off = 0
    m0 = m1 = 0.0
    res = np.zeros((nr, nt))
    for r in range(nr):
        e1 = np.abs(obs[r]).sum()                         # :62-66
        e2 = np.abs(syn[r]).sum()
        if e2 / (e1 + 1e-12) > 10 or e1 / (e2 + 1e-12) > 10 or not _offsets_keep(sr, rmin, rmax, r):
            continue                                      # :77-78
        A = obs[r][::-1]                                  # :84-88
        m0 += np.abs(A - syn[r]).sum()
        full = np.convolve(A, syn[r])                     # :90-101, length 2nt-1
        x = full[off:off + nt]                            # :103-112
        b = T * x                                         # :125
        tt1, tt2 = b @ b, x @ x                           # :127-128
        if tt2 == 0.0:
            continue
        alpha = tt1 / tt2
        m1 += alpha                                       # :131
        y = (T * b - x * alpha) / tt2 * 2.0               # :133-134
        full = np.convolve(obs[r], y)                     # :141-157
        res[r] = full[off:off + nt]                       # :159-169
    return np.array([m0, m1]), res


def _mf_ops(a, nt, nm, middle, diff, eps0):
    """A, A^T, D, D^T of lsq_decon_cg.c:3-79 for one trace (a = 100 obs)."""
    h = nm // 2
    ar = a[::-1]

    def A(w):                                             # fun_Ax + Ax_slicing
        full = np.convolve(a, w[:nm])
        return full[h:h + nt] if middle else full[:nt]

    def At(r):                                            # fun_Ax(fftAt, .) + Atr_slicing
        full = np.convolve(ar, r[:nt])
        lo = nt - h - 1 if middle else nt - 1
        out = np.zeros(nm)
        seg = full[lo:lo + nm]
        out[:len(seg)] = seg
        return out

    def D(w):                                             # fun_Dx
        if diff:
            r1 = np.empty(nm)
            r1[:-1] = w[1:nm] - w[:nm - 1]
            r1[-1] = w[nm - 1]
        else:
            r1 = w[:nm].copy()
        return eps0 * r1

    def Dt_add(r1, z):                                    # fun_Dtr
        if diff:
            z[1:nm - 1] += (r1[:nm - 2] - r1[1:nm - 1]) * eps0
            z[0] -= r1[0] * eps0
            z[nm - 1] += (r1[nm - 1] + r1[nm - 2]) * eps0
        else:
            z += r1 * eps0
        return z

    return A, At, D, Dt_add


def mf1d(obs, syn, nmodel, niter, mode="middle", reg="diff", tor0=0.0, eps0=1.0, Tdepressor=1.0,
         return_w=False):
    """MF1D, lsq_decon_cg.c:82-392.  Returns (misfit[2], res) (and the filters w)."""
    obs = np.asarray(obs, np.float64)
    syn = np.asarray(syn, np.float64)
    nr, nt = obs.shape
    nm = nmodel
    middle = mode == "middle"
    diff = reg == "diff"
    h = nm // 2
    T = (np.abs(np.arange(nm) - h) if middle else np.arange(nm)).astype(np.float64) * Tdepressor  # :147-164
    m0 = float(((obs - syn) ** 2).sum())                 # :168-172
    m1 = 0.0
    res = np.zeros((nr, nt))
    ws = []
    for ix in range(nr):
        a = obs[ix] * 100.0                               # :183-184
        b = syn[ix] * 100.0
        A, At, D, Dt_add = _mf_ops(a, nt, nm, middle, diff, eps0)
        x = np.zeros(nm)                                  # :187
        r0 = b - A(x)                                     # :192-201
        r1 = -D(x)                                        # :203-205
        z = Dt_add(r1, At(r0))                            # :208-213
        p = z.copy()
        for _ in range(niter):                            # :220-274
            rn0 = np.sqrt(r0 @ r0 + r1 @ r1)
            w0, w1 = A(p), D(p)
            num = z @ z
            den = w0 @ w0 + w1 @ w1
            if num == 0.0 or den == 0.0:
                break
            al = num / den
            x = x + al * p
            r0 = r0 - al * w0
            r1 = r1 - al * w1
            z = Dt_add(r1, At(r0))
            be = (z @ z) / num
            p = z + be * p
            rn1 = np.sqrt(r0 @ r0 + r1 @ r1)
            if abs(rn1 - rn0) < tor0:
                break
        ws.append(x)
        z0 = T * x                                        # :280-290
        num, den = z0 @ z0, x @ x
        if den == 0.0:
            continue
        alpha = num / den
        m1 += alpha
        c = (T * z0 - x * alpha) / den * 2.0
        y0 = np.zeros(nt)                                 # :277-278
        y1 = np.zeros(nm)
        r = c - Dt_add(y1, At(y0))                        # :297-304
        p0, p1 = A(r), D(r)                               # :309-315
        for _ in range(niter):                            # :317-360
            rn0 = np.sqrt(r @ r)
            nn = rn0 * rn0
            dd = p0 @ p0 + p1 @ p1
            if nn == 0.0 or dd == 0.0:
                break
            al = nn / dd
            y0 = y0 + al * p0
            y1 = y1 + al * p1
            z = Dt_add(p1, At(p0))
            r = r - al * z
            rr2 = r @ r
            be = rr2 / nn
            p0 = be * p0 + A(r)
            p1 = be * p1 + D(r)
            if abs(np.sqrt(rr2) - rn0) < tor0:
                break
        res[ix] = y0                                      # :376-378
    out = (np.array([m0, m1]), res)
    return out + (np.array(ws),) if return_w else out


def mf_dense(a, nt, nm, middle, diff, eps0):
    """Dense matrices A (nt x nm) and D (nm x nm) built from the definitions
    (A w = the nt-sample slice of the linear convolution a * w; D the
    eps0-scaled forward difference with the last row kept, or identity)."""
    h = nm // 2
    off = h if middle else 0
    Am = np.zeros((nt, nm))
    for i in range(nt):
        for k in range(nm):
            j = i + off - k
            if 0 <= j < len(a):
                Am[i, k] = a[j]
    if diff:
        Dm = -np.eye(nm) + np.eye(nm, k=1)
        Dm[-1, -1] = 1.0
    else:
        Dm = np.eye(nm)
    return Am, eps0 * Dm
