"""ctypes view of the CPU oracle (oracle/fwi_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.  Every function
mirrors one C entry point; the C file header lists the reference file:line
each one restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# CPML meta of the reference (fwi_2d.c:257 with fwi_sys.h MAX_F, MAX_V, COEF_R)
MAX_F, MAX_V, COEF_R = 15.0, 6000.0, 0.001


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Oracle:
    """One precision of the oracle: dtype is np.float32 or np.float64."""

    def __init__(self, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        # FWI_ORACLE_VARIANT=san: the sanitizer build (make san)
        var = os.environ.get("FWI_ORACLE_VARIANT", "")
        name = f"liboracle{'_' + var if var else ''}_{'f32' if self.dtype == np.float32 else 'f64'}.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        self.R = ctypes.c_float if self.dtype == np.float32 else ctypes.c_double
        assert self.lib.orc_real_size() == self.dtype.itemsize

    # -- helpers -----------------------------------------------------------
    def _a(self, x, dtype=None):
        dt = self.dtype if dtype is None else dtype
        a = np.ascontiguousarray(x, dtype=dt)
        return a

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(ctypes.c_void_p)

    def _r(self, v):
        return self.R(v)

    # -- setup kernels -----------------------------------------------------
    def cpml_coe(self, npml, dx, dt, maxf=MAX_F, maxv=MAX_V, coefr=COEF_R):
        a = np.zeros(2 * npml + 1, self.dtype)
        b = np.zeros(2 * npml + 1, self.dtype)
        self.lib.orc_cpml_coe(self._p(a), self._p(b), ctypes.c_int(npml), self._r(dx), self._r(dt),
                              self._r(maxf), self._r(maxv), self._r(coefr))
        return a, b

    def cpml_tables(self, npml, dx, dz, dt, maxf=MAX_F, maxv=MAX_V, coefr=COEF_R):
        tab = np.zeros((6, 2 * npml + 1), self.dtype)
        self.lib.orc_cpml_tables(ctypes.c_int(npml), self._r(dx), self._r(dz), self._r(dt),
                                 self._r(maxf), self._r(maxv), self._r(coefr), self._p(tab))
        return tab

    def fdcoe(self, dx, dz):
        out = np.zeros(8, self.dtype)
        self.lib.orc_fdcoe(self._r(dx), self._r(dz), self._p(out))
        return out

    def model_ext(self, model, npml):
        model = self._a(model)
        nz, nx = model.shape
        em = np.zeros((nz + npml, nx + 2 * npml), self.dtype)
        self.lib.orc_model_ext(self._p(model), ctypes.c_int(nx), ctypes.c_int(nz), self._p(em),
                               ctypes.c_int(nx + 2 * npml), ctypes.c_int(nz + npml))
        return em

    def ekp_from_vel(self, vel, npml, dt):
        vel = self._a(vel)
        nz, nx = vel.shape
        ekp = np.zeros((nz + npml, nx + 2 * npml), self.dtype)
        self.lib.orc_ekp_from_vel(self._p(vel), ctypes.c_int(nx), ctypes.c_int(nz), ctypes.c_int(npml),
                                  self._r(dt), self._p(ekp))
        return ekp

    # -- per-step kernels ----------------------------------------------------
    def new_pml_state(self, pnx, pnz, npml):
        d = self.dtype
        return dict(
            u=np.zeros((pnz, pnx + 1), d), v=np.zeros((pnz + 1, pnx), d),
            ux=np.zeros((pnz, pnx), d), vz=np.zeros((pnz, pnx), d),
            u0=np.zeros((pnz, npml + 1), d), u1=np.zeros((pnz, npml + 1), d),
            v1=np.zeros((npml + 1, pnx), d), ux0=np.zeros((pnz, npml), d),
            ux1=np.zeros((pnz, npml), d), vz1=np.zeros((npml, pnx), d))

    def op_pml(self, p, st, coe, tab, npml):
        """a_lap_opt_2d_pml on p[pnz][pnx]; updates st in place."""
        p = self._a(p)
        pnz, pnx = p.shape
        self.lib.orc_op_pml(self._p(p), self._p(st["u"]), self._p(st["v"]), self._p(st["ux"]),
                            self._p(st["vz"]), self._p(self._a(coe)), self._p(self._a(tab)),
                            ctypes.c_int(npml), ctypes.c_int(pnx), ctypes.c_int(pnz),
                            self._p(st["u0"]), self._p(st["u1"]), self._p(st["v1"]),
                            self._p(st["ux0"]), self._p(st["ux1"]), self._p(st["vz1"]))
        return st

    def new_adj_state(self, pnx, pnz, npml):
        d = self.dtype
        return dict(
            u0=np.zeros((pnz, npml + 1), d), u1=np.zeros((pnz, npml + 1), d),
            v1=np.zeros((npml + 1, pnx), d), ux0=np.zeros((pnz, npml), d),
            ux1=np.zeros((pnz, npml), d), vz1=np.zeros((npml, pnx), d),
            tu=np.zeros((pnz, pnx + 1), d), tv=np.zeros((pnz + 1, pnx), d), tw=np.zeros((pnz, pnx), d))

    def op_pml_adj(self, w, st, coe, tab, npml):
        w = self._a(w)
        pnz, pnx = w.shape
        out = np.zeros_like(w)
        self.lib.orc_op_pml_adj(self._p(w), self._p(out), self._p(self._a(coe)), self._p(self._a(tab)),
                                ctypes.c_int(npml), ctypes.c_int(pnx), ctypes.c_int(pnz),
                                self._p(st["u0"]), self._p(st["u1"]), self._p(st["v1"]),
                                self._p(st["ux0"]), self._p(st["ux1"]), self._p(st["vz1"]),
                                self._p(st["tu"]), self._p(st["tv"]), self._p(st["tw"]))
        return out

    def op_interior(self, p, coe, u=None, v=None, ux=None, vz=None):
        p = self._a(p)
        pnz, pnx = p.shape
        d = self.dtype
        u = np.zeros((pnz, pnx + 1), d) if u is None else u
        v = np.zeros((pnz + 1, pnx), d) if v is None else v
        ux = np.zeros((pnz, pnx), d) if ux is None else ux
        vz = np.zeros((pnz, pnx), d) if vz is None else vz
        self.lib.orc_op_interior(self._p(p), self._p(u), self._p(v), self._p(ux), self._p(vz),
                                 self._p(self._a(coe)), ctypes.c_int(pnx), ctypes.c_int(pnz))
        return u, v, ux, vz

    def time_integral(self, pnow, pold, pxx, pzz, ekp):
        pnow, pold, pxx, pzz, ekp = (self._a(x) for x in (pnow, pold, pxx, pzz, ekp))
        m, n = pnow.shape
        pnew = np.zeros_like(pnow)
        self.lib.orc_time_integral(self._p(pnew), self._p(pnow), self._p(pold), self._p(pxx),
                                   self._p(pzz), self._p(ekp), ctypes.c_int(m), ctypes.c_int(n))
        return pnew

    def write_bdr(self, pf, nx, nz, npml):
        pf = self._a(pf)
        ud = np.zeros((6, nx), self.dtype)
        lr = np.zeros((nz, 6), self.dtype)
        self.lib.orc_write_bdr(self._p(ud), self._p(lr), self._p(pf), ctypes.c_int(nx), ctypes.c_int(nz),
                               ctypes.c_int(npml))
        return ud, lr

    def read_bdr(self, ud, lr, pb):
        pb = self._a(pb).copy()
        nz, nx = pb.shape
        self.lib.orc_read_bdr(self._p(self._a(ud)), self._p(self._a(lr)), self._p(pb), ctypes.c_int(nx),
                              ctypes.c_int(nz))
        return pb

    # -- driver ------------------------------------------------------------------
    def _geom_args(self, g):
        return (ctypes.c_int(g.nx), ctypes.c_int(g.nz), ctypes.c_int(g.npml), self._r(g.dx), self._r(g.dz),
                self._r(g.dt), self._r(MAX_F), self._r(MAX_V), self._r(COEF_R))

    def forward(self, g, vel, shot, want_strips=False, want_xall=False):
        """Returns dict(syn[nr][nt], pstrip, xstrip, plast, xall)."""
        vel = self._a(vel)
        nt, nr = shot.nt, shot.nr
        sl = 6 * g.nx + 6 * g.nz
        syn = np.zeros((nr, nt), self.dtype)
        ps = np.zeros((nt, sl), self.dtype) if want_strips else None
        xs = np.zeros((nt, sl), self.dtype) if want_strips else None
        pl = np.zeros((2, g.nz, g.nx), self.dtype) if want_strips else None
        xa = np.zeros((nt, g.nz, g.nx), self.dtype) if want_xall else None
        nul = ctypes.c_void_p(0)
        wav = self._a(shot.wav)
        irx = np.ascontiguousarray(shot.irx, np.int32)
        irz = np.ascontiguousarray(shot.irz, np.int32)
        self.lib.orc_forward(*self._geom_args(g), self._p(vel), ctypes.c_int(nt), ctypes.c_int(shot.isx),
                             ctypes.c_int(shot.isz), self._p(wav), ctypes.c_int(nr), self._p(irx), self._p(irz),
                             self._p(syn), self._p(ps) if ps is not None else nul,
                             self._p(xs) if xs is not None else nul, self._p(pl) if pl is not None else nul,
                             self._p(xa) if xa is not None else nul)
        return dict(syn=syn, pstrip=ps, xstrip=xs, plast=pl, xall=xa)

    def backward(self, g, vel, shot, res, fwd, use_xall=False):
        """Returns dict(gk, grad, musrc)."""
        vel = self._a(vel)
        nt, nr = shot.nt, shot.nr
        gk = np.zeros((g.nz, g.nx), np.float64)
        grad = np.zeros((g.nz, g.nx), np.float64)
        musrc = np.zeros(nt, self.dtype)
        nul = ctypes.c_void_p(0)
        res = self._a(res)
        wav = self._a(shot.wav)
        irx = np.ascontiguousarray(shot.irx, np.int32)
        irz = np.ascontiguousarray(shot.irz, np.int32)
        if use_xall:
            ps = xs = pl = nul
            xa = self._p(fwd["xall"])
        else:
            ps, xs, pl = self._p(fwd["pstrip"]), self._p(fwd["xstrip"]), self._p(fwd["plast"])
            xa = nul
        self.lib.orc_backward(*self._geom_args(g), self._p(vel), ctypes.c_int(nt), ctypes.c_int(shot.isx),
                              ctypes.c_int(shot.isz), self._p(wav), ctypes.c_int(nr), self._p(irx), self._p(irz),
                              self._p(res), ps, xs, pl, xa, self._p(gk), self._p(grad), self._p(musrc))
        return dict(gk=gk, grad=grad, musrc=musrc)

    def misfit_l2(self, syn, obs):
        syn, obs = self._a(syn), self._a(obs)
        res = np.zeros_like(syn)
        self.lib.orc_misfit_l2.restype = ctypes.c_double
        J = self.lib.orc_misfit_l2(self._p(syn), self._p(obs), self._p(res), ctypes.c_long(syn.size))
        return J, res

    def gradient(self, g, vel, shot, obs):
        """forward + L2 misfit + backward with reconstruction: (J, syn, grad, gk)."""
        vel = self._a(vel)
        syn = np.zeros((shot.nr, shot.nt), self.dtype)
        grad = np.zeros((g.nz, g.nx), np.float64)
        gk = np.zeros((g.nz, g.nx), np.float64)
        wav = self._a(shot.wav)
        irx = np.ascontiguousarray(shot.irx, np.int32)
        irz = np.ascontiguousarray(shot.irz, np.int32)
        self.lib.orc_gradient.restype = ctypes.c_double
        J = self.lib.orc_gradient(*self._geom_args(g), self._p(vel), ctypes.c_int(shot.nt),
                                  ctypes.c_int(shot.isx), ctypes.c_int(shot.isz), self._p(wav),
                                  ctypes.c_int(shot.nr), self._p(irx), self._p(irz), self._p(self._a(obs)),
                                  self._p(syn), self._p(grad), self._p(gk))
        return J, syn, grad, gk
