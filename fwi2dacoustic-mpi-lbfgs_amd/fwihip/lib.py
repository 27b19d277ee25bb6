"""ctypes binding of libfwi_hip.so (include/fwi_hip.h).

The library is the compute path; this module only marshals arguments.  It
fails loudly when the library has not been built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("FWI_HIP_LIBRARY") or os.path.join(PKG, "libfwi_hip.so")
# the same library linked against PyTorch's bundled HIP runtime: chosen when
# torch is already imported, so that the process holds one HIP runtime
LIB_TORCH_PATH = os.path.join(PKG, "libfwi_hip_torch.so")

OK, EINVAL, ENOMEM, EHIP, ESTATE = 0, 1, 2, 3, 4

# CPML meta of the reference (fwi_sys.h: MAX_F, MAX_V, COEF_R)
MAX_F, MAX_V, COEF_R = 15.0, 6000.0, 0.001

EXPORTS = (
    "fwi_hip_create", "fwi_hip_destroy", "fwi_hip_last_error", "fwi_hip_bytes", "fwi_hip_set_model",
    "fwi_hip_set_shots", "fwi_hip_forward", "fwi_hip_forward_tm", "fwi_hip_set_residual_tm", "fwi_hip_set_observed", "fwi_hip_misfit_l2",
    "fwi_hip_set_residual", "fwi_hip_adjoint", "fwi_hip_synchronize", "fwi_hip_set_timing",
    "fwi_hip_kernel_time", "fwi_hip_read_state", "fwi_hip_check_model", "fwi_hip_op_pml", "fwi_hip_op_interior",
    "fwi_hip_time_integral", "fwi_hip_op_pml_noalloc", "fwi_hip_write_bdr", "fwi_hip_write_bdr_fs_no", "fwi_hip_model_ext_fs_no", "fwi_hip_read_bdr", "fwi_hip_cpml_coe",
    "fwi_hip_set_graph", "fwi_hip_set_kernels", "fwi_hip_set_rec_group", "fwi_hip_set_imaging", "fwi_hip_set_bwd_chunk", "fwi_hip_set_bwd_lanes", "fwi_hip_set_bwd_general_rows", "fwi_hip_set_checkpointing", "fwi_hip_set_revolve", "fwi_hip_backward_plan", "fwi_hip_plan_memory", "fwi_hip_revolve_schedule", "fwi_hip_set_bwd_overlap", "fwi_hip_set_bwd_fused", "fwi_hip_set_concurrent", "fwi_hip_set_interior_wgs", "fwi_hip_set_pass_shape", "fwi_hip_set_fwd_mixed", "fwi_hip_set_bwd_mixed", "fwi_hip_device_count", "fwi_hip_dmalloc", "fwi_hip_dfree", "fwi_hip_host_alloc", "fwi_hip_host_free", "fwi_hip_memcpy",
)


class FwiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fwi_hip error {code}: {msg}")
        self.code = code


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int), ("npml", ctypes.c_int),
                ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
                ("max_f", ctypes.c_float), ("max_v", ctypes.c_float), ("coef_r", ctypes.c_float)]


class Desc(ctypes.Structure):
    _fields_ = [("grid", Grid), ("max_shots", ctypes.c_int), ("max_nt", ctypes.c_int),
                ("max_nr", ctypes.c_int), ("tile_z", ctypes.c_int), ("shots_per_group", ctypes.c_int),
                ("hbm_budget", ctypes.c_size_t), ("stream", ctypes.c_void_p)]


class ShotC(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("isx", ctypes.c_int), ("isz", ctypes.c_int), ("nr", ctypes.c_int),
                ("irx", ctypes.c_void_p), ("irz", ctypes.c_void_p), ("wav", ctypes.c_void_p)]


_lib = None


def load():
    """Load the library (raises if it is missing: the HIP path is the product)."""
    global _lib
    if _lib is None:
        import sys
        path = LIB_PATH
        if "torch" in sys.modules and not os.environ.get("FWI_HIP_LIBRARY"):
            path = LIB_TORCH_PATH
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built: run __graft_entry__.build() "
                                    "or python fwi2dacoustic-mpi-lbfgs_amd/build.py")
        lib = ctypes.CDLL(path)
        for name in EXPORTS:
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
        lib.fwi_hip_last_error.restype = ctypes.c_char_p
        lib.fwi_hip_bytes.restype = ctypes.c_size_t
        lib.fwi_hip_destroy.restype = None
        _lib = lib
    return _lib


def make_grid(g, max_f=MAX_F, max_v=MAX_V, coef_r=COEF_R):
    return Grid(g.nx, g.nz, g.npml, g.dx, g.dz, g.dt, max_f, max_v, coef_r)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


BLOCK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int)


class Context:
    """One fwi_hip context: a batch of shots on one model."""

    def __init__(self, geom, max_shots, max_nt, max_nr, tile_z=0, shots_per_group=0, hbm_budget=0):
        self.lib = load()
        self.geom = geom
        d = Desc(make_grid(geom), max_shots, max_nt, max_nr, tile_z, shots_per_group, hbm_budget, None)
        h = ctypes.c_void_p()
        rc = self.lib.fwi_hip_create(ctypes.byref(d), ctypes.byref(h))
        if rc:
            raise FwiError(rc, self.lib.fwi_hip_last_error(None).decode())
        self.h = h
        self._keep = []
        self.shots = []

    def _chk(self, rc):
        if rc:
            raise FwiError(rc, self.lib.fwi_hip_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.fwi_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bytes(self):
        return self.lib.fwi_hip_bytes(self.h)

    def set_model(self, vel):
        vel = np.ascontiguousarray(vel, np.float32)
        assert vel.shape == (self.geom.nz, self.geom.nx)
        self._chk(self.lib.fwi_hip_set_model(self.h, _p(vel)))

    def set_shots(self, shots):
        arr = (ShotC * len(shots))()
        keep = []
        for i, s in enumerate(shots):
            irx = np.ascontiguousarray(s.irx, np.int32)
            irz = np.ascontiguousarray(s.irz, np.int32)
            wav = np.ascontiguousarray(s.wav, np.float32)
            keep += [irx, irz, wav]
            arr[i] = ShotC(s.nt, s.isx, s.isz, len(irx), irx.ctypes.data, irz.ctypes.data, wav.ctypes.data)
        self._chk(self.lib.fwi_hip_set_shots(self.h, len(shots), arr))
        self.shots = list(shots)

    def ntrace(self):
        return sum(s.nr * s.nt for s in self.shots)

    def split(self, packed):
        out, o = [], 0
        for s in self.shots:
            out.append(packed[o:o + s.nr * s.nt].reshape(s.nr, s.nt))
            o += s.nr * s.nt
        return out

    def pack(self, traces):
        return np.ascontiguousarray(np.concatenate([np.asarray(t, np.float32).ravel() for t in traces]))

    def forward(self, want_traces=True):
        out = np.empty(self.ntrace(), np.float32) if want_traces else None
        self._chk(self.lib.fwi_hip_forward(self.h, _p(out)))
        return out

    def forward_tm(self, callback=None, out=None):
        """fwi_hip_forward_tm: the traces time-major (split_tm), and
        callback(k0, k1) -> int for each block of rows as it arrives (the
        packed array is readable and writable inside the callback)."""
        out = np.empty(self.ntrace(), np.float32) if out is None else out
        assert out.dtype == np.float32 and out.size == self.ntrace() and out.flags.c_contiguous
        fn = BLOCK_FN(lambda user, k0, k1: int(callback(k0, k1))) if callback is not None else None
        self._chk(self.lib.fwi_hip_forward_tm(self.h, _p(out), fn, None))
        return out

    def split_tm(self, packed):
        out, o = [], 0
        for s in self.shots:
            out.append(packed[o:o + s.nr * s.nt].reshape(s.nt, s.nr))
            o += s.nr * s.nt
        return out

    def set_residual_tm(self, res_tm_packed):
        """fwi_hip_set_residual_tm: the array is kept until the next adjoint."""
        res = np.ascontiguousarray(res_tm_packed, np.float32)
        assert res.size == self.ntrace()
        self._res_tm = res
        self._chk(self.lib.fwi_hip_set_residual_tm(self.h, _p(res)))

    def set_observed(self, obs_packed):
        obs = np.ascontiguousarray(obs_packed, np.float32)
        assert obs.size == self.ntrace()
        self._chk(self.lib.fwi_hip_set_observed(self.h, _p(obs)))

    def misfit_l2(self):
        J = ctypes.c_double()
        self._chk(self.lib.fwi_hip_misfit_l2(self.h, ctypes.byref(J)))
        return J.value

    def set_residual(self, res_packed):
        res = np.ascontiguousarray(res_packed, np.float32)
        assert res.size == self.ntrace()
        self._chk(self.lib.fwi_hip_set_residual(self.h, _p(res)))

    def adjoint(self, want=True):
        g = self.geom
        grad = np.empty((g.nz, g.nx), np.float64) if want else None
        gk = np.empty((g.nz, g.nx), np.float64) if want else None
        try:
            self._chk(self.lib.fwi_hip_adjoint(self.h, _p(grad), _p(gk)))
        finally:
            self._res_tm = None
        return grad, gk

    def synchronize(self):
        self._chk(self.lib.fwi_hip_synchronize(self.h))

    def set_kernels(self, kind):
        """0: 2-D tile kernels (default), 1: z-marching kernels."""
        self._chk(self.lib.fwi_hip_set_kernels(self.h, int(kind)))

    def set_rec_group(self, shots):
        """Shots per workgroup of the tile reconstruction + imaging kernel."""
        self._chk(self.lib.fwi_hip_set_rec_group(self.h, int(shots)))

    def set_checkpointing(self, every, storage="fp32"):
        """Snapshots every `every` steps (0: boundary-saving reconstruction),
        stored as "fp32", "fp16" or "bf16"."""
        code = {"fp32": 0, "fp16": 1, "bf16": 2}.get(storage, -1)
        self._chk(self.lib.fwi_hip_set_checkpointing(self.h, int(every), code))

    def set_revolve(self, snaps, storage="fp32"):
        """Binomial checkpointing with `snaps` slots (-1: chosen from the HBM
        budget, 0: off), snapshots stored as "fp32", "fp16" or "bf16"."""
        code = {"fp32": 0, "fp16": 1, "bf16": 2}.get(storage, -1)
        self._chk(self.lib.fwi_hip_set_revolve(self.h, int(snaps), code))

    def backward_plan(self):
        """(method, snaps, recomputed forward steps): method 0 boundary frames,
        1 uniform snapshots, 2 binomial checkpoints."""
        m, sn, rec = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        self._chk(self.lib.fwi_hip_backward_plan(self.h, ctypes.byref(m), ctypes.byref(sn), ctypes.byref(rec)))
        return m.value, sn.value, rec.value

    def set_concurrent(self, on):
        """General strips on a second stream beside the interior launch."""
        self._chk(self.lib.fwi_hip_set_concurrent(self.h, int(bool(on))))

    def set_interior_wgs(self, wgs):
        """Interior workgroups per CU beside the general strips (0: no cap)."""
        self._chk(self.lib.fwi_hip_set_interior_wgs(self.h, int(wgs)))

    def set_pass_shape(self, pass_, rows=8, interior_wgs=3):
        """Tile shape of the forward (0) or backward (1) pass: rows per wave (4 or 8)
        and interior workgroups per CU beside the general strips."""
        self._chk(self.lib.fwi_hip_set_pass_shape(self.h, int(pass_), int(rows), int(interior_wgs)))

    def set_fwd_mixed(self, on):
        """The forward step as one launch of interior and general strips."""
        self._chk(self.lib.fwi_hip_set_fwd_mixed(self.h, int(bool(on))))

    def set_bwd_mixed(self, on):
        """The fused backward step as one launch (general tiles imaged one step late)."""
        self._chk(self.lib.fwi_hip_set_bwd_mixed(self.h, int(bool(on))))

    def set_bwd_fused(self, on):
        """One fused kernel per backward step (adjoint + reconstruction + imaging)."""
        self._chk(self.lib.fwi_hip_set_bwd_fused(self.h, int(bool(on))))

    def set_bwd_overlap(self, on):
        """Adjoint step k beside imaging step k+1 on a second stream."""
        self._chk(self.lib.fwi_hip_set_bwd_overlap(self.h, int(bool(on))))

    def set_bwd_chunk(self, shots):
        """Shots per backward pass (0: all shots in one pass; -1: the library's choice by cache fit)."""
        self._chk(self.lib.fwi_hip_set_bwd_chunk(self.h, int(shots)))

    def set_bwd_general_rows(self, rows):
        """Rows per wave of the fused backward step's general side under 4-row strips (4 or 8)."""
        self._chk(self.lib.fwi_hip_set_bwd_general_rows(self.h, int(rows)))

    def set_bwd_lanes(self, lanes):
        """Backward chunks in flight: 2 runs chunks in pairs on two stream sets, 1 one after another."""
        self._chk(self.lib.fwi_hip_set_bwd_lanes(self.h, int(lanes)))

    def set_imaging(self, jump=1, k0=0, k1=2**31 - 1):
        """Image only the reconstruction steps k0 <= k <= k1 with (k - k0) % jump == 0."""
        self._chk(self.lib.fwi_hip_set_imaging(self.h, int(jump), int(k0), int(k1)))

    def set_graph(self, steps):
        self._chk(self.lib.fwi_hip_set_graph(self.h, int(steps)))

    def set_timing(self, on):
        self._chk(self.lib.fwi_hip_set_timing(self.h, int(on)))

    def kernel_time(self, which):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._chk(self.lib.fwi_hip_kernel_time(self.h, which, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def read_state(self, what, shot):
        g = self.geom
        if what <= 3:
            out = np.empty((g.pnz, g.pnx), np.float32)
        else:
            out = np.empty((self.shots[shot].nt, 6 * g.nx + 6 * g.nz), np.float32)
        self._chk(self.lib.fwi_hip_read_state(self.h, what, shot, _p(out)))
        return out


def cpml_coe(npml, dx, dt, max_f=MAX_F, max_v=MAX_V, coef_r=COEF_R):
    lib = load()
    a = np.zeros(2 * npml + 1, np.float32)
    b = np.zeros(2 * npml + 1, np.float32)
    rc = lib.fwi_hip_cpml_coe(_p(a), _p(b), npml, ctypes.c_float(dx), ctypes.c_float(dt), ctypes.c_float(max_f),
                              ctypes.c_float(max_v), ctypes.c_float(coef_r))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())
    return a, b


def plan_memory(geom, max_shots, max_nt, max_nr, hbm_budget, storage="fp32"):
    """(method, snaps, bytes) a context takes under set_revolve(-1) within the
    budget (no device call)."""
    lib = load()
    d = Desc(make_grid(geom), max_shots, max_nt, max_nr, 0, 0, hbm_budget, None)
    m, sn, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    code = {"fp32": 0, "fp16": 1, "bf16": 2}[storage]
    rc = lib.fwi_hip_plan_memory(ctypes.byref(d), code, ctypes.byref(m), ctypes.byref(sn), ctypes.byref(b))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())
    return m.value, sn.value, b.value


def check_model(geom, vel):
    """The host model check of set_model (no device call): raises FwiError
    (EINVAL) naming the first non-finite, non-positive or unstable cell."""
    lib = load()
    g = make_grid(geom)
    v = np.ascontiguousarray(vel, np.float32)
    if v.size != geom.nx * geom.nz:
        raise ValueError(f"model has {v.size} cells, the grid {geom.nx * geom.nz}")
    cell = ctypes.c_longlong()
    rc = lib.fwi_hip_check_model(ctypes.byref(g), v.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cell))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())


def revolve_schedule(nt, snaps):
    """The binomial schedule: (actions as an (n, 4) array of op, a, b, slot;
    forward steps).  op 0 advance a -> b, 1 snapshot at a, 2 restore at a,
    3 turn at a."""
    lib = load()
    n, fs = ctypes.c_int(), ctypes.c_longlong()
    rc = lib.fwi_hip_revolve_schedule(int(nt), int(snaps), None, 0, ctypes.byref(n), ctypes.byref(fs))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())
    acts = np.zeros((n.value, 4), np.int32)
    rc = lib.fwi_hip_revolve_schedule(int(nt), int(snaps), _p(acts), n.value, ctypes.byref(n), ctypes.byref(fs))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())
    return acts, fs.value


def device_count():
    n = ctypes.c_int(0)
    load().fwi_hip_device_count(ctypes.byref(n))
    return n.value


class DevArray:
    """A float32 device buffer owned through the library's own HIP runtime
    (callers must not mix in a second HIP runtime, e.g. torch's bundled one,
    in the same process)."""

    def __init__(self, shape, init=None):
        self.lib = load()
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = ctypes.c_void_p()
        rc = self.lib.fwi_hip_dmalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes))
        if rc:
            raise FwiError(rc, self.lib.fwi_hip_last_error(None).decode())
        if init is not None:
            self.upload(init)

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.size * 4 == self.nbytes
        rc = self.lib.fwi_hip_memcpy(self.ptr, _p(a), ctypes.c_size_t(self.nbytes), 1)
        if rc:
            raise FwiError(rc, self.lib.fwi_hip_last_error(None).decode())

    def numpy(self):
        out = np.empty(self.shape, np.float32)
        rc = self.lib.fwi_hip_memcpy(_p(out), self.ptr, ctypes.c_size_t(self.nbytes), 2)
        if rc:
            raise FwiError(rc, self.lib.fwi_hip_last_error(None).decode())
        return out

    def __del__(self):
        try:
            if self.ptr:
                self.lib.fwi_hip_dfree(self.ptr)
                self.ptr = ctypes.c_void_p()
        except Exception:
            pass


def _dp(t):
    return t.ptr if isinstance(t, DevArray) else ctypes.c_void_p(t.data_ptr())

def op_pml(geom, p, u, v, ux, vz, u0, u1, v1, ux0, ux1, vz1):
    """a_lap_opt_2d_pml on torch device tensors (in place)."""
    lib = load()
    gr = make_grid(geom)
    rc = lib.fwi_hip_op_pml(ctypes.byref(gr), *[_dp(t) for t in (p, u, v, ux, vz, u0, u1, v1, ux0, ux1, vz1)],
                            ctypes.c_void_p(0))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())


def op_pml_noalloc(geom, p, u, v, ux, vz, u0, u1, v1, ux0, ux1, vz1):
    """a_lap_opt_2d_pml like op_pml, with the CPML tables in the library's
    static device array instead of a fresh allocation."""
    lib = load()
    gr = make_grid(geom)
    rc = lib.fwi_hip_op_pml_noalloc(ctypes.byref(gr), *[_dp(t) for t in (p, u, v, ux, vz, u0, u1, v1, ux0, ux1,
                                                                          vz1)], ctypes.c_void_p(0))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())


def op_interior(geom, p, u, v, ux, vz):
    lib = load()
    gr = make_grid(geom)
    rc = lib.fwi_hip_op_interior(ctypes.byref(gr), *[_dp(t) for t in (p, u, v, ux, vz)], ctypes.c_void_p(0))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())


def time_integral(pnew, pnow, pold, pxx, pzz, ekp):
    lib = load()
    m, n = pnow.shape
    rc = lib.fwi_hip_time_integral(*[_dp(t) for t in (pnew, pnow, pold, pxx, pzz, ekp)], ctypes.c_size_t(m),
                                   ctypes.c_size_t(n), ctypes.c_void_p(0))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())


def write_bdr(geom, ud, lr, pf):
    lib = load()
    gr = make_grid(geom)
    rc = lib.fwi_hip_write_bdr(ctypes.byref(gr), _dp(ud), _dp(lr), _dp(pf), ctypes.c_void_p(0))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())


def write_bdr_fs_no(geom, ud, lr, pf):
    lib = load()
    gr = make_grid(geom)
    rc = lib.fwi_hip_write_bdr_fs_no(ctypes.byref(gr), _dp(ud), _dp(lr), _dp(pf), ctypes.c_void_p(0))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())


def model_ext_fs_no(model, en1, en2):
    """model[n2][n1] extended to [en2][en1] (host)."""
    lib = load()
    m = np.ascontiguousarray(model, np.float32)
    out = np.empty((en2, en1), np.float32)
    rc = lib.fwi_hip_model_ext_fs_no(m.ctypes.data_as(ctypes.c_void_p), m.shape[1], m.shape[0],
                                     out.ctypes.data_as(ctypes.c_void_p), en1, en2)
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())
    return out


def read_bdr(geom, ud, lr, pb):
    lib = load()
    gr = make_grid(geom)
    rc = lib.fwi_hip_read_bdr(ctypes.byref(gr), _dp(ud), _dp(lr), _dp(pb), ctypes.c_void_p(0))
    if rc:
        raise FwiError(rc, lib.fwi_hip_last_error(None).decode())
