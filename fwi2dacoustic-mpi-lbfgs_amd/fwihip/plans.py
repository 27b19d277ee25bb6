"""ctypes mirror of the plan structs in include/fwi_plan.h (the reference's
header.h:21-185 layout) and of the drop-in driver's entry points
(grad4_1shot, grad_all_shots, their settings), for Python callers of the
drop-in boundary: the tests and the benchmark."""
import ctypes

import numpy as np

from fwihip import lib as L

c_size_t, c_float, c_int = ctypes.c_size_t, ctypes.c_float, ctypes.c_int
PF = ctypes.POINTER(c_float)
PPF = ctypes.POINTER(PF)
PI = ctypes.POINTER(c_int)


class FwiBase(ctypes.Structure):
    _fields_ = [(n, c_size_t) for n in ("dim", "mm", "npml", "nx", "nz", "nt", "nt_mem", "free_surface")] + \
               [(n, c_float) for n in ("dx", "dz", "dt")]


class FwiShape(ctypes.Structure):
    _fields_ = [(n, c_size_t) for n in ("mnx", "mnz", "pnx", "pnz", "unx", "unz", "vnx", "vnz")]


class FwiSpecs(ctypes.Structure):
    _fields_ = [("base", ctypes.POINTER(FwiBase)), ("f", ctypes.POINTER(FwiShape)), ("b", ctypes.POINTER(FwiShape))]


class FwiModel(ctypes.Structure):
    _fields_ = [("md", PPF), ("kp", PPF), ("ekp", PPF), ("btry", PF), ("alpha_btry", c_float), ("ibtry", PI)]


class FwiMsf(ctypes.Structure):
    _fields_ = [(n, c_size_t) for n in ("nt", "nr", "jump", "rectx", "rectt", "times")] + \
               [("alpha_fbt0", c_int), ("alpha_fbt1", c_int)] + \
               [(n, c_float) for n in ("dr", "dt", "os0", "os1", "alpha_msf")] + \
               [(n, PPF) for n in ("obs0", "obs", "obs_t", "syn", "syn0", "syn_t")] + \
               [("fbt", PF), ("val", PF), ("ios0", c_int), ("ios1", c_int), ("ifbt", PI)]


class FwiSou(ctypes.Structure):
    _fields_ = [("iarea", c_float), ("xyz_sr", PF), ("ixyz_sr", PI), ("wav0", PF), ("wav", PF), ("res", PPF)]


class FwiGrad(ctypes.Structure):
    _fields_ = [("id", c_size_t)] + [(n, c_float) for n in ("end_t", "start_t", "alpha_grad", "perc")] + \
               [(n, c_size_t) for n in ("jump", "rectx1", "rectz1", "repeat1", "rectx2", "rectz2", "repeat2")] + \
               [("body", PF)]


class BdState(ctypes.Structure):
    _fields_ = [("path_bd_wr", ctypes.c_char * 300), ("bc_ud", ctypes.POINTER(PPF)),
                ("bc_lr", ctypes.POINTER(PPF)), ("offset", c_int)]


class FwiPlan(ctypes.Structure):
    _fields_ = [("grid", ctypes.POINTER(FwiSpecs)), ("model", ctypes.POINTER(FwiModel)),
                ("fw", ctypes.c_void_p), ("bw", ctypes.c_void_p), ("sou", ctypes.POINTER(FwiSou)),
                ("msf", ctypes.POINTER(FwiMsf)), ("pml", ctypes.c_void_p), ("coe", ctypes.c_void_p),
                ("path", ctypes.c_void_p), ("grad", ctypes.POINTER(FwiGrad)), ("bd", BdState)]


class RwiModel(ctypes.Structure):
    _fields_ = [(n, PPF) for n in ("m0", "m1", "kp_total", "kp_0", "ekp_total", "ekp_0")] + \
               [("btry", PF), ("alpha_btry", c_float), ("ibtry", PI)]


class RwiBdState(ctypes.Structure):
    _fields_ = [("path_bd0_wr", ctypes.c_char * 300), ("path_bd1_wr", ctypes.c_char * 300)] + \
               [(n, ctypes.POINTER(PPF)) for n in ("bc_ud_inc", "bc_lr_inc", "bc_ud_per", "bc_lr_per")] + \
               [("offset", c_int)]


class RwiPlan(ctypes.Structure):
    _fields_ = [("grid", ctypes.POINTER(FwiSpecs)), ("model", ctypes.POINTER(RwiModel))] + \
               [(n, ctypes.c_void_p) for n in ("fw_inc", "bw_inc", "fw_per", "bw_per")] + \
               [("sou", ctypes.POINTER(FwiSou)), ("msf", ctypes.POINTER(FwiMsf))] + \
               [(n, ctypes.c_void_p) for n in ("pml_inc", "pml_per", "coe", "path")] + \
               [("grad", ctypes.POINTER(FwiGrad)), ("bd", RwiBdState)]


class EotPar(ctypes.Structure):
    """eot_msf_par (include/fwi_msf.h), the parameters of grad4_set_misfit's
    entropic-OT misfit."""
    _fields_ = [("nt", c_int), ("nr", c_int), ("max_iter", c_int), ("M", PPF), ("u", PF), ("v", PF),
                ("eps", c_float), ("tor", c_float), ("dt", c_float)]


def lib():
    h = L.load()
    h.fwi_plan_alloc.restype = ctypes.POINTER(FwiPlan)
    h.fwi_plan_alloc.argtypes = [c_int] * 6 + [c_float] * 3
    h.fwi_plan_free.argtypes = [ctypes.POINTER(FwiPlan)]
    h.fwi_plan_free.restype = None
    h.grad4_1shot.argtypes = [ctypes.POINTER(FwiPlan), c_int]
    h.grad4_1shot.restype = None
    h.grad_all_shots.argtypes = [ctypes.POINTER(ctypes.POINTER(FwiPlan)), c_int, c_int]
    h.grad_all_shots.restype = c_int
    h.grad4_last_error.restype = ctypes.c_char_p
    h.grad4_phase_ms.argtypes = [ctypes.POINTER(ctypes.c_double), c_int]
    h.grad4_phase_ms.restype = c_int
    h.grad4_set_checkpointing.argtypes = [c_int, c_int]
    h.grad4_set_checkpointing.restype = c_int
    h.fwi_read_f32.argtypes = [ctypes.c_char_p, PF, c_size_t, c_int]
    h.fwi_read_f32.restype = c_int
    h.fwi_write_f32.argtypes = [ctypes.c_char_p, PF, c_size_t]
    h.fwi_write_f32.restype = c_int
    h.fwi_plan_read_inputs.argtypes = [ctypes.POINTER(FwiPlan), ctypes.c_char_p, c_int, ctypes.c_char_p, c_int,
                                       c_int, c_float]
    h.fwi_plan_read_inputs.restype = c_int
    h.rwi_plan_read_inputs.argtypes = [ctypes.POINTER(RwiPlan), ctypes.c_char_p, c_int, ctypes.c_char_p,
                                       ctypes.c_char_p, ctypes.c_char_p, c_int, c_int, c_float]
    h.rwi_plan_read_inputs.restype = c_int
    h.fwi_trcinp.argtypes = [PPF, c_int, c_float, PPF, c_int, c_float, c_int]
    h.fwi_trcinp.restype = c_int
    h.fwi_io_last_error.restype = ctypes.c_char_p
    h.grad4_set_residual_mode.argtypes = [c_int]
    h.grad4_set_residual_mode.restype = c_int
    h.grad4_set_misfit.argtypes = [c_int, ctypes.c_void_p, c_float, c_float]
    h.grad4_set_misfit.restype = c_int
    h.grad4_set_bandpass.argtypes = [c_int, c_int, c_int, c_int, c_float, c_float]
    h.grad4_set_bandpass.restype = c_int
    h.grad4_set_fbt_mute.argtypes = [c_int, c_float, c_float, c_int]
    h.grad4_set_fbt_mute.restype = c_int
    h.rwi_plan_alloc.restype = ctypes.POINTER(RwiPlan)
    h.rwi_plan_alloc.argtypes = [c_int] * 6 + [c_float] * 3
    h.rwi_plan_free.argtypes = [ctypes.POINTER(RwiPlan)]
    h.rwi_plan_free.restype = None
    h.grad4_1shot_rfl.argtypes = [ctypes.POINTER(RwiPlan), c_int]
    h.grad4_1shot_rfl.restype = None
    h.grad_all_shots_rfl.argtypes = [ctypes.POINTER(ctypes.POINTER(RwiPlan)), c_int, c_int, PF]
    h.grad_all_shots_rfl.restype = c_int
    return h


def rows(a):
    """row-pointer array (float **) over a C-contiguous float32 2-D numpy array"""
    assert a.dtype == np.float32 and a.flags.c_contiguous and a.ndim == 2
    base = a.ctypes.data
    return (PF * a.shape[0])(*[ctypes.cast(base + i * a.strides[0], PF) for i in range(a.shape[0])])


def view(ptr, shape):
    """numpy view of a float slab (row 0 of a row-pointer array)."""
    n = int(np.prod(shape))
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(c_float * n)).contents).reshape(shape)


def make_plan(h, g, shot, obs0, vel, jump=1):
    """A plan as fwi_init would leave it (fwi_2d.c:341-371), without files."""
    nt0 = obs0.shape[1]
    p = h.fwi_plan_alloc(g.nx, g.nz, g.npml, shot.nt, jump, shot.nr, g.dx, g.dz, g.dt)
    assert p, "fwi_plan_alloc failed"
    P = p.contents
    view(P.model.contents.md[0], (g.nz, g.nx))[:] = vel
    sou = P.sou.contents
    ix = np.ctypeslib.as_array(ctypes.cast(sou.ixyz_sr, ctypes.POINTER(c_int * (2 + 2 * shot.nr))).contents)
    ix[0], ix[1] = shot.isx, shot.isz
    ix[2::2] = shot.irx
    ix[3::2] = shot.irz
    view(sou.wav0, (shot.nt,))[:] = shot.wav
    msf = P.msf.contents
    assert msf.nt == nt0 and msf.nr == shot.nr and msf.jump == jump
    view(msf.obs0[0], (shot.nr, nt0))[:] = obs0
    return p


def make_rwi_plan(h, g, shot, obs0, m0, m1, jump=1):
    """An RWI plan as rwi_init would leave it (fwi_2d.c:398-719), without files."""
    nt0 = obs0.shape[1]
    p = h.rwi_plan_alloc(g.nx, g.nz, g.npml, shot.nt, jump, shot.nr, g.dx, g.dz, g.dt)
    assert p, "rwi_plan_alloc failed"
    P = p.contents
    view(P.model.contents.m0[0], (g.nz, g.nx))[:] = m0
    view(P.model.contents.m1[0], (g.nz, g.nx))[:] = m1
    sou = P.sou.contents
    ix = np.ctypeslib.as_array(ctypes.cast(sou.ixyz_sr, ctypes.POINTER(c_int * (2 + 2 * shot.nr))).contents)
    ix[0], ix[1] = shot.isx, shot.isz
    ix[2::2] = shot.irx
    ix[3::2] = shot.irz
    view(sou.wav0, (shot.nt,))[:] = shot.wav
    msf = P.msf.contents
    assert msf.nt == nt0 and msf.nr == shot.nr and msf.jump == jump
    view(msf.obs0[0], (shot.nr, nt0))[:] = obs0
    return p
