"""ctypes bindings of the host trace conditioning (include/fwi_sigproc.h) and
misfit pool (include/fwi_msf.h) in libfwi_hip.so, and of the reference's own
Butterworth/mute sources compiled under oracle/_ref (absent when the
reference sources were not there to build it)."""
import ctypes
import os

import numpy as np

from fwihip import lib as L

c_int, c_float, c_bool = ctypes.c_int, ctypes.c_float, ctypes.c_bool
PF = ctypes.POINTER(c_float)
PPF = ctypes.POINTER(PF)
VP = ctypes.c_void_p

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SIGPROC = os.path.join(ROOT, "oracle", "_ref", "libref_sigproc.so")
REF_RANK = os.path.join(ROOT, "oracle", "_ref", "libref_rank.so")
PI = ctypes.POINTER(c_int)
PPPF = ctypes.POINTER(PPF)


class ParXcorr1d(ctypes.Structure):
    _fields_ = [("nr", c_int), ("nt", c_int), ("npad4Ab", c_int), ("conv_how_to_slice", ctypes.c_char_p)]


class Epsdf1(ctypes.Structure):
    _fields_ = [(n, c_int) for n in ("nr", "nt", "nmodel", "npad4Ax", "npad4Atr", "niter")] + \
               [("conv_how_to_slice", ctypes.c_char_p), ("reg_type", ctypes.c_char_p)] + \
               [(n, c_float) for n in ("tor0", "eps0", "Tdepressor")]


class EotPar(ctypes.Structure):
    _fields_ = [("nt", c_int), ("nr", c_int), ("max_iter", c_int), ("M", PPF), ("u", PF), ("v", PF),
                ("eps", c_float), ("tor", c_float), ("dt", c_float)]


class EotLogPar(ctypes.Structure):
    _fields_ = [("bar", c_float), ("M", PPF), ("u", PF), ("v", PF), ("logp", PF), ("logq", PF),
                ("delta_u", PF), ("delta_v", PF), ("vmax0", c_float)]


class DiwPar(ctypes.Structure):
    _fields_ = [("nl", c_int), ("nw", c_int), ("eps", c_float), ("alpha_win", c_float), ("sim_barrier", c_float),
                ("lag_barrier", c_float)]


class SoftRankingVars(ctypes.Structure):
    _fields_ = [("sorted", PF), ("sol", PF), ("soft_rank", PF), ("hard_rank", PI), ("perm", PI), ("target", PI)]


def _sig(f, argtypes, restype=c_int):
    f.argtypes = argtypes
    f.restype = restype


TAPER_ARGS = [PPF, PF, c_int, c_int, c_float, c_float, c_float, c_int]
TANH_ARGS = [PPF, PF, c_int, c_int, c_float, c_float, c_float, c_float, PPF]
SIGPROC = {
    "sfbandpass": [PF, c_int, c_int, c_float, c_bool, c_int, c_int, c_float, c_float, PF],
    "sfbandpass_adjoint": [PF, c_int, c_int, c_float, c_bool, c_int, c_int, c_float, c_float, PF],
    "fbt_muter": [PF, PF, c_int, c_int, c_float, c_float, c_float, c_bool],
    "win_apply": [PF, PF, c_int, c_int, c_float, c_float, PF],
    "upper_off_srcood": TANH_ARGS,
    "lower_off_srcood": TANH_ARGS,
    "linear_upper_off": TAPER_ARGS,
    "linear_lower_off": TAPER_ARGS,
    "quadratic_upper_off": TAPER_ARGS,
    "quadratic_lower_off": TAPER_ARGS,
}


def lib():
    h = L.load()
    for name, args in SIGPROC.items():
        _sig(getattr(h, name), args)
    _sig(h.fwi_sigproc_last_error, [], ctypes.c_char_p)
    _sig(h.L2norm, [PPF, PPF, c_int, c_int, PF, c_float, c_float, PF, PPF])
    _sig(h.NCCnorm, [PPF, PPF, c_int, c_int, PF, c_float, c_float, PF, PPF])
    _sig(h.XCORR1D, [PPF, PPF, PF, c_float, c_float, VP, PF, PPF])
    _sig(h.MF1D, [PPF, PPF, VP, PF, PPF])
    _sig(h.eot_msf_init, [ctypes.POINTER(EotPar)])
    _sig(h.eot_msf_destroy, [ctypes.POINTER(EotPar)], None)
    _sig(h.eot_log_msf_init, [ctypes.POINTER(EotPar), ctypes.POINTER(EotLogPar)])
    _sig(h.entropy_ot, [PF, PF, ctypes.POINTER(EotPar), PF, PF])
    _sig(h.entropy_ot_log, [PF, PF, ctypes.POINTER(EotPar), ctypes.POINTER(EotLogPar), PF, PF])
    _sig(h.entropy_ot_2d_trace_by_trace, [PPF, PPF, ctypes.POINTER(EotPar), PF, PPF])
    _sig(h.entropy_ot_log_2d_trace_by_trace, [PPF, PPF, ctypes.POINTER(EotPar), ctypes.POINTER(EotLogPar), PF, PPF])
    _sig(h.fwi_msf_last_error, [], ctypes.c_char_p)
    PSR = ctypes.POINTER(SoftRankingVars)
    _sig(h.gaussian_win, [ctypes.c_size_t, c_float, PF], None)
    _sig(h.lcc_err, [PPF, PPF, PF, PF, c_float, c_int, c_int, c_int, c_int, c_float, PPPF])
    for name in ("ac_3pts_fw", "ac_3pts_bw", "ac_5pts_fw", "ac_5pts_bw"):
        _sig(getattr(h, name), [PPF, PPF, c_int, c_int], None)
    _sig(h.bt_fw_3pts, [PPF, c_int, c_int, PF], None)
    _sig(h.bt_fw_5pts, [PPF, PPF, c_int, c_int, PF], None)
    _sig(h.bt_bw_5pts, [PPF, PPF, c_int, c_int, PF], None)
    _sig(h.lcc_diw_msf, [PPF, PPF, PF, PF, c_float, c_int, c_int, c_int, c_int, c_float, c_float, c_float, PF, PPF])
    _sig(h.sorting_flt_argsort, [c_int, PF, c_int, PI, PI, PF])
    _sig(h.isotonic_l2, [c_int, PF, PF, PI])
    _sig(h.soft_ranking_init, [c_int, PSR], None)
    _sig(h.soft_ranking_free, [PSR], None)
    _sig(h.soft_ranking, [c_int, PF, c_float, PSR])
    _sig(h.soft_ranking_msf, [PF, PF, c_int, c_float, PSR, PSR, PF, PF])
    _sig(h.soft_ranking_2d_trace_by_trace, [PPF, PPF, PF, c_float, c_float, ctypes.POINTER(PSR), c_float, c_int, c_int,
                                            PF, PPF])
    _sig(h.fwi_msf_rank_last_error, [], ctypes.c_char_p)
    return h


def ref_rank_lib():
    """The reference's soft_ranking.c (argsort, isotonic regression), or None."""
    if not os.path.exists(REF_RANK):
        return None
    h = ctypes.CDLL(REF_RANK)
    _sig(h.sorting_flt_argsort, [c_int, PF, c_int, PI, PI, PF])
    _sig(h.isotonic_l2, [c_int, PF, PF, PI], None)
    return h


def ref_lib():
    """The reference's butterworth_filter.c + mute_fun.c (void functions), or None."""
    if not os.path.exists(REF_SIGPROC):
        return None
    h = ctypes.CDLL(REF_SIGPROC)
    for name, args in SIGPROC.items():
        if name != "sfbandpass_adjoint":
            _sig(getattr(h, name), args, None)
    return h


def fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(PF)


def rows(a):
    """row-pointer array (float **) over a C-contiguous float32 2-D array"""
    assert a.dtype == np.float32 and a.flags.c_contiguous and a.ndim == 2
    base = a.ctypes.data
    return (PF * a.shape[0])(*[ctypes.cast(base + i * a.strides[0], PF) for i in range(a.shape[0])])
