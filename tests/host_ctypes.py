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
