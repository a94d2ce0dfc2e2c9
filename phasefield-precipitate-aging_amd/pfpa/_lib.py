"""Loader and ctypes prototypes of libpfpa.so (the C-ABI of include/pfpa.h).

The library is the product: there is no fallback.  If it has not been built,
or no GPU is visible for a device call, the call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfpa.so")

PF_OK = 0
PF_ERR_ARG = 1
PF_ERR_HIP = 2
PF_ERR_OVERFLOW = 3
PF_ERR_NO_DEVICE = 4
PF_BC_NOFLUX = 0
PF_BC_PERIODIC = 1

_d = ctypes.c_double
_i = ctypes.c_int
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


class PfParams(ctypes.Structure):
    _fields_ = [("dt", _d), ("DCr", _d * 2), ("DNb", _d * 2), ("kappa", _d * 2),
                ("omega", _d * 2), ("Lmob", _d * 2), ("alpha", _d), ("mask", _d * 9)]


class PfNucParams(ctypes.Structure):
    _fields_ = [("sigma_del", _d), ("sigma_lav", _d), ("unit_a", _d), ("ifce_width", _d),
                ("dx", _d), ("dy", _d), ("dz", _d), ("dt", _d), ("seed", ctypes.c_uint64),
                ("call", ctypes.c_uint64), ("instance", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class PfStepItem(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("bc", ctypes.c_int),
                ("prm", ctypes.POINTER(PfParams)), ("old_fields", ctypes.c_void_p * 4),
                ("new_fields", ctypes.c_void_p * 4)]


class PfMobilityFields(ctypes.Structure):
    _fields_ = [("conc", _vp * 6), ("mob_Cr", _vp * 8), ("mob_Nb", _vp * 8),
                ("phi_del", _vp), ("phi_lav", _vp)]


class PfError(RuntimeError):
    pass


# name -> (restype, argtypes)
PROTOTYPES = {
    "pf_ctx_create": (_i, [ctypes.POINTER(_vp)]),
    "pf_ctx_destroy": (None, [_vp]),
    "pf_last_error": (ctypes.c_char_p, []),
    "pf_device_count": (_i, []),
    "pf_ctx_set_tile": (_i, [_vp, _i, _i, _i]),
    "pf_ctx_set_storage_offset": (_i, [_vp, ctypes.c_double, ctypes.c_double]),
    "pf_step_f64": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 +
                    [ctypes.POINTER(PfNucParams), _vp]),
    "pf_step_f32": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 +
                    [ctypes.POINTER(PfNucParams), _vp]),
    "pf_step2_f64": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 + [_vp]),
    "pf_step2_f32": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 + [_vp]),
    "pf_step3_f64": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 + [_vp]),
    "pf_step3_f32": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 + [_vp]),
    "pf_step4_f64": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 + [_vp]),
    "pf_step4_f32": (_i, [_vp, _i, _i, _i, ctypes.POINTER(PfParams)] + [_vp] * 8 + [_vp]),
    "pf_ctx_set_tile2": (_i, [_vp, _i, _i]),
    "pf_default_rows2": (_i, [_i, _i, _i]),
    "pf_ctx_set_tile3": (_i, [_vp, _i, _i]),
    "pf_default_rows3": (_i, [_i, _i, _i]),
    "pf_ctx_set_tile4": (_i, [_vp, _i, _i]),
    "pf_default_rows4": (_i, [_i, _i, _i]),
    "pf_nucleation_f64": (_i, [_vp, _i, _i] + [_vp] * 4 + [ctypes.POINTER(PfNucParams), _vp]),
    "pf_nucleation_f32": (_i, [_vp, _i, _i] + [_vp] * 4 + [ctypes.POINTER(PfNucParams), _vp]),
    "pf_nucleation_status": (_i, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_i)]),
    "pf_summarize_f64": (_i, [_vp, _i, _i] + [_vp] * 5 + [_vp]),
    "pf_summarize_f32": (_i, [_vp, _i, _i] + [_vp] * 5 + [_vp]),
    "pf_dataviz_f64": (_i, [_i, _i] + [_vp] * 6 + [_vp]),
    "pf_energy_f64": (_i, [_vp, _i, _i, _i, _d, _d, _dp, _dp, _d] + [_vp] * 7 + [_vp]),
    "pf_energy_f32": (_i, [_vp, _i, _i, _i, _d, _d, _dp, _dp, _d] + [_vp] * 7 + [_vp]),
    "pf_thermo_eval": (_i, [_i, _i, _vp, _vp, _vp]),
    "pf_philox_eval": (_i, [_i, _vp, _vp, _vp]),
    "pf_default_rows": (_i, [_i, _i]),
    "pf_step_loop_f64": (_i, [_vp, _i, ctypes.POINTER(PfStepItem), _vp]),
    "pf_step2_loop_f64": (_i, [_vp, _i, ctypes.POINTER(PfStepItem), _vp]),
    "pf_step3_loop_f64": (_i, [_vp, _i, ctypes.POINTER(PfStepItem), _vp]),
    "pf_step4_loop_f64": (_i, [_vp, _i, ctypes.POINTER(PfStepItem), _vp]),
    "pf_mobility_divergence_f64": (_i, [_i, _i, _d, _d, ctypes.POINTER(PfMobilityFields), _vp,
                                        _vp, _vp]),
    "pf_host_p": (_d, [_d]),
    "pf_host_pPrime": (_d, [_d]),
    "pf_host_g": (_d, [_i, _d, _d]),
    "pf_host_D": (_d, [_i, _i, _d, _d]),
    "pf_host_fict": (None, [_d, _d, _d, _d, _d, _dp]),
    "pf_host_xe": (_d, [_i, _i]),
    "pf_host_nucleation_probability": (None, [_d] * 10 + [_dp, _dp]),
    "pf_host_max_effective_diffusivity": (_d, [_dp, _dp]),
    "pf_host_step_coefficients": (_i, [ctypes.POINTER(PfParams), _dp]),
}

_lib = None


def lib():
    """Load libpfpa.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PfError("libpfpa.so not built (%s): run __graft_entry__.build() or "
                          "python phasefield-precipitate-aging_amd/build.py" % LIB_PATH)
        # PyTorch ships its own HIP runtime (same soname as /opt/rocm's).  Load
        # it first, so that libpfpa binds to that one runtime: loaded the other
        # way round the process would hold two, and torch could not see the GPU
        # or share streams and allocations with the library.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != PF_OK:
        msg = lib().pf_last_error()
        msg = msg.decode() if msg else ""
        raise PfError("%s failed (code %d): %s" % (what or "libpfpa call", rc, msg))
    return rc
