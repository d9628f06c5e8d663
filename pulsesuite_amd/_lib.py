"""ctypes binding of libpss.so (the C ABI declared in include/pss.h).

The library is the only compute path: importing this module on a machine
without the built library raises immediately, and every call checks its
status and raises RuntimeError with the library's message.
"""

import contextlib
import ctypes as C
import gc
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpss.so")

# mb flag bits (pss.h)
OPTICS = 1 << 0
EXCITONS = 1 << 1
EHS = 1 << 2
SCREENED = 1 << 3
PHONON = 1 << 4
DCTRANS = 1 << 5
LF = 1 << 6
DIAGDPH = 1 << 8
OFFDIAGDPH = 1 << 9
RECOMB = 1 << 10
IGNOREWIRE = 1 << 12
LORENTZDELTA = 1 << 14

FP64 = 0
MIXED = 1

# ABI version the signatures below describe (pss_version() of include/pss.h)
ABI_VERSION = "pss 0.3"

_vp = C.c_void_p
_d = C.c_double
_i = C.c_int32


class WireDesc(C.Structure):
    _fields_ = (
        [("Nk", _i), ("Nr", _i), ("nrr", _i), ("w", _i), ("flags", _i), ("precision", _i), ("pad0", _i), ("pad1", _i)]
        + [(n, _d) for n in ("dt", "L", "gap", "ehint", "area", "dcv", "dkr", "dr_r", "drr", "yw", "zsign",
                             "gam_e", "gam_h", "gam_eh", "Vscale", "N0", "hOph", "hGph", "Gph",
                             "density_max", "gate")]
        + [("Edc", _d), ("rsp", _d * 4), ("pad2", _d * 3)]
        + [(n, _vp) for n in ("kr", "Ee", "Eh", "win", "r", "V0", "chie", "chih", "qe", "qh", "ktab", "xe", "xh",
                              "iqv", "twid", "sa_x", "sa_t", "sa_b", "sa_h", "sa_end", "sa_idx", "sa_dx",
                              "sb_t", "sb_b", "sb_h", "sb_end", "sb_idx", "sb_dx")]
        + [("p", _vp * 3), ("C", _vp * 3), ("D", _vp * 3), ("ctl", _vp), ("on", _vp)]
        + [(n, _vp) for n in ("Exx", "Eyy", "Ezz", "Pxx", "Pyy", "Pzz", "Rho", "ws")]
    )


class Grid(C.Structure):
    _fields_ = (
        [(n, _i) for n in ("Nx", "Ny", "Nz", "npml_x", "npml_y", "npml_z", "precision", "pad1")]
        + [("dt", _d), ("v2", _d), ("mu0", _d)]
        + [(n, _vp) for n in ("kx", "ky", "kz", "bx", "cx", "ikx", "by", "cy", "iky", "bz", "cz", "ikz",
                              "twx", "twy", "twz", "tdev")]
    )


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m pulsesuite_amd._build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "pss_sbe_workspace_bytes": (C.c_int64, [_i, _i, _i, _i]),
        "pss_sbe_tune": (C.c_int, [_i]),
        "pss_cpml_psi_len": (C.c_int64, [C.POINTER(Grid)]),
        "pss_sbe_step": (C.c_int, [_vp, _i, _i, _i, _i, _i, _vp]),
        "pss_sbe_step_p": (C.c_int, [_vp, _i, _i, _i, _i, _i, _i, _i, _vp]),
        "pss_vint": (C.c_int, [_vp, _i, _vp, _i, _d, _d, _d, _d, _vp, _vp]),
        "pss_sbe_stages": (C.c_int, [_vp, _i, _i, _i, _i, _i, _i, _i, _i, _vp]),
        "pss_calc_h": (C.c_int, [_vp, _i, _i, _i, _i, _i, _i, _vp]),
        "pss_sbe_rhs": (C.c_int, [_vp, _i, _i, _i, _i, _i, _i, _vp]),
        "pss_coll_rates": (C.c_int, [_vp, _i, _i, _i, _i, _i, _i, _vp]),
        "pss_qwpol": (C.c_int, [_vp, _i, _i, _i, _i, _i, _vp]),
        "pss_rhs_dense": (C.c_int, [_i, _i] + [_vp] * 10 + [_vp]),
        "pss_sbe_get_state": (C.c_int, [_vp, _i, _vp, _vp]),
        "pss_sbe_set_state": (C.c_int, [_vp, _i, _vp, _vp]),
        "pss_update_e3d": (C.c_int, [C.POINTER(Grid)] + [_vp] * 12 + [_vp]),
        "pss_update_e3d_sample": (C.c_int, [C.POINTER(Grid)] + [_vp] * 12 + [_vp] * 4 + [_vp, _i, _i, _vp, _d, _i,
                                                                                      _vp]),
        "pss_update_b3d": (C.c_int, [C.POINTER(Grid)] + [_vp] * 6 + [_vp]),
        "pss_fft3d": (C.c_int, [C.POINTER(Grid), _vp, _i, _vp]),
        "pss_fft_line": (C.c_int, [C.POINTER(Grid), _vp, _vp, _i, _i, _vp]),
        "pss_fft_line_batch": (C.c_int, [C.POINTER(Grid), C.POINTER(_vp), C.POINTER(_vp), _i, _i, _i, _vp]),
        "pss_ksample": (C.c_int, [C.POINTER(Grid), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pss_b_half": (C.c_int, [C.POINTER(Grid), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, C.POINTER(_d), _vp,
                                 _vp, _vp, _vp, _vp]),
        "pss_cpml_e_curl": (C.c_int, [C.POINTER(Grid), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, C.POINTER(_d),
                                      _vp, _vp, _vp, _vp, _vp]),
        "pss_cpml_half": (C.c_int, [C.POINTER(Grid), _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i,
                                    C.POINTER(_d), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pss_fill_line": (C.c_int, [_vp, _vp, _i, _vp]),
        "pss_mask_half": (C.c_int, [C.POINTER(Grid), _vp, _vp, _vp, _vp, _vp, _vp, _i, C.POINTER(_d), _vp, _vp, _vp,
                                    _vp, _vp]),
        "pss_fft_planes_acc": (C.c_int, [C.POINTER(Grid), C.POINTER(_vp), C.POINTER(_vp), _vp, _vp, _vp, _d, _vp]),
        "pss_fft_pass_acc": (C.c_int, [C.POINTER(Grid), C.POINTER(_vp), C.POINTER(_vp), _i, _i, _i, _i, _vp, _vp,
                                       _vp, _d, _vp]),
        "pss_jline": (C.c_int, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _i, _i, _vp]),
        "pss_elong_k": (C.c_int, [C.POINTER(Grid), _vp, _vp, _vp, _vp, C.POINTER(_vp), _d, _d, _vp]),
        "pss_qw_place": (C.c_int, [_i, _i, _i, _i, _i, C.POINTER(_vp), C.POINTER(C.c_int64), C.POINTER(_vp),
                                   C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_d), C.POINTER(_i), _d, _vp, _vp]),
        "pss_kfilter": (C.c_int, [C.POINTER(Grid), _vp, _d, _d, _d, _vp]),
        "pss_ksample_noalloc": (C.c_int, [C.POINTER(Grid), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pss_pass_timing_arm": (C.c_int, [_i]),
        "pss_pass_timing_read": (C.c_int, [C.POINTER(C.c_float), _i, C.POINTER(_i)]),
        "pss_drive_uniform": (C.c_int, [_vp, _vp, _i, _vp]),
        "pss_add_scalar": (C.c_int, [_vp, _d, _vp]),
        "pss_rk4_stage": (C.c_int, [_i, C.c_int64, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pss_last_error": (C.c_char_p, []),
        "pss_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    ver = lib.pss_version().decode(errors="replace")
    if not ver.startswith(ABI_VERSION):
        raise RuntimeError(f"{LIB_PATH} reports {ver!r}; this binding needs {ABI_VERSION!r}: rebuild the library")
    return lib


lib = _load()

EXPORTS = ("pss_sbe_workspace_bytes", "pss_sbe_tune", "pss_sbe_step", "pss_sbe_step_p", "pss_sbe_stages", "pss_calc_h", "pss_sbe_rhs", "pss_coll_rates",
           "pss_qwpol", "pss_rhs_dense", "pss_sbe_get_state", "pss_sbe_set_state", "pss_vint", "pss_update_e3d", "pss_update_e3d_sample", "pss_update_b3d",
           "pss_fft3d", "pss_fft_line", "pss_fft_line_batch", "pss_fft_pass_acc", "pss_fft_planes_acc", "pss_ksample", "pss_ksample_noalloc", "pss_pass_timing_arm", "pss_pass_timing_read", "pss_b_half", "pss_cpml_e_curl", "pss_cpml_half", "pss_mask_half", "pss_cpml_psi_len", "pss_fill_line", "pss_jline", "pss_elong_k", "pss_qw_place", "pss_kfilter", "pss_rk4_stage", "pss_drive_uniform", "pss_add_scalar",
           "pss_last_error", "pss_version")


def check(rc, what):
    if rc != 0:
        msg = lib.pss_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


@contextlib.contextmanager
def gc_paused():
    """Keep the cyclic garbage collector off until the block ends.

    Wraps every HIP graph capture: a collection that starts inside a capture runs the
    destructors of dead graphs, events and pooled tensors (graph-exec destroy, pool
    release) in the middle of the capture, where they are not allowed, and that has
    aborted the process.  Garbage is left for the next collection after the capture (a
    full collection before each capture cost the serving loop, which captures per solve,
    about a second per 48 requests)."""
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
