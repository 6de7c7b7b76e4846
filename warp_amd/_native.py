"""ctypes binding of libwarpmi.so (include/warpmi.h).

The library is built in-tree by ``python -m warp_amd.build``.  Loading it
does not need a GPU (the conductor set-up entry points run on the host);
every device entry point fails loudly if no HIP device is present.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "lib", "libwarpmi.so")

_i64 = C.c_int64
_d = C.c_double
_P = C.c_void_p
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int64)


class WmIntercepts(C.Structure):
    _fields_ = [
        ("xmmin", _d), ("ymmin", _d), ("zmmin", _d),
        ("dx", _d), ("dy", _d), ("dz", _d),
        ("nx", _i64), ("ny", _i64), ("nz", _i64),
        ("ix", _i64), ("iy", _i64), ("iz", _i64), ("mglevel", _i64),
        ("nxicpt", _i64), ("nyicpt", _i64), ("nzicpt", _i64),
        ("xintercepts", _pd), ("yintercepts", _pd), ("zintercepts", _pd),
        ("xvoltages", _pd), ("yvoltages", _pd), ("zvoltages", _pd),
        ("xcondids", _pi), ("ycondids", _pi), ("zcondids", _pi),
    ]


class WmConductorLists(C.Structure):
    _fields_ = [
        ("levels", _i64),
        ("int_n", _i64), ("int_indx", _pi), ("int_volt", _pd), ("int_istart", _pi),
        ("sub_n", _i64 * 2), ("sub_indx", _pi * 2), ("sub_dels", _pd * 2),
        ("sub_volt", _pd * 2), ("sub_istart", _pi * 2),
    ]


# name -> (restype, argtypes); every entry point listed in include/warpmi.h.
SIGNATURES = {
    "wm_last_error": (C.c_char_p, []),
    "wm_device_info": (C.c_int, [C.c_char_p, _i64, _P]),
    "wm_stream_probe": (C.c_int, [_P, _P, _i64, _i64, _P]),
    "wm_mg3d_create": (C.c_int, [_P, _i64, _i64, _i64, _d, _d, _d, _P, _i64]),
    "wm_mg3d_destroy": (C.c_int, [_P]),
    "wm_mg3d_nlevels": (C.c_int, [_P, _P]),
    "wm_mg3d_level_dims": (C.c_int, [_P, _i64, _P, _P]),
    "wm_mg3d_set_conductors": (C.c_int, [_P, _P]),
    "wm_mg3d_solve": (C.c_int, [_P, _P, _P, _d, _i64, _i64, _d, _i64, _i64, _i64,
                                _i64, _i64, _P, _P, _P]),
    "wm_phi_extrapolate": (C.c_int, [_P, _P, _i64, _P]),
    "wm_mg3d_set_timing": (C.c_int, [_P, _i64]),
    "wm_mg3d_set_sor_form": (C.c_int, [_P, _i64]),
    "wm_mg3d_get_timing": (C.c_int, [_P, _P, _P]),
    "wm_mg3d_timing_fused": (C.c_int, [_P, _P]),
    "wm_mg3d_sorhalfpass": (C.c_int, [_P, _i64, _i64, _P, _P, _d, _i64, _i64, _i64, _P]),
    "wm_mg3d_sorpass": (C.c_int, [_P, _i64, _P, _P, _d, _i64, _i64, _i64, _P]),
    "wm_mg3d_residual": (C.c_int, [_P, _i64, _P, _P, _P, _d, _i64, _i64, _i64, _i64, _P]),
    "wm_mg3d_restrict": (C.c_int, [_P, _i64, _P, _P, _P]),
    "wm_mg3d_expand": (C.c_int, [_P, _i64, _P, _P, _P]),
    "wm_applyboundaryconditions3d": (C.c_int, [_i64, _i64, _i64, _P, _P, _i64, _i64, _P]),
    "wm_setrho3d": (C.c_int, [_P, _i64, _P, _P, _P, _d, _d, _d, _i64, _i64, _i64, _d, _d,
                              _d, _d, _d, _d, _P, _P]),
    "wm_setrho3d_sorted_worksize": (_i64, [_i64, _i64, _i64, _i64]),
    "wm_setrho3d_sorted": (C.c_int, [_P, _i64, _P, _P, _P, _d, _d, _d, _i64, _i64, _i64,
                                     _d, _d, _d, _d, _d, _d, _P, _P]),
    "wm_applyrhoboundaryconditions3d": (C.c_int, [_i64, _i64, _i64, _P, _P, _P]),
    "wm_sete3d": (C.c_int, [_P, _i64, _P, _P, _P, _d, _d, _d, _d, _d, _d, _d, _i64, _i64,
                            _i64, _P, _P, _P, _P]),
    "wm_epush3d": (C.c_int, [_i64, _P, _P, _P, _P, _P, _P, _d, _d, _d, _P]),
    "wm_xpush3d": (C.c_int, [_i64, _P, _P, _P, _P, _P, _P, _P, _d, _P]),
    "wm_gather_push3d": (C.c_int, [_P, _i64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _d,
                                   _d, _d, _d, _d, _d, _d, _i64, _i64, _i64, _d, _d, _d,
                                   _d, _d, _d, _P, _P]),
    "wm_padvnc3d": (C.c_int, [_P, _i64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _d, _d, _d,
                              _d, _d, _d, _d, _i64, _i64, _i64, _d, _d, _d, _d, _d, _d, _P,
                              _i64, _P]),
    "wm_padvnc3d_selfe": (C.c_int, [_P, _i64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _d, _d,
                                    _d, _d, _d, _d, _d, _i64, _i64, _i64, _d, _d, _d, _d, _d,
                                    _d, _P, _i64, _P]),
    "wm_sete3d_from_e": (C.c_int, [_P, _i64, _P, _P, _P, _d, _d, _d, _d, _d, _d, _d, _i64,
                                   _i64, _i64, _P, _P, _P, _P]),
    "wm_condfield_create": (C.c_int, [_P, _i64, _i64, _i64, _P]),
    "wm_condfield_destroy": (C.c_int, [_P]),
    "wm_condfield_efield": (C.c_int, [_P, _P, _d, _d, _d, _P]),
    "wm_condfield_fixefield": (C.c_int, [_P, _P, _P]),
    "wm_condfield_read": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "wm_sete3d_conductor": (C.c_int, [_P, _P, _P, _i64, _i64, _P, _P, _P, _d, _d, _d, _d, _d,
                                      _d, _d, _P, _P, _P, _P]),
    "wm_padvnc3d_conductor": (C.c_int, [_P, _P, _P, _i64, _i64, _P, _P, _P, _P, _P, _P, _P,
                                        _P, _P, _P, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d,
                                        _d, _d, _d, _P, _i64, _P]),
    "wm_getzmmnt_worksize": (_i64, [_i64]),
    "wm_getzmmnt": (C.c_int, [_i64, _P, _P, _P, _P, _P, _P, _P, _i64, _P, _P, _P]),
    "wm_mg3d_reuse_phi": (C.c_int, [_P]),
    "wm_mg3d_last_reused": (C.c_int, [_P, _P]),
    "wm_mg3d_tail_start": (C.c_int, [_P, _P]),
    "wm_gtlchg3d": (C.c_int, [_P, _i64, _i64, _i64, _d, _d, _d, _d, _d, _d, _d, _d, _i64, _i64,
                              _i64, _P, _P, _P]),
    "wm_getese3d_worksize": (_i64, [_i64, _i64]),
    "wm_getese3d": (C.c_int, [_P, _P, _i64, _i64, _i64, _d, _d, _d, _i64, _i64, _P, _P, _P]),
    "wm_getese3d_noalloc": (C.c_int, [_P, _P, _i64, _i64, _i64, _d, _d, _d, _i64, _i64, _P, _P]),
    "wm_unet_conv3": (C.c_int, [_P, _i64, _P, _P, _i64, _i64, _i64, _i64, _P, _P, _d, _P]),
    "wm_unet_conv3_wgrad": (C.c_int, [_P, _i64, _P, _i64, _i64, _i64, _i64, _P, _P, _P]),
    "wm_unet_act_bwd": (C.c_int, [_P, _P, _d, _i64, _P]),
    "wm_unet_pack_size": (_i64, [_i64, _i64]),
    "wm_unet_pack": (C.c_int, [_P, _i64, _i64, _P, _P, _P]),
    "wm_unet_pool": (C.c_int, [_P, _i64, _i64, _i64, _i64, _P, _i64, _P]),
    "wm_unet_up": (C.c_int, [_P, _i64, _i64, _i64, _i64, _P, _i64, _i64, _P]),
    "wm_unet_chcopy": (C.c_int, [_P, _i64, _i64, _P, _i64, _i64, _i64, _i64, _P]),
    "wm_unet_upcat": (C.c_int, [_P, _i64, _i64, _i64, _i64, _P, _i64, _P, _P]),
    "wm_unet_add": (C.c_int, [_P, _P, _i64, _P]),
    "wm_unet_adam": (C.c_int, [_P, _P, _P, _P, _i64, _d, _d, _d, _d, _i64, _P]),
    "wm_unet_input": (C.c_int, [_P, _P, _P, _i64, _i64, _i64, _d, _d, _P, _P]),
    "wm_unet_apply": (C.c_int, [_P, _P, _i64, _i64, _i64, _d, _P, _P]),
    "wm_unet_mse": (C.c_int, [_P, _P, _i64, _P, _P, _P]),
    "wm_unet_half_mean": (C.c_int, [_P, _i64, _i64, _i64, _i64, _P, _P]),
    "wm_unet_second_diff": (C.c_int, [_P, _P, _P, _i64, _P, _P, _P]),
    "wm_unet_learn_input": (C.c_int, [_P, _P, _P, _P, _i64, _P, _P]),
    "wm_unet_learn_target": (C.c_int, [_P, _P, _P, _i64, _i64, _i64, _P, _d, _P, _P]),
    "wm_unet_learn_apply": (C.c_int, [_P, _P, _d, _i64, _i64, _i64, _P, _P, _P]),
    "wm_particle_sort_worksize": (_i64, [_i64, _i64, _i64, _i64]),
    "wm_particle_sort": (C.c_int, [_i64, _P, _P, _P, _d, _d, _d, _d, _d, _d, _i64, _i64, _i64,
                                   _P, _P, _P]),
    "wm_permute_columns": (C.c_int, [_i64, _i64, _P, _P, _P, _P]),
    "wm_getselfe3d": (C.c_int, [_P, _i64, _i64, _i64, _P, _d, _d, _d, _i64, _P]),
    "wm_getgrid3d": (C.c_int, [_i64, _P, _P, _P, _P, _i64, _i64, _i64, _P, _d, _d, _d, _d, _d,
                               _d, _i64, _i64, _P]),
    "wm_clearpart_worksize": (_i64, [_i64]),
    "wm_particleboundaries3d": (C.c_int, [_i64, _i64, _P, _d, _d, _d, _d, _d, _d, _i64, _i64,
                                          _i64, _d, _i64, _i64, _P, _P, _P, _P]),
    "wm_clearpart": (C.c_int, [_i64, _i64, _P, _i64, _P, _P, _P, _P]),
    "wm_zcylinder_intercepts": (C.c_int, [_d, _d, _d, _d, _d, _P, _d, _i64]),
    "wm_zsrfrv_intercepts": (C.c_int, [_i64, _P, _P, _d, _d, _d, _P]),
    "wm_condlist_create": (C.c_int, [_P]),
    "wm_condlist_destroy": (C.c_int, [_P]),
    "wm_conductordelfromintercepts": (C.c_int, [_P, _P, _d, _d]),
    "wm_checkconductors": (C.c_int, [_P, _i64, _i64, _i64, _d, _d, _d]),
    "wm_condlist_view": (C.c_int, [_P, _P]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """Load libwarpmi.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise NativeError(
                "libwarpmi.so is not built (%s); run `python -m warp_amd.build`" % LIBPATH)
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    """Raise NativeError with the library's message for a nonzero code."""
    if rc != 0:
        msg = lib().wm_last_error().decode(errors="replace")
        raise NativeError("%s failed (%d): %s" % (what, rc, msg))


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc


def ptr(t):
    """Device or host address of a contiguous torch tensor or numpy array."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return t.data_ptr()
    return t.ctypes.data


def stream_of(device):
    """hipStream_t of torch's current stream on device."""
    import torch
    return torch.cuda.current_stream(device).cuda_stream
