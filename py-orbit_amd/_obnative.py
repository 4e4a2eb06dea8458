"""Loader for liborbit_mi355x.so, the C ABI declared in include/orbit_mi355x.h.

The library is the only compute path: if it is missing or no device is visible, every
call raises.  There is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborbit_mi355x.so")

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_vp = C.c_void_p
_i, _d, _i64 = C.c_int, C.c_double, C.c_int64

# name -> argtypes (all functions return int status unless listed in _RET)
SIGNATURES = {
    "ob_init": [_i],
    "ob_synchronize": [],
    "ob_bunch_create": [C.POINTER(_vp)],
    "ob_bunch_destroy": [_vp],
    "ob_bunch_set_mass": [_vp, _d],
    "ob_bunch_set_charge": [_vp, _d],
    "ob_bunch_set_classical_radius": [_vp, _d],
    "ob_bunch_set_macrosize": [_vp, _d],
    "ob_bunch_get_attrs": [_vp, _D],
    "ob_bunch_set_kin_energy": [_vp, _d],
    "ob_bunch_set_sync_momentum": [_vp, _d],
    "ob_bunch_get_sync": [_vp, _D],
    "ob_bunch_set_sync_time": [_vp, _d],
    "ob_bunch_add_particle": [_vp, _d, _d, _d, _d, _d, _d],
    "ob_bunch_set_coords": [_vp, _i64, _D, _D, _D, _D, _D, _D],
    "ob_bunch_get_coords": [_vp, _D, _D, _D, _D, _D, _D],
    "ob_bunch_get_coord": [_vp, _i64, _i, _D],
    "ob_bunch_set_coord": [_vp, _i64, _i, _d],
    "ob_bunch_get_size": [_vp, C.POINTER(_i64)],
    "ob_bunch_delete_particle": [_vp, _i64],
    "ob_bunch_compress": [_vp],
    "ob_bunch_set_macrosizes": [_vp, _i64, _D],
    "ob_bunch_get_macrosizes": [_vp, _i64, _D],
    "ob_bunch_moments": [_vp, _D],
    "ob_bunch_moments_noalloc": [_vp, _D],
    "ob_bunch_extrema": [_vp, _D],
    "ob_bunch_extrema_r": [_vp, _D],
    "ob_grid1d_bin_bunch": [_vp, _i, _i, _i, _d, _d, _D],
    "ob_bunch_xy_moments": [_vp, _i, _D, _D],
    "ob_bunch_device_coords": [_vp, C.POINTER(_vp)],
    "ob_bunch_flush": [_vp],
    "ob_tpb_quad3": [_vp, _d, _d, _i],
    "ob_tpb_drifti": [_vp, _i64, _d],
    "ob_tpb_wedgerotate": [_vp, _d, _i],
    "ob_tpb_wedgedrift": [_vp, _d, _i],
    "ob_tpb_wedgebend": [_vp, _d, _i, _d, _i],
    "ob_tpb_bend1": [_vp, _d, _d],
    "ob_tpb_bend2": [_vp, _d],
    "ob_tpb_bend3": [_vp, _d],
    "ob_tpb_bend4": [_vp, _d],
    "ob_tpb_bendfringeIN": [_vp, _d],
    "ob_tpb_bendfringeOUT": [_vp, _d],
    "ob_tpb_soln": [_vp, _d, _d, _i],
    "ob_tpb_wedgebendCF": [_vp, _d, _i, _d, _i, C.POINTER(_i), _D, C.POINTER(_i), _i, _i],
    "ob_tpb_RingRF": [_vp, _d, _i, _d, _d, _i],
    "ob_track_sequence": [_vp, _vp, _i],
    "ob_tpb_rotatexy": [_vp, _d],
    "ob_tpb_drift": [_vp, _d],
    "ob_tpb_wrapbunch": [_vp, _d],
    "ob_tpb_kick": [_vp, _d, _d, _d, _i],
    "ob_tpb_multp": [_vp, _i, _d, _i, _i],
    "ob_tpb_multpfringeIN": [_vp, _i, _d, _i, _i],
    "ob_tpb_multpfringeOUT": [_vp, _i, _d, _i, _i],
    "ob_tpb_quad1": [_vp, _d, _d, _i],
    "ob_tpb_quad2": [_vp, _d],
    "ob_tpb_quadfringeIN": [_vp, _d, _i],
    "ob_tpb_quadfringeOUT": [_vp, _d, _i],
    "ob_sc2p5d_create": [_i, _i, _i, _d, C.POINTER(_vp)],
    "ob_sc2p5d_destroy": [_vp],
    "ob_sc2p5d_track": [_vp, _vp, _d],
    "ob_sc2p5d_get_grid": [_vp, _i, _D, _D],
    "ob_sc2p5d_get_state": [_vp, _D],
    "ob_sc2p5d_track_boundary": [_vp, _vp, _d, _vp],
    "ob_slice2d_create": [_i, _i, _i, _d, C.POINTER(_vp)],
    "ob_slice2d_destroy": [_vp],
    "ob_slice2d_track": [_vp, _vp, _d, _vp],
    "ob_slice2d_get_grid": [_vp, _i, _D, _D],
    "ob_slice2d_get_state": [_vp, _D],
    "ob_boundary2d_create": [_i, _i, C.POINTER(_vp)],
    "ob_boundary2d_create_shaped": [_i, _i, C.c_char_p, _d, _d, C.POINTER(_vp)],
    "ob_boundary2d_destroy": [_vp],
    "ob_boundary2d_set_point": [_vp, _i, _d, _d],
    "ob_boundary2d_get_point": [_vp, _i, _D, _D],
    "ob_boundary2d_initialize": [_vp],
    "ob_boundary2d_info": [_vp, _D, C.POINTER(_i)],
    "ob_boundary2d_is_inside": [_vp, _d, _d, C.POINTER(_i)],
    "ob_boundary2d_add_potential": [_vp, _D, _i, _i, _D],
    "ob_sc2p5drb_create": [_i, _i, _i, _d, C.POINTER(_vp)],
    "ob_sc2p5drb_destroy": [_vp],
    "ob_sc2p5drb_track": [_vp, _vp, _d, _d],
    "ob_sc2p5drb_get_grid": [_vp, _i, _D, _D],
    "ob_sc2p5drb_get_moments": [_vp, _D],
    "ob_sc2p5drb_set_long_avg_n": [_vp, _i],
    "ob_sc2p5drb_get_long_avg_n": [_vp, C.POINTER(_i)],
    "ob_sc3d_create": [_i, _i, _i, C.POINTER(_vp)],
    "ob_sc3d_destroy": [_vp],
    "ob_sc3d_track": [_vp, _vp, _d],
    "ob_sc3d_get_grid": [_vp, _i, _D, _D],
    "ob_sc3d_set_ratio_limit": [_vp, _d],
    "ob_sc3d_get_ratio_limit": [_vp, _D],
    "ob_sc3d_green_rebuilds": [_vp, C.POINTER(_i)],
    "ob_sc3d_set_deposit_mode": [_vp, _i, _i],
    "ob_sc3d_set_num_ext_bunches": [_vp, _i],
    "ob_sc3d_get_num_ext_bunches": [_vp, C.POINTER(_i)],
    "ob_sc3d_set_freq": [_vp, _d],
    "ob_sc3d_get_freq": [_vp, _D],
    "ob_batch_create": [C.POINTER(_vp), _i, C.POINTER(_vp)],
    "ob_batch_destroy": [_vp],
    "ob_batch_set_sc2p5drb": [_vp, C.POINTER(_vp)],
    "ob_batch_track_sequence": [_vp, _vp, _i],
    "ob_batch_sc2p5drb_track": [_vp, _d, _d],
    "ob_batch_flush": [_vp],
    "ob_batch_synchronize": [_vp],
    "ob_batch_set_tuning": [_vp, _i, _i],
    "ob_batch_get_tuning": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)],
    "ob_batch_capture_begin": [_vp],
    "ob_batch_capture_end": [_vp],
    "ob_batch_replay": [_vp, _i],
    "ob_sc2p5drb_find_potential": [_vp, _D, _d, _d, _d, _d, _D],
    "ob_sc2p5drb_set_denoiser": [_vp, _vp],
    "ob_dn_create": [_i, _i, _i, C.POINTER(_vp)],
    "ob_dn_destroy": [_vp],
    "ob_dn_num_params": [_vp, C.POINTER(_i64)],
    "ob_dn_set_params": [_vp, _F, _i64],
    "ob_dn_get_params": [_vp, _F, _i64],
    "ob_dn_forward": [_vp, _F, _F, _i],
    "ob_dn_forward_dev": [_vp, _vp, _vp, _i],
    "ob_dn_loss_grad": [_vp, _F, _F, _i, _F, _F],
    "ob_dn_train_step_dev": [_vp, _vp, _vp, _i, C.c_float],
    "ob_dn_last_loss": [_vp, _F],
    "ob_dn_set_loss": [_vp, _i, _F],
    "ob_dn_make_pairs": [_i, _i, _i, _F, C.c_uint64, _i64, _i64, _vp, _vp],
    "ob_dn_pair_from_bunch": [_i, _i, _vp, _i64, _vp, _vp],
    "ob_dn_gather": [_vp, _i64, _vp, _i64, C.POINTER(_i), _i],
    "ob_dn_alloc": [_i64, C.POINTER(_vp)],
    "ob_dn_free": [_vp],
    "ob_dn_copy": [_vp, _vp, _i64, _i],
    "ob_profile_enable": [_i],
    "ob_profile_report": [C.c_char_p, _i],
    "ob_calibrate_copy": [_i64, _i, _D],
    "ob_calibrate_copy_wide": [_i64, _i, _D, C.POINTER(C.c_int)],
}
_RET = {"ob_last_error": C.c_char_p, "ob_version": C.c_char_p}

_lib = None


class OrbitError(RuntimeError):
    """A failure reported by liborbit_mi355x (the reference would abort the process)."""


def load():
    """Load the native library (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OrbitError(f"{LIB_PATH} is not built: run python py-orbit_amd/build_lib.py")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    for name, res in _RET.items():
        f = getattr(L, name)
        f.argtypes = []
        f.restype = res
    _lib = L
    return L


def call(name, *args):
    """Call an entry point; raise OrbitError with the library's message on failure."""
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.ob_last_error().decode(errors="replace")
        raise OrbitError(f"{name}: {msg}")
    return rc


def dptr(a):
    return a.ctypes.data_as(_D)


def fptr(a):
    return a.ctypes.data_as(_F)


def synchronize():
    call("ob_synchronize")
