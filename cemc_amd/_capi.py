"""ctypes binding of libcemcgpu.so (the C ABI in include/cemc_gpu.h).

The library is loaded from cemc_amd/lib/.  A missing library raises
ImportError: there is no CPU fallback for any operation of this package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libcemcgpu.so")

CEMC_OK, CEMC_EINVAL, CEMC_ERUNTIME, CEMC_EDEVICE, CEMC_EFULL = 0, 1, 2, 3, 4


class PoolFull(RuntimeError):
    """A pool has no free slot or device memory for one more chain."""
CF_C0, CF_SINGLET, CF_CLUSTER = 0, 1, 2
MOVE_CANONICAL, MOVE_SGC = 0, 1
PATH_AUTO, PATH_EVALUATE, PATH_FIELDS = 0, 1, 2

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i8p = ctypes.POINTER(ctypes.c_int8)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("n_sites", ctypes.c_int32), ("n_species", ctypes.c_int32), ("n_groups", ctypes.c_int32),
        ("n_cf", ctypes.c_int32), ("n_entries", ctypes.c_int32), ("n_trans_cols", ctypes.c_int32),
        ("basis", _dp), ("site_group", _i32p), ("cf_kind", _i32p), ("cf_dec", _i32p),
        ("cf_entry", _i32p), ("cf_eq_off", _i32p), ("eq_decos", _i32p), ("entry_group", _i32p),
        ("entry_size", _i32p), ("entry_M", _i32p), ("entry_norm", _i32p), ("entry_mem_off", _i32p),
        ("members", _i32p), ("entry_ord_off", _i32p), ("order", _i32p), ("trans", _i32p),
        ("periodic_dims", _i32p), ("col_cells", _i32p), ("eci", _dp), ("background", _i32p),
    ]


class ChainDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("temperature", ctypes.c_double), ("chem_pot", _dp),
        ("seed", ctypes.c_uint64), ("uid", ctypes.c_uint64), ("occupation", _i8p),
        ("first_move", ctypes.c_uint64),
    ]


class Thermo(ctypes.Structure):
    _fields_ = [
        ("n_samples", ctypes.c_int64), ("energy", ctypes.c_double), ("mean_energy", ctypes.c_double),
        ("var_energy", ctypes.c_double), ("n_singlets", ctypes.c_int32),
        ("singlet", ctypes.c_double * 7), ("var_singlet", ctypes.c_double * 7),
        ("cov_singlet_energy", ctypes.c_double * 7),
    ]


class HooksDesc(ctypes.Structure):
    _fields_ = [
        ("fixed_species", ctypes.c_uint32), ("max_attempts", ctypes.c_int32),
        ("allowed", ctypes.POINTER(ctypes.c_uint8)), ("pair_species", ctypes.c_int32 * 2),
        ("pair_members", ctypes.c_int32), ("pair_neighbours", ctypes.POINTER(ctypes.c_int32)),
        ("bias_species", ctypes.c_int32), ("recycle_waste", ctypes.c_int32),
        ("bias_table", ctypes.POINTER(ctypes.c_double)),
    ]


class WLDesc(ctypes.Structure):
    _fields_ = [
        ("n_bins", ctypes.c_int32), ("current_bin", ctypes.c_int32), ("is_first", ctypes.c_int32),
        ("converged", ctypes.c_int32), ("e_min", ctypes.c_double), ("e_max", ctypes.c_double),
        ("f", ctypes.c_double), ("f_min", ctypes.c_double), ("flatness", ctypes.c_double),
    ]


class WLScheme(ctypes.Structure):
    _fields_ = [
        ("inverse_time", ctypes.c_int32), ("adaptive", ctypes.c_int32), ("min_width", ctypes.c_int32),
        ("inv_active", ctypes.c_int32), ("upper", ctypes.c_int32), ("n_overlap", ctypes.c_int32),
        ("n_edges", ctypes.c_int32), ("has_first", ctypes.c_int32), ("inv_factor", ctypes.c_double),
        ("bin_change", ctypes.c_double), ("steps", ctypes.c_int64), ("since_flat", ctypes.c_int64),
        ("prev_conv", ctypes.c_int64), ("n_walk", ctypes.c_int64),
    ]


# (name, restype, argtypes) of every entry point declared in include/cemc_gpu.h
SIGNATURES = [
    ("cemc_last_error", ctypes.c_char_p, []),
    ("cemc_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("cemc_model_create", ctypes.c_int, [ctypes.POINTER(ModelDesc), ctypes.POINTER(_vp)]),
    ("cemc_model_set_ecis", ctypes.c_int, [_vp, _dp]),
    ("cemc_model_is_periodic", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    ("cemc_model_destroy", ctypes.c_int, [_vp]),
    ("cemc_updater_create", ctypes.c_int, [_vp, _i8p, _dp, _dp, ctypes.POINTER(_vp)]),
    ("cemc_updater_destroy", ctypes.c_int, [_vp]),
    ("cemc_updater_update_cf", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("cemc_updater_calculate", ctypes.c_int, [_vp, ctypes.c_int32, _i32p, _i32p, _i32p, _dp]),
    ("cemc_updater_undo_changes", ctypes.c_int, [_vp]),
    ("cemc_updater_clear_history", ctypes.c_int, [_vp]),
    ("cemc_updater_get_energy", ctypes.c_int, [_vp, _dp]),
    ("cemc_updater_get_cf", ctypes.c_int, [_vp, _dp]),
    ("cemc_updater_set_ecis", ctypes.c_int, [_vp, _dp]),
    ("cemc_updater_get_singlets", ctypes.c_int, [_vp, _dp, ctypes.c_int32, _i32p]),
    ("cemc_updater_add_linear_vib_correction", ctypes.c_int, [_vp, ctypes.c_int32, _i32p, _dp]),
    ("cemc_updater_vib_energy", ctypes.c_int, [_vp, ctypes.c_double, _dp]),
    ("cemc_updater_set_num_threads", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("cemc_updater_set_resident", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("cemc_updater_get_occupation", ctypes.c_int, [_vp, _i8p]),
    ("cemc_updater_history_size", ctypes.c_int, [_vp, _i32p]),
    ("cemc_updater_set_occupation", ctypes.c_int, [_vp, _i8p]),
    ("cemc_batch_create", ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(_vp), ctypes.POINTER(ChainDesc),
                                         ctypes.POINTER(_vp)]),
    ("cemc_batch_destroy", ctypes.c_int, [_vp]),
    ("cemc_batch_run", ctypes.c_int, [_vp, _i64p, ctypes.c_int64]),
    ("cemc_batch_run_timed", ctypes.c_int, [_vp, _i64p, ctypes.c_int64, _dp]),
    ("cemc_batch_run_sliced", ctypes.c_int, [_vp, _i64p, ctypes.c_int64, ctypes.c_double, _dp]),
    ("cemc_batch_run_logged", ctypes.c_int, [_vp, _i64p, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint32),
                                             ctypes.c_int64]),
    ("cemc_batch_get_energy", ctypes.c_int, [_vp, _dp]),
    ("cemc_batch_get_occupation", ctypes.c_int, [_vp, ctypes.c_int32, _i8p]),
    ("cemc_batch_get_stats", ctypes.c_int, [_vp, _i64p, _i64p, _dp, _dp]),
    ("cemc_batch_reset_stats", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("cemc_batch_get_thermo", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(Thermo)]),
    ("cemc_batch_get_cf", ctypes.c_int, [_vp, ctypes.c_int32, _dp, _dp]),
    ("cemc_batch_get_cf_noalloc", ctypes.c_int, [_vp, ctypes.c_int32, _dp, _dp]),
    ("cemc_batch_set_waves", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("cemc_batch_set_moves_per_round", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("cemc_batch_set_energy_path", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("cemc_batch_get_energy_path", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    ("cemc_batch_get_launch_shape", ctypes.c_int, [_vp, ctypes.c_int32, _i32p, _i32p, _i32p]),
    ("cemc_batch_set_temperature", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double]),
    ("cemc_batch_set_hooks", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(HooksDesc)]),
    ("cemc_batch_get_hooks_state", ctypes.c_int, [_vp, ctypes.c_int32, _i64p, _i32p, _dp]),
    ("cemc_batch_wl_set", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(WLDesc), _dp, _i32p,
                                         ctypes.POINTER(ctypes.c_uint8)]),
    ("cemc_batch_wl_get", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(WLDesc), _dp, _i32p,
                                         ctypes.POINTER(ctypes.c_uint8), _i64p]),
    ("cemc_batch_wl_set_scheme", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(WLScheme)]),
    ("cemc_batch_wl_get_scheme", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(WLScheme), _i32p, _i32p, _dp,
                                                ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint8)]),
    ("cemc_batch_wl_run_shared", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.POINTER(ctypes.c_uint32), ctypes.c_int64, _dp]),
    ("cemc_batch_wl_run", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_uint32), ctypes.c_int64, _dp]),
    ("cemc_pool_create", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.POINTER(_vp)]),
    ("cemc_pool_footprint", ctypes.c_int, [_vp, ctypes.c_int32, _i64p, _i64p]),
    ("cemc_pool_add", ctypes.c_int, [_vp, _vp, ctypes.POINTER(ChainDesc), _i32p]),
    ("cemc_pool_remove", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("cemc_pt_create", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("cemc_pt_destroy", ctypes.c_int, [_vp]),
    ("cemc_pt_run", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _dp]),
    ("cemc_pt_get", ctypes.c_int, [_vp, _i32p, _dp, _i64p, _i64p, _i64p, _dp, _dp]),
]

_lib = None


def library_path():
    return _LIB_PATH


def load():
    """Load libcemcgpu.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError("libcemcgpu.so not found at %s: build it with `python -m cemc_amd.build`" % _LIB_PATH)
    lib = ctypes.CDLL(_LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    """Raise the Python exception matching a C-ABI status code."""
    if rc == CEMC_OK:
        return
    msg = _lib.cemc_last_error().decode(errors="replace") if _lib is not None else "libcemcgpu error"
    if rc == CEMC_EINVAL:
        raise ValueError(msg)
    if rc == CEMC_EFULL:
        raise PoolFull(msg)
    raise RuntimeError(msg)


def device_count():
    lib = load()
    n = ctypes.c_int(0)
    check(lib.cemc_device_count(ctypes.byref(n)))
    return n.value


def require_gpu():
    """Raise RuntimeError when no GPU is visible (nothing here runs on the CPU)."""
    if device_count() < 1:
        raise RuntimeError("no GPU visible: the cluster-expansion engine runs on the GPU only")


def i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def i8(a):
    a = np.ascontiguousarray(a, dtype=np.int8)
    return a, a.ctypes.data_as(_i8p)


def i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_i64p)
