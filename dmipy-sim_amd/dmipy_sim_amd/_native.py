"""ctypes binding of the HIP walk library (include/dmipy_walk.h).

The library ``libdmw.so`` is built in-tree next to this file.  There is no
fallback: when the library or a GPU is missing every entry point raises.
"""
import ctypes as C
import os
import weakref

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdmw.so")
_lib = None

DMW_OK, DMW_EINVAL, DMW_EUNSUPPORTED, DMW_ENOMEM, DMW_ENODEV, DMW_EHIP = range(6)
# dmw_phase_form: how a walk accumulated the phase
PHASE_FORMS = ("float32 registers", "float32 Kahan-compensated, registers",
               "fp16 hi/lo split on matrix cores, fp32 accumulate",
               "float32 Kahan-compensated, parked in HBM between samples")

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


class Myelin_t(C.Structure):
    _fields_ = [("step_len", C.c_float * 3), ("two_kappa_over_D", C.c_float * 4),
                ("dt_over_T2", C.c_float * 3), ("spin_weight", C.c_float * 3),
                ("n_start", C.c_int64 * 2)]


class PackedMyelin_t(C.Structure):
    _fields_ = [("inner_radii", _fp), ("step_len", _fp), ("two_kappa_over_D", _fp),
                ("dt_over_T2", C.c_float * 3), ("rho_over_D", C.c_float),
                ("myelin_weight", C.c_float), ("start", C.POINTER(C.c_int64))]


class Geometry_t(C.Structure):
    _fields_ = [("kind", C.c_int32), ("wall", C.c_int32), ("n_obj", C.c_int32),
                ("size", C.c_float), ("cell", C.c_float), ("eps", C.c_float),
                ("nudge", C.c_float), ("rho_over_D", C.c_float), ("kappa_over_D", C.c_float),
                ("centers", _fp), ("radii", _fp), ("axes", C.c_float * 3), ("aux", C.c_float),
                ("myelin", C.POINTER(Myelin_t)), ("packed_myelin", C.POINTER(PackedMyelin_t))]


class Waveform_t(C.Structure):
    _fields_ = [("n_meas", C.c_int32), ("n_t", C.c_int32), ("G", _fp), ("chi", _fp),
                ("n_echo", C.c_int32), ("echo_idx", _ip)]


class Run_t(C.Structure):
    _fields_ = [("n_walkers", C.c_int64), ("seed", C.c_uint64), ("n_sub", C.c_int32),
                ("step_len", C.c_float), ("gamma_dt_sub", C.c_float), ("dt_sub", C.c_float),
                ("inv_T2", C.c_float), ("inv_T1", C.c_float), ("has_t2", C.c_int32),
                ("has_t1", C.c_int32), ("kahan", C.c_int32), ("meas_block", C.c_int32),
                ("r0", _fp), ("inj_normals", _fp), ("inj_uniforms", _fp),
                ("first_walker", C.c_int64)]


class Result_t(C.Structure):
    _fields_ = [("signal", _dp), ("signal_im", _dp), ("final_pos", _fp),
                ("comp_origin", _ip), ("comp_final", _ip), ("comp_full", _ip),
                ("pos_full", _fp), ("n_cross", _up), ("log_w", _fp), ("phi", _fp),
                ("kernel_ms", C.c_double), ("n_hits", C.c_uint64), ("n_cand", C.c_uint64),
                ("want_counters", C.c_int32), ("phase_form", C.c_int32)]


class Instance_t(C.Structure):
    _fields_ = [("geom", Geometry_t), ("n_walkers", C.c_int64), ("n_sub", C.c_int32),
                ("step_len", C.c_float), ("gamma_dt_sub", C.c_float), ("dt_sub", C.c_float),
                ("inv_T2", C.c_float), ("has_t2", C.c_int32), ("wave_set", C.c_int32),
                ("stream", C.c_int32), ("seed_offset", C.c_uint64)]


class Replay_t(C.Structure):
    _fields_ = [("n_meas", C.c_int32), ("n_t", C.c_int32), ("G", _dp), ("scale", C.c_double),
                ("chi", _dp), ("chi_rows", C.c_int32), ("log_w0", _dp),
                ("rho_over_D", C.c_double), ("meas_tiles", C.c_int32)]


class BatchStats_t(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double),
                ("walker_steps", C.c_uint64), ("n_units", C.c_int32), ("grid", C.c_int32)]


class BatchTuning_t(C.Structure):
    _fields_ = [("waves_per_simd", C.c_int32), ("walkers_per_wave", C.c_int32),
                ("unroll", C.c_int32)]


#: Every symbol include/dmipy_walk.h declares, with (restype, argtypes).
SIGNATURES = {
    "dmw_device_count": (C.c_int, []),
    "dmw_set_device": (C.c_int, [C.c_int]),
    "dmw_device_arch": (C.c_int, [C.c_int32, C.c_char_p, C.c_int32]),
    "dmw_last_error": (C.c_char_p, []),
    "dmw_walk": (C.c_int, [C.POINTER(Geometry_t), C.POINTER(Waveform_t), C.POINTER(Run_t),
                           C.POINTER(Result_t)]),
    "dmw_walk_noalloc": (C.c_int, [C.POINTER(Geometry_t), C.POINTER(Waveform_t),
                                   C.POINTER(Run_t), C.POINTER(Result_t)]),
    "dmw_plan_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(Geometry_t),
                                  C.POINTER(Waveform_t), C.POINTER(Run_t)]),
    "dmw_plan_execute": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(Result_t)]),
    "dmw_plan_destroy": (C.c_int, [C.c_void_p]),
    "dmw_plan_advance": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(Result_t),
                                   C.POINTER(C.c_int32)]),
    "dmw_plan_rewind": (C.c_int, [C.c_void_p]),
    "dmw_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(Instance_t),
                                   C.c_int32, C.POINTER(Waveform_t), C.POINTER(BatchTuning_t),
                                   C.c_int32]),
    "dmw_batch_execute": (C.c_int, [C.c_void_p, C.c_uint64, _dp, C.POINTER(BatchStats_t)]),
    "dmw_batch_destroy": (C.c_int, [C.c_void_p]),
    "dmw_traj_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(Geometry_t), C.POINTER(Run_t),
                                  C.c_int32, C.c_int32, _fp]),
    "dmw_traj_from_host": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64, C.c_int32,
                                     C.POINTER(C.c_uint16), C.POINTER(C.c_uint16)]),
    "dmw_traj_download": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16),
                                    C.POINTER(C.c_uint16)]),
    "dmw_traj_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32), _dp]),
    "dmw_traj_destroy": (C.c_int, [C.c_void_p]),
    "dmw_replay_run": (C.c_int, [C.c_void_p, C.POINTER(Replay_t), _dp, _dp, _dp, _dp]),
    "dmw_mlp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32,
                                 C.c_int32]),
    "dmw_mlp_destroy": (C.c_int, [C.c_void_p]),
    "dmw_mlp_num_params": (C.c_int64, [C.c_void_p]),
    "dmw_mlp_set_params": (C.c_int, [C.c_void_p, _fp, C.c_int64]),
    "dmw_mlp_get_params": (C.c_int, [C.c_void_p, _fp, C.c_int64]),
    "dmw_mlp_forward": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    "dmw_mlp_gradients": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, _fp, _fp]),
    "dmw_mlp_set_data": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int64]),
    "dmw_mlp_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, _fp, C.c_float,
                                C.c_float, C.c_float, _fp, _dp]),
    "dmw_mlp_predict": (C.c_int, [C.c_void_p, C.c_int64, _fp, _dp]),
}


class NativeLibraryMissing(RuntimeError):
    """The HIP walk library has not been built (run __graft_entry__.build())."""


def lib():
    """Load ``libdmw.so`` once; raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise NativeLibraryMissing(
                f"HIP walk library not found at {_LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        handle = C.CDLL(_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def device_count():
    return int(lib().dmw_device_count())


# Live library objects that own device memory (plans, batches, networks, trajectory
# sets), so that gpu.free_gpu_memory(aggressive=True) can release them all.
_live = weakref.WeakSet()


def track(obj):
    """Register an object with a close() that frees its device memory."""
    _live.add(obj)
    return obj


def live_objects():
    return list(_live)


def device_arch(device=0):
    """The device's architecture name with its feature flags sanitised for a
    file name ("gfx950_sramecc+_xnack-"), or None without a device."""
    if device_count() <= device:
        return None
    buf = C.create_string_buffer(128)
    check(lib().dmw_device_arch(int(device), buf, len(buf)))
    return buf.value.decode().replace(":", "_") or None


def check(rc):
    """Map a dmw_status to the reference's exception types (ValueError,
    NotImplementedError, RuntimeError with "out of memory")."""
    if rc == DMW_OK:
        return
    msg = lib().dmw_last_error().decode(errors="replace")
    if rc == DMW_EINVAL:
        raise ValueError(msg)
    if rc == DMW_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == DMW_ENOMEM:
        raise RuntimeError(msg if "out of memory" in msg else msg + " (out of memory)")
    raise RuntimeError(msg)


def fptr(a):
    return None if a is None else a.ctypes.data_as(_fp)


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)


def uptr(a):
    return None if a is None else a.ctypes.data_as(_up)


def f32c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)
