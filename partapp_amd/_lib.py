"""ctypes view of libpartapp_ps.so (include/partapp_ps.h).

The library is built in-tree (partapp_amd/build.py) and loaded from
partapp_amd/lib/.  There is no fallback: a missing library is an error.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpartapp_ps.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)

PS_OK = 0
STATUS = {
    0: "PS_OK",
    -1: "PS_ERR_INVALID_ARG",
    -2: "PS_ERR_TREE",
    -3: "PS_ERR_COVARIANCE",
    -4: "PS_ERR_KERNEL_SIZE",
    -5: "PS_ERR_UNSUPPORTED",
    -6: "PS_ERR_WARP",
    -7: "PS_ERR_HIP",
    -8: "PS_ERR_NO_DEVICE",
    -9: "PS_ERR_OUT_OF_MEMORY",
}


class ps_exp_param(ctypes.Structure):
    _fields_ = [
        ("min_part_rotation", ctypes.c_float),
        ("max_part_rotation", ctypes.c_float),
        ("num_rotation_steps", ctypes.c_uint),
        ("min_object_scale", ctypes.c_float),
        ("max_object_scale", ctypes.c_float),
        ("num_scale_steps", ctypes.c_uint),
        ("strip_border_detections", ctypes.c_float),
    ]


class ps_joint(ctypes.Structure):
    _fields_ = [
        ("child_idx", ctypes.c_int),
        ("parent_idx", ctypes.c_int),
        ("offset_c", ctypes.c_double * 2),
        ("offset_p", ctypes.c_double * 2),
        ("C", ctypes.c_double * 4),
        ("rot_mean", ctypes.c_double),
        ("rot_sigma", ctypes.c_double),
    ]


class ps_model_desc(ctypes.Structure):
    _fields_ = [
        ("num_parts", ctypes.c_int),
        ("is_detect", c_int_p),
        ("is_upright", c_int_p),
        ("is_root", c_int_p),
        ("num_joints", ctypes.c_int),
        ("joints", ctypes.POINTER(ps_joint)),
        ("exp_param", ps_exp_param),
        ("is_sparse", ctypes.c_int),
    ]


class ps_instance(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int),
        ("width", ctypes.c_int),
        ("unary_offset", ctypes.c_longlong),
        ("marginal_offset", ctypes.c_longlong),
        ("root_offset", ctypes.c_longlong),
        ("joints", ctypes.POINTER(ps_joint)),
    ]


class ps_map(ctypes.Structure):
    _fields_ = [
        ("rotidx", ctypes.c_int),
        ("y", ctypes.c_int),
        ("x", ctypes.c_int),
        ("score", ctypes.c_float),
    ]


class ps_local_max(ctypes.Structure):
    _fields_ = [
        ("d", ctypes.c_int),
        ("x", ctypes.c_int),
        ("y", ctypes.c_int),
        ("score", ctypes.c_float),
    ]


# name -> (restype, argtypes); the exported surface of include/partapp_ps.h
SIGNATURES = {
    "ps_model_create": (ctypes.c_int, [ctypes.POINTER(ps_model_desc), ctypes.POINTER(ctypes.c_void_p)]),
    "ps_model_destroy": (None, [ctypes.c_void_p]),
    "ps_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ps_instance), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_void_p)]),
    "ps_plan_destroy": (None, [ctypes.c_void_p]),
    "ps_plan_scratch_bytes": (ctypes.c_size_t, [ctypes.c_void_p]),
    "ps_plan_execute": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "ps_plan_execute_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    "ps_plan_set_method": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_plan_method_levels": (ctypes.c_int, [ctypes.c_void_p, c_int_p, c_int_p]),
    "ps_plan_keep_layout": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                                           ctypes.POINTER(ctypes.c_longlong)]),
    "ps_plan_execute_pruned": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "ps_plan_set_map_guard": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    "ps_plan_read_map_guard": (ctypes.c_int, [ctypes.c_void_p, c_int_p, ctypes.POINTER(ctypes.c_uint),
                                              ctypes.c_void_p]),
    "ps_plan_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "ps_plan_tune_class": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ps_plan_classes": (ctypes.c_int, [ctypes.c_void_p, c_int_p, c_int_p, ctypes.c_int]),
    "ps_plan_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_plan_set_profiling_classes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint]),
    "ps_plan_read_profile": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_int_p]),
    "ps_marginals": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ps_instance), ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ps_map), ctypes.c_void_p]),
    "ps_message": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                  c_double_p, c_double_p, c_double_p, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_void_p]),
    "ps_message_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_int, c_double_p, c_double_p, c_double_p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]),
    "ps_find_local_max": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), c_int_p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "ps_ingest_scores": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_double_p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ps_add_torso_pos_prior": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_double_p,
                                              ctypes.c_float, ctypes.c_void_p]),
    "ps_add_dpm_score": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_void_p]),
    "ps_add_extra_unary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_float,
                                          ctypes.c_void_p]),
    "ps_mat_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "ps_mat_close": (None, [ctypes.c_void_p]),
    "ps_mat_var": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_longlong, c_int_p, c_int_p,
                                  ctypes.POINTER(ctypes.c_longlong), c_double_p, ctypes.c_longlong]),
    "ps_mat_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, c_double_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_longlong), ctypes.c_int, ctypes.c_int]),
    "ps_load_joint_mat": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ps_joint), c_int_p]),
    "ps_synth_unary_dev": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ps_last_error_string": (ctypes.c_char_p, []),
    "ps_version": (ctypes.c_char_p, []),
    "ps_debug_message_plan": (ctypes.c_int, [ctypes.POINTER(ps_exp_param), ctypes.c_int, ctypes.c_int, c_double_p,
                                             c_double_p, c_double_p, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                             c_float_p, c_int_p, c_int_p]),
    "ps_synth_joints": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ps_joint),
                                       c_int_p]),
    "ps_synth_unary": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, c_float_p, ctypes.c_int]),
}

_lib = None


class PsError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = lib().ps_last_error_string().decode(errors="replace")
        super().__init__("%s failed: %s (%s)" % (where, STATUS.get(status, status), msg))


def lib():
    """Load libpartapp_ps.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("%s is missing: build it with `python -m partapp_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != PS_OK:
        raise PsError(status, where)
