"""ctypes binding of libkifmm2d.so (the C-ABI declared in include/kifmm2d.h).

The library is built in-tree by ``python -m pykifmm2d_amd.build``.  Importing the compute
entry points without it raises: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkifmm2d.so")

KF_OK = 0
KF_LAPLACE = 0
KF_MODHELM = 1
KF_TABLE = 2
KF_PREC_FP64 = 0
KF_PREC_SPLIT = 1
KF_PREC_SPLIT16 = 2

c_i32, c_i64, c_f64, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
P = ctypes.POINTER


class KernelDesc(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("k", c_f64), ("scale", c_f64), ("table", c_vp), ("table_im", c_vp)]


class TreeInfo(ctypes.Structure):
    _fields_ = [("levels", c_i32), ("key_depth", c_i32), ("n_points", c_i64), ("n_boxes", c_i64),
                ("root_min", c_f64), ("root_width", c_f64)]


class PlanStats(ctypes.Structure):
    _fields_ = [("p2p_pairs", c_i64), ("p2m_evals", c_i64), ("l2p_evals", c_i64), ("m2l_pairs", c_i64),
                ("m2m_boxes", c_i64), ("l2l_boxes", c_i64), ("c2e_up_rows", c_i64), ("c2e_down_rows", c_i64),
                ("leaves", c_i64),
                ("p", c_i32), ("p_padded", c_i32)]


class Tuning(ctypes.Structure):
    _fields_ = [("near_grid", c_i32), ("near_fork", c_i32), ("m2l_pipeline", c_i32), ("level_rows", c_i32),
                ("near_split", c_i32), ("chain", c_i32), ("batch_form", c_i32),
                ("near_form", c_i32)]


# name -> (restype, argtypes); every symbol include/kifmm2d.h declares
SIGNATURES = {
    "kf_version": (c_i32, []),
    "kf_last_error": (ctypes.c_char_p, []),
    "kf_k0": (c_f64, [c_f64]),
    "kf_tree_build": (c_i32, [c_vp, c_vp, c_i64, c_i32, c_vp, P(c_vp)]),
    "kf_tree_get_info": (c_i32, [c_vp, P(TreeInfo)]),
    "kf_tree_level_size": (c_i32, [c_vp, c_i32, P(c_i64)]),
    "kf_tree_export_level": (c_i32, [c_vp, c_i32] + [c_vp] * 10),
    "kf_tree_export_points": (c_i32, [c_vp, c_vp, c_vp, c_vp]),
    "kf_tree_device_points": (c_i32, [c_vp, P(c_vp), P(c_vp), P(c_vp)]),
    "kf_tree_from_sorted_keys": (c_i32, [c_vp, c_i64, c_i32, c_f64, c_f64, c_i32, P(c_vp)]),
    "kf_tree_destroy": (None, [c_vp]),
    "kf_tree_locate": (c_i32, [c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp]),
    "kf_tree_export_depths": (c_i32, [c_vp, c_i32, c_vp]),
    "kf_tree_level_real": (c_i32, [c_vp, c_i32, c_vp]),
    "kf_plan_create": (c_i32, [c_vp, P(KernelDesc), c_i32, c_i32, c_vp, P(c_vp)]),
    "kf_plan_stats_get": (c_i32, [c_vp, P(PlanStats)]),
    "kf_plan_verify": (c_i32, [c_vp, c_vp]),
    "kf_plan_expand": (c_i32, [c_vp, c_vp, c_i64, c_i32, c_vp]),
    "kf_plan_self": (c_i32, [c_vp, c_vp, c_i64, c_vp]),
    "kf_plan_eval": (c_i32, [c_vp, c_vp, c_i64, c_i32, c_vp, c_i64, c_vp]),
    "kf_plan_eval_ptrs": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_vp]),
    "kf_tree_build_batch": (c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp]),
    "kf_forest_create": (c_i32, [c_vp, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp]),
    "kf_plan_reserve": (c_i32, [c_vp, c_i32, c_vp]),
    "kf_plan_set_graphs": (c_i32, [c_vp, c_i32]),
    "kf_plan_get_tuning": (c_i32, [c_vp, c_vp]),
    "kf_plan_set_tuning": (c_i32, [c_vp, c_vp]),
    "kf_plan_autotune": (c_i32, [c_vp, c_i32, c_vp, c_vp, c_vp]),
    "kf_plan_footprint": (c_i32, [c_vp, c_i32, c_vp]),
    "kf_workspace_create": (c_i32, [P(c_vp)]),
    "kf_workspace_destroy": (c_i32, [c_vp]),
    "kf_workspace_bytes": (c_i32, [c_vp, P(c_i64)]),
    "kf_workspace_fill": (c_i32, [c_vp, c_f64, c_vp]),
    "kf_plan_attach_workspace": (c_i32, [c_vp, c_vp]),
    "kf_plan_slab_bytes": (c_i32, [c_vp, c_i32, P(c_i64)]),
    "kf_plan_targets": (c_i32, [c_vp, c_vp, c_vp, c_i64, c_i32, c_vp, c_vp]),
    "kf_plan_profile": (c_i32, [c_vp, c_i32, c_vp, c_i32]),
    "kf_plan_destroy": (None, [c_vp]),
    "kf_direct": (c_i32, [P(KernelDesc), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_i32, c_vp, c_vp]),
    "kf_direct_noalloc": (c_i32, [P(KernelDesc), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_i32, c_vp,
                                  c_vp]),
    "kf_kernel_form": (c_i32, [P(KernelDesc), c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "kf_plan_store_matrices": (c_i32, [c_vp, c_i32, c_vp]),
    "kf_ktable_create": (c_i32, [c_vp, c_i32, c_i32, c_f64, c_f64, c_f64, P(c_vp)]),
    "kf_ktable_destroy": (None, [c_vp]),
    "kf_ktable_eval": (c_i32, [c_vp, c_vp, c_i64, c_vp]),
}

_lib = None


class KifmmError(RuntimeError):
    """A libkifmm2d call returned an error code."""

    def __init__(self, code, msg):
        super().__init__("libkifmm2d error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libkifmm2d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libkifmm2d.so is not built; run `python -m pykifmm2d_amd.build` "
                              "(there is no CPU fallback)")
        # torch first: it bundles its own libamdhip64.so.7, and a process holds one library per
        # soname, so loading this library first would bind torch to the system HIP runtime
        # instead; with torch first, the device memory torch hands over and this library's
        # kernels live in torch's runtime
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != KF_OK:
        raise KifmmError(rc, lib().kf_last_error().decode(errors="replace"))
    return rc
