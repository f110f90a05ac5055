"""ctypes binding of libidwarp_hip.so (include/idwarp_hip.h).

The library is built in-tree by ``idwarp_amd.build.build_library()``; there is
no CPU fallback: if the shared object is missing or the device is not a
gfx950, calls raise.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libidwarp_hip.so")

IDW_OK, IDW_EINVAL, IDW_ESTATE, IDW_EHIP, IDW_ENOMEM, IDW_ENODEV = 0, -1, -2, -3, -4, -5
EVAL_EXACT, EVAL_FAST = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

# every symbol declared in include/idwarp_hip.h
EXPORTS = (
    "idw_default_options", "idw_last_error", "idw_create", "idw_destroy", "idw_set_options",
    "idw_set_volume_nodes", "idw_point_reduce", "idw_initialize_warping", "idw_set_surface_coordinates",
    "idw_get_surface_coordinates", "idw_warp_mesh", "idw_get_volume_coordinates",
    "idw_set_external_mesh_indices", "idw_get_solver_dof", "idw_warp_to_solver_grid", "idw_solver_to_warp_grid",
    "idw_warp_deriv", "idw_get_dxs", "idw_warp_mesh_dev", "idw_warp_deriv_dev", "idw_get_sizes", "idw_get_ldef0",
    "idw_get_denominator0", "idw_get_timings", "idw_fast_counts", "idw_warp_deriv_fwd", "idw_warp_deriv_fwd_dev",
    "idw_set_symmetry_planes", "idw_batch_create", "idw_batch_destroy", "idw_batch_add_warp",
    "idw_batch_add_warp_deriv", "idw_batch_run", "idw_batch_info", "idw_set_restart_file",
    "idw_get_restart_info", "idw_mixed_near_wall_floor", "idw_batch_stats", "idw_device_bytes",
    "idw_warp_deriv_dev_noalloc", "idw_warp_mesh_async", "idw_warp_deriv_async", "idw_wait",
    "idw_tune", "idw_get_tuning", "idw_set_tuning",
)


class IdwOptions(ctypes.Structure):
    _fields_ = [
        ("aExp", ctypes.c_double),
        ("bExp", ctypes.c_double),
        ("LdefFact", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("errTol", ctypes.c_double),
        ("cornerAngle", ctypes.c_double),
        ("evalMode", ctypes.c_int32),
        ("useRotations", ctypes.c_int32),
        ("zeroCornerRotations", ctypes.c_int32),
        ("bucketSize", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("reusePlan", ctypes.c_int32),
        ("nearWallFactor", ctypes.c_double),
        ("deterministic", ctypes.c_int32),
        ("storeWeights", ctypes.c_int32),
    ]


class IdwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("idwarp_hip error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load():
    """Load the library and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("%s is missing: build it with idwarp_amd.build.build_library()" % LIB_PATH)
    # torch ships its own HIP runtime (torch/lib/libamdhip64.so, soname
    # libamdhip64.so.7) and links it by the file name libamdhip64.so.  If this
    # library were loaded first, /opt/rocm's runtime would be in the process
    # and torch would load a second one, which then finds no device.  With
    # torch loaded first this library binds to torch's runtime by soname and
    # the process has a single HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "idw_default_options": (None, [P(IdwOptions)]),
        "idw_last_error": (ctypes.c_char_p, []),
        "idw_create": (i32, [P(vp), i32, P(IdwOptions)]),
        "idw_destroy": (i32, [vp]),
        "idw_set_options": (i32, [vp, P(IdwOptions)]),
        "idw_set_volume_nodes": (i32, [vp, vp, i64]),
        "idw_point_reduce": (i32, [vp, i64, d, vp, vp, P(i64)]),
        "idw_initialize_warping": (i32, [vp, vp, i64, vp, i64, vp, vp, i64, vp, i64]),
        "idw_set_surface_coordinates": (i32, [vp, vp, i64]),
        "idw_get_surface_coordinates": (i32, [vp, vp, i64]),
        "idw_warp_mesh": (i32, [vp]),
        "idw_get_volume_coordinates": (i32, [vp, vp, i64]),
        "idw_set_external_mesh_indices": (i32, [vp, vp, i64]),
        "idw_get_solver_dof": (i32, [vp, P(i64)]),
        "idw_warp_to_solver_grid": (i32, [vp, vp, i64, vp, i64]),
        "idw_solver_to_warp_grid": (i32, [vp, vp, i64, vp, i64]),
        "idw_warp_deriv": (i32, [vp, vp, i64]),
        "idw_get_dxs": (i32, [vp, vp, i64]),
        "idw_warp_mesh_dev": (i32, [vp, vp, vp, vp]),
        "idw_warp_deriv_dev": (i32, [vp, vp, vp, vp, vp]),
        "idw_warp_deriv_fwd": (i32, [vp, vp, i64, vp, i64]),
        "idw_warp_deriv_fwd_dev": (i32, [vp, vp, vp, vp, vp]),
        "idw_get_sizes": (i32, [vp, P(i64), P(i64), P(i64), P(i64)]),
        "idw_get_ldef0": (i32, [vp, P(d)]),
        "idw_get_denominator0": (i32, [vp, vp, i64]),
        "idw_get_timings": (i32, [vp, vp]),
        "idw_fast_counts": (i32, [vp, vp]),
        "idw_set_symmetry_planes": (i32, [vp, vp, vp, i64]),
        "idw_batch_create": (i32, [P(vp), i32]),
        "idw_batch_destroy": (i32, [vp]),
        "idw_batch_add_warp": (i32, [vp, vp, vp, vp]),
        "idw_batch_add_warp_deriv": (i32, [vp, vp, vp, vp, vp]),
        "idw_batch_run": (i32, [vp, vp]),
        "idw_batch_info": (i32, [vp, P(i32), vp]),
        "idw_set_restart_file": (i32, [vp, ctypes.c_char_p]),
        "idw_get_restart_info": (i32, [vp, P(i32), vp, i64]),
        "idw_mixed_near_wall_floor": (d, [d, d, d]),
        "idw_batch_stats": (i32, [vp, P(i64), P(i64)]),
        "idw_device_bytes": (i32, [vp, P(i64)]),
        "idw_warp_deriv_dev_noalloc": (i32, [vp, vp, vp, vp, vp]),
        "idw_warp_mesh_async": (i32, [vp]),
        "idw_warp_deriv_async": (i32, [vp, vp, i64]),
        "idw_wait": (i32, [vp]),
        "idw_tune": (i32, [vp, i32]),
        "idw_get_tuning": (i32, [vp, vp, vp, i64]),
        "idw_set_tuning": (i32, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != IDW_OK:
        raise IdwError(rc, load().idw_last_error().decode())
    return rc


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else ctypes.c_void_p(0)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def point_reduce(pts, tol=1e-12):
    """pointreduce: (uniquePts (nU,3), link 1-based, nU)."""
    L = load()
    pts = f64(pts).reshape(-1, 3)
    uniq = np.zeros_like(pts)
    link = np.zeros(len(pts), np.int32)
    nu = ctypes.c_int64(0)
    check(L.idw_point_reduce(ptr(pts), len(pts), float(tol), ptr(uniq), ptr(link), ctypes.byref(nu)))
    return uniq[: nu.value].copy(), link, nu.value
