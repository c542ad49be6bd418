"""ctypes binding of ``libklampt_gpu.so`` (declarations in include/klampt_gpu.h).

There is no CPU fallback: if the shared library is missing this raises at
import of the first handle, naming the build command.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libklampt_gpu.so")

_c_int, _c_double, _c_float = ctypes.c_int, ctypes.c_double, ctypes.c_float
_vp = ctypes.c_void_p
_pd = ctypes.POINTER(ctypes.c_double)
_pf = ctypes.POINTER(ctypes.c_float)
_pi = ctypes.POINTER(ctypes.c_int32)
_pu8 = ctypes.POINTER(ctypes.c_uint8)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pvp = ctypes.POINTER(ctypes.c_void_p)


class DistSettings(ctypes.Structure):
    _fields_ = [("relErr", _c_double), ("absErr", _c_double), ("upperBound", _c_double)]


class DistResult(ctypes.Structure):
    _fields_ = [("d", _c_double), ("hasClosestPoints", _c_int), ("hasGradients", _c_int),
                ("cp1", _c_double * 3), ("cp2", _c_double * 3), ("grad1", _c_double * 3),
                ("grad2", _c_double * 3), ("elem1", _c_int), ("elem2", _c_int)]


class Params(ctypes.Structure):
    """kg_params: engine parameters of a context (include/klampt_gpu.h)."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "cpw", "leaf_min", "scap", "waves_per_cu", "static_pct", "nsub", "tail_lanes", "tail_reps", "spill_cap",
        "dist_ipw", "edge_w0", "sdf_lanes", "prep_group", "force_recheck", "edge_budget_mb",
        "edge_growth")]


# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "kg_last_error": [],
    "kg_version": [],
    "kg_ctx_create": [_c_int, _pvp],
    "kg_ctx_destroy": [_vp],
    "kg_ctx_synchronize": [_vp],
    "kg_ctx_get_params": [_vp, ctypes.POINTER(Params)],
    "kg_ctx_set_params": [_vp, ctypes.POINTER(Params)],
    "kg_ctx_status_dev": [_vp, _vp, _vp],
    "kg_geom_create_trimesh": [_vp, _pd, _c_int, _pi, _c_int, _pvp],
    "kg_geom_create_sdf": [_vp, _pd, _pd, _pi, _pf, _pvp],
    "kg_geom_destroy": [_vp],
    "kg_geom_clone": [_vp, _pvp],
    "kg_geom_get_bvh": [_vp, _pf, _c_int, _pi],
    "kg_distance_jacobian_batch": [_vp, _c_int, _pf, _c_int, _c_float, _pf, _pf, _pf],
    "kg_distance_jacobian_batch_dev": [_vp, _c_int, _vp, _c_int, _c_float, _vp, _vp, _vp, _vp],
    "kg_geom_create_pointcloud": [_vp, _pd, _c_int, _pvp],
    "kg_geom_mesh_to_sdf": [_vp, _c_double, _c_int, _pvp],
    "kg_geom_get_sdf": [_vp, _pi, _pd, _pd, _pf],
    "kg_geom_set_transform": [_vp, _pd, _pd],
    "kg_geom_get_transform": [_vp, _pd, _pd],
    "kg_geom_get_bb": [_vp, _pd, _pd],
    "kg_geom_set_collision_margin": [_vp, _c_double],
    "kg_geom_get_collision_margin": [_vp, _pd],
    "kg_geom_collide": [_vp, _vp, _pi],
    "kg_geom_within_distance": [_vp, _vp, _c_double, _pi],
    "kg_geom_distance": [_vp, _vp, ctypes.POINTER(DistSettings), ctypes.POINTER(DistResult)],
    "kg_geom_distance_point": [_vp, _pd, ctypes.POINTER(DistSettings), ctypes.POINTER(DistResult)],
    "kg_sdf_eval": [_vp, _pf, _c_int, _pf, _pf],
    "kg_robot_create": [_vp, _c_int, _pi, _pd, _pd, _pi, _pd, _pd, _pvp],
    "kg_robot_destroy": [_vp],
    "kg_robot_set_link_geom": [_vp, _c_int, _vp],
    "kg_robot_enable_self_collision": [_vp, _c_int, _c_int, _c_int],
    "kg_robot_self_collision_enabled": [_vp, _c_int, _c_int, _pi],
    "kg_robot_set_config": [_vp, _pd],
    "kg_robot_get_config": [_vp, _pd],
    "kg_robot_get_link_transform": [_vp, _c_int, _pd, _pd],
    "kg_robot_self_collides": [_vp, _pi],
    "kg_fk_batch": [_vp, _pf, _c_int, _pf],
    "kg_fk_batch_dev": [_vp, _vp, _c_int, _vp, _vp],
    "kg_fk_batch64": [_vp, _pd, _c_int, _pd],
    "kg_fk_batch64_dev": [_vp, _vp, _c_int, _vp, _vp],
    "kg_robot_set_self_pairs": [_vp, _pi, _c_int],
    "kg_world_create": [_vp, _pvp],
    "kg_world_destroy": [_vp],
    "kg_world_add_terrain": [_vp, _vp, _pi],
    "kg_world_add_rigid_object": [_vp, _vp, _pd, _pd, _pi],
    "kg_world_set_rigid_object_transform": [_vp, _c_int, _pd, _pd],
    "kg_world_set_terrain_geom": [_vp, _c_int, _vp],
    "kg_world_set_rigid_object_geom": [_vp, _c_int, _vp],
    "kg_world_add_robot": [_vp, _vp, _pi],
    "kg_world_num_ids": [_vp],
    "kg_world_set_collision_enabled": [_vp, _c_int, _c_int, _c_int],
    "kg_world_get_collision_enabled": [_vp, _c_int, _c_int, _pi],
    "kg_world_pair_list": [_vp, _c_int, _pi, _c_int, _pi],
    "kg_feasible_batch": [_vp, _c_int, _pf, _c_int, _pu8, _pu8, _pu8, _pi],
    "kg_feasible_batch_dev": [_vp, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp],
    "kg_feasible_batch_f64": [_vp, _c_int, _pd, _c_int, _pu8, _pu8, _pu8, _pi],
    "kg_feasible_batch_f64_dev": [_vp, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp],
    "kg_feasible_batch_dev_noalloc": [_vp, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp],
    "kg_distance_batch": [_vp, _c_int, _pf, _c_int, _pf],
    "kg_distance_batch_dev": [_vp, _c_int, _vp, _c_int, _vp, _vp],
    "kg_edge_batch": [_vp, _c_int, _pf, _pf, _c_int, _c_double, _c_int, _pf],
    "kg_edge_batch_f64": [_vp, _c_int, _pd, _pd, _c_int, _c_double, _c_int, _pf],
    "kg_edge_batch_dev": [_vp, _c_int, _vp, _vp, _c_int, _c_double, _c_int, _c_int, _vp, _vp, _vp],
    "kg_edge_batch_f64_dev": [_vp, _c_int, _vp, _vp, _c_int, _c_double, _c_int, _c_int, _vp, _vp, _vp],
    "kg_edge_sample_index": [_c_int, _c_int],
    "kg_sdf_penalty_fwd": [_vp, _vp, _pf, _c_int, _pf, _pi, _c_float, _pf],
    "kg_sdf_penalty_bwd": [_vp, _vp, _pf, _c_int, _pf, _pi, _c_float, _pf, _pf],
    "kg_sdf_penalty_fwd_dev": [_vp, _vp, _vp, _c_int, _vp, _pi, _c_float, _vp, _vp],
    "kg_sdf_penalty_bwd_dev": [_vp, _vp, _vp, _c_int, _vp, _pi, _c_float, _vp, _vp, _vp],
    "kg_world_last_stats": [_vp, _pi64, _c_int],
    "kg_feasible_multi": [_c_int, _pvp, _pi, _pi, _pf, _pu8],
    "kg_feasible_multi_dev": [_c_int, _pvp, _pi, _pi, _vp, _vp, _vp],
    "kg_autotune": [_c_int, _pvp, _pi, _pi, _vp, _pi, _pf],
    "kg_serve_create": [_c_int, _pvp, _pi, _pi, _vp, _vp, _c_int, _pvp, _pi, _pi, _vp, _vp, _c_double, _c_int,
                        _c_int, _vp, _pvp],
    "kg_serve_run": [_vp, _vp],
    "kg_world_num_obstacles": [_vp, _pi],
    "kg_sweep_batch": [_vp, _c_int, _pd, _pd, _c_int, _c_double, _pd, _pd, _pd],
    "kg_sweep_batch_dev": [_vp, _c_int, _vp, _vp, _c_int, _c_double, _vp, _vp, _vp, _vp],
    "kg_serve_captures": [_vp, _pi],
    "kg_serve_destroy": [_vp],
    "kg_ctx_set_tuning": [_vp, _c_int, _c_int, _c_int, _c_int],
    "kg_ctx_kernel_times": [_vp, _c_int, _pi64, _pi],
    "kg_ctx_set_timing": [_vp, _c_int],
}
_RESTYPE = {"kg_last_error": ctypes.c_char_p}

_lib = None


class KlamptGPUError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _share_torch_hip_runtime():
    """PyTorch-ROCm wheels bundle their own libamdhip64 / libhsa-runtime64
    with the same SONAMEs as /opt/rocm.  Whichever is loaded first serves the
    whole process; two copies cannot share a device.  Loading torch first
    makes this library bind to torch's runtime, so tensors and kernels live
    in one HIP runtime.  Without torch, /opt/rocm's runtime is used."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is None:
        _share_torch_hip_runtime()
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                f"or `python klampt_amd/build.py` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().kg_last_error()
        raise KlamptGPUError(rc, msg.decode() if msg else "unknown error")
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))
