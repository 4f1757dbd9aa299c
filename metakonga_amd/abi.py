"""ctypes mirror of include/xdem.h and the loader of libxdem.so.

The layouts follow the reference structs of include/mphysics_cuda_dec.cuh:17-99
(CUDA vector_types alignment: double3 = 24 B / 8-aligned, double4 = 32 B /
16-aligned); sizes are asserted at import.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxdem.so")


class uint3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]


class double3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]

    def as_tuple(self):
        return (self.x, self.y, self.z)


class double4(C.Structure):
    # 16-byte alignment of CUDA's double4 is reproduced by explicit padding in
    # the structs that contain one (ctypes cannot over-align).
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("w", C.c_double)]


def _d3(v):
    return double3(float(v[0]), float(v[1]), float(v[2]))


def _d4(v):
    return double4(float(v[0]), float(v[1]), float(v[2]), float(v[3]))


class DeviceParameters(C.Structure):
    """device_parameters — include/mphysics_cuda_dec.cuh:17-30 (112 B)."""
    _fields_ = [
        ("rollingCondition", C.c_bool),
        ("np", C.c_uint32),
        ("nsphere", C.c_uint32),
        ("ncell", C.c_uint32),
        ("grid_size", uint3),
        ("dt", C.c_double),
        ("half2dt", C.c_double),
        ("cell_size", C.c_double),
        ("cohesion", C.c_double),
        ("gravity", double3),
        ("world_origin", double3),
    ]


class PolygonInfo(C.Structure):
    """device_polygon_info — include/mphysics_cuda_dec.cuh:32-41 (152 B)."""
    _fields_ = [("id", C.c_int32), ("P", double3), ("Q", double3), ("R", double3),
                ("V", double3), ("W", double3), ("N", double3)]


class PlaneInfo(C.Structure):
    """device_plane_info — include/mphysics_cuda_dec.cuh:43-55 (232 B)."""
    _fields_ = [("l1", C.c_double), ("l2", C.c_double)] + [
        (n, double3) for n in ("u1", "u2", "uw", "xw", "pa", "pb", "w2", "w3", "w4")]


class PolygonMassInfo(C.Structure):
    """device_polygon_mass_info — include/mphysics_cuda_dec.cuh:57-65 (160 B)."""
    _fields_ = [("origin", double3), ("_pad0", C.c_double), ("ep", double4), ("vel", double3),
                ("omega", double3), ("force", double3), ("moment", double3)]


class CylinderInfo(C.Structure):
    """device_cylinder_info — include/mphysics_cuda_dec.cuh:67-76 (176 B)."""
    _fields_ = [("len", C.c_double), ("rbase", C.c_double), ("rtop", C.c_double),
                ("pbase", double3), ("ptop", double3), ("origin", double3), ("vel", double3),
                ("omega", double3), ("ep", double4)]


class DrumInfo(C.Structure):
    """xdem_drum_info (80 B): inner cylindrical shell with optional end caps,
    spinning at omega rad/s about +axis."""
    _fields_ = [("origin", double3), ("axis", double3), ("radius", C.c_double),
                ("half_length", C.c_double), ("omega", C.c_double), ("caps", C.c_int32),
                ("pad_", C.c_int32)]


class Motion(C.Structure):
    """xdem_motion (64 B): constant velocity and angular velocity from
    start_time (to end_time when end_time > start_time)."""
    _fields_ = [("velocity", double3), ("omega", double3), ("start_time", C.c_double),
                ("end_time", C.c_double)]


OBJECT_PLANE, OBJECT_POLYGON = 0, 1


class ContactProperty(C.Structure):
    """device_contact_property — include/mphysics_cuda_dec.cuh:78-88 (88 B)."""
    _fields_ = [(n, C.c_double) for n in
                ("Ei", "Ej", "pri", "prj", "Gi", "Gj", "rest", "fric", "rfric", "coh", "sratio")]


class ForceConstant(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("kn", "vn", "ks", "vs", "mu", "ms")]


class Material(C.Structure):
    _fields_ = [("density", C.c_double), ("youngs", C.c_double), ("poisson", C.c_double),
                ("shear", C.c_double)]


class ContactPairDesc(C.Structure):
    _fields_ = [("restitution", C.c_double), ("friction", C.c_double),
                ("rolling_friction", C.c_double), ("cohesion", C.c_double),
                ("stiffness_ratio", C.c_double), ("wall", Material)]


class SimDesc(C.Structure):
    _fields_ = [("np", C.c_uint32), ("dt", C.c_double), ("gravity", double3),
                ("world_origin", double3), ("cell_size", C.c_double), ("grid_size", uint3),
                ("rolling", C.c_bool), ("tcm", C.c_int32), ("particle", Material),
                ("pp", ContactPairDesc)]


# the older entry points (include/xdem_setb.h, cu_dem_dec.cuh:41-87)
class BDeviceParameters(C.Structure):
    _fields_ = [("np", C.c_uint32), ("nsp", C.c_uint32), ("ncell", C.c_uint32),
                ("grid_size", uint3), ("dt", C.c_double), ("half2dt", C.c_double),
                ("cell_size", C.c_double), ("gravity", double3), ("world_origin", double3),
                ("kn", C.c_double), ("vn", C.c_double), ("ks", C.c_double), ("vs", C.c_double),
                ("mu", C.c_double), ("cohesive", C.c_double)]


class BPlaneInfo(C.Structure):
    _fields_ = [("l1", C.c_double), ("l2", C.c_double)] + [
        (n, double3) for n in ("u1", "u2", "uw", "xw", "pa", "pb")]


class BPolygon(C.Structure):
    _fields_ = [(n, double3) for n in ("P", "Q", "R", "V", "W", "N")]


class BMassInfo(C.Structure):
    _fields_ = [("pos", double3), ("vel", double3)]


_SIZES = {DeviceParameters: 112, PolygonInfo: 152, PlaneInfo: 232, PolygonMassInfo: 160,
          CylinderInfo: 176, ContactProperty: 88, BDeviceParameters: 144, BPlaneInfo: 160,
          BPolygon: 144, BMassInfo: 48, DrumInfo: 80, Motion: 64}
for _t, _s in _SIZES.items():
    assert C.sizeof(_t) == _s, (_t.__name__, C.sizeof(_t), _s)
assert DeviceParameters.dt.offset == 32 and DeviceParameters.world_origin.offset == 88
assert PolygonMassInfo.ep.offset == 32 and CylinderInfo.ep.offset == 144


def struct_array(ctype, items):
    arr = (ctype * len(items))()
    for i, it in enumerate(items):
        arr[i] = it
    return arr


def ptr(a):
    """Host pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


class XdemError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Load libxdem.so (built in-tree by __graft_entry__.build()).  Raises if absent:
    there is no CPU fallback for the product path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise XdemError(f"{p} not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(p)
    _declare(lib)
    if path is None:
        _lib = lib
    return lib


def check(lib, rc):
    if rc != 0:
        msg = lib.xdem_last_error()
        raise XdemError(f"xdem status {rc}: {msg.decode() if msg else ''}")
    return rc


def _declare(lib):
    vp, u32, i32, f64 = C.c_void_p, C.c_uint32, C.c_int32, C.c_double
    lib.xdem_last_error.restype = C.c_char_p
    lib.xdem_last_error.argtypes = []
    sig = {
        "xdem_version": ([], i32),
        "xdem_device_count": ([], i32),
        "xdem_synchronize": ([], i32),
        "xdem_malloc": ([C.POINTER(vp), C.c_size_t], i32),
        "xdem_free": ([vp], i32),
        "xdem_memset": ([vp, C.c_int, C.c_size_t], i32),
        "xdem_memcpy_htod": ([vp, vp, C.c_size_t], i32),
        "xdem_memcpy_dtoh": ([vp, vp, C.c_size_t], i32),
        "xdem_memcpy_dtod": ([vp, vp, C.c_size_t], i32),
        "xdem_device_memory": ([C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], i32),
        "xdem_copy_kernel": ([vp, vp, C.c_size_t], i32),
        "xdem_set_symbolic_parameter": ([C.POINTER(DeviceParameters)], i32),
        "xdem_get_symbolic_parameter": ([C.POINTER(DeviceParameters)], i32),
        "xdem_set_stream": ([vp], i32),
        "xdem_vv_update_position_f64": ([vp, vp, vp, u32], i32),
        "xdem_vv_update_velocity_f64": ([vp] * 8 + [u32], i32),
        "xdem_calculate_hash_and_index_f64": ([vp, vp, vp, u32], i32),
        "xdem_calculate_hash_and_index_polygon_sphere_f64": ([vp, vp, u32, u32, vp], i32),
        "xdem_reorder_data_and_find_cell_start": ([vp] * 5 + [u32, u32], i32),
        "xdem_reorder_data_and_find_cell_start_noalloc": ([vp] * 5 + [u32, u32], i32),
        "xdem_calculate_p2p_f64": ([i32] + [vp] * 10 + [u32], i32),
        "xdem_plane_contact_force_f64": ([i32] + [vp] * 7 + [u32, vp], i32),
        "xdem_cube_contact_force_f64": ([i32] + [vp] * 7 + [u32, vp], i32),
        "xdem_particle_polygon_object_collision_f64":
            ([i32, vp, vp, vp, u32, u32] + [vp] * 10 + [u32], i32),
        "xdem_cylinder_hertzian_contact_force_f64":
            ([i32] + [vp] * 7 + [u32] + [vp] * 6, i32),
        "xdem_drum_contact_force_f64": ([i32] + [vp] * 7 + [u32] + [vp] * 5, i32),
        "xdem_reduction_d3": ([vp, u32, vp], i32),
        "xdem_update_polygon_object_data_f64": ([vp, vp, vp, vp, u32], i32),
        "xdem_grid_create": ([C.POINTER(vp)], i32),
        "xdem_grid_destroy": ([vp], i32),
        "xdem_grid_set_world_origin": ([vp, f64, f64, f64], i32),
        "xdem_grid_set_cell_size": ([vp, f64], i32),
        "xdem_grid_set_grid_size": ([vp, u32, u32, u32], i32),
        "xdem_grid_initialize": ([vp, u32], i32),
        "xdem_grid_detection": ([vp, vp, vp, u32, u32], i32),
        "xdem_grid_ncell": ([vp], u32),
        "xdem_grid_sorted_id": ([vp], vp),
        "xdem_grid_cell_start": ([vp], vp),
        "xdem_grid_cell_end": ([vp], vp),
        "xdem_grid_cell_id": ([vp], vp),
        "xdem_grid_body_id": ([vp], vp),
        "xdem_grid_get_cell_number": ([vp, f64, f64, f64, vp], i32),
        "xdem_grid_get_hash": ([vp, vp], u32),
        "xdem_sim_create": ([C.POINTER(SimDesc), vp, C.POINTER(vp)], i32),
        "xdem_sim_destroy": ([vp], i32),
        "xdem_sim_add_plane": ([vp, vp, vp], i32),
        "xdem_sim_add_cube": ([vp, vp, vp], i32),
        "xdem_sim_add_cylinder": ([vp, vp, vp], i32),
        "xdem_sim_set_polygons": ([vp, vp, vp, u32, vp, vp, u32], i32),
        "xdem_sim_add_drum": ([vp, vp, vp], i32),
        "xdem_sim_set_motion": ([vp, i32, u32, vp], i32),
        "xdem_sim_get_polygons": ([vp, vp, vp], i32),
        "xdem_sim_get_objects": ([vp, vp], i32),
        "xdem_sim_steps": ([vp], C.c_uint64),
        "xdem_sim_set_object_state": ([vp, u32, vp], i32),
        "xdem_sim_step": ([vp, u32], i32),
        "xdem_write_part_f64": ([C.c_char_p, f64, u32, vp], i32),
        "xdem_write_part_f32": ([C.c_char_p, f64, u32, vp], i32),
        "xdem_read_part": ([C.c_char_p, vp, vp, vp, vp, u32], i32),
        "xdem_write_bfr_f64": ([C.c_char_p, f64, u32, vp, vp, vp], i32),
        "xdem_write_bfr_f32": ([C.c_char_p, f64, u32, vp, vp, vp], i32),
        "xdem_read_bfr": ([C.c_char_p, u32, vp, vp, vp, vp, vp], i32),
        "xdem_sim_save_part": ([vp, C.c_char_p, u32], i32),
        "xdem_sim_save_final": ([vp, C.c_char_p], i32),
        "xdem_sim_load_bfr": ([vp, C.c_char_p], i32),
        "xdem_sim_save_checkpoint": ([vp, C.c_char_p], i32),
        "xdem_sim_load_checkpoint": ([vp, C.c_char_p], i32),
        "xdem_batch_save_checkpoint": ([vp, C.c_char_p], i32),
        "xdem_batch_load_checkpoint": ([vp, C.c_char_p, vp, u32, vp], i32),
        "xdem_checkpoint_verify": ([C.c_char_p, vp], i32),
        "xdem_sim_get_state": ([vp] * 8, i32),
        "xdem_sim_set_state": ([vp] * 8, i32),
        "xdem_sim_device_arrays": ([vp] + [C.POINTER(vp)] * 9, i32),
        "xdem_sim_np": ([vp], u32),
        "xdem_sim_set_release": ([vp, u32, u32], i32),
        "xdem_sim_set_release_groups": ([vp, u32, vp, u32], i32),
        "xdem_sim_active": ([vp], u32),
        "xdem_sim_enable_timing": ([vp, i32], i32),
        "xdem_sim_phase_times": ([vp, vp, vp], i32),
        "xdem_contact_pairs_f64": ([vp] * 4 + [u32, u32, vp, vp], i32),
        "xdem_sim_grid": ([vp], vp),
        "xdem_b_set_symbolic_parameter": ([C.POINTER(BDeviceParameters)], i32),
        "xdem_b_get_symbolic_parameter": ([C.POINTER(BDeviceParameters)], i32),
        "xdem_b_vv_update_position": ([vp, vp, vp, u32], i32),
        "xdem_b_vv_update_velocity": ([vp] * 6 + [u32], i32),
        "xdem_b_mergedata": ([vp, vp, u32, vp, u32], i32),
        "xdem_b_calculate_hash_and_index": ([vp, vp, vp, u32], i32),
        "xdem_b_reorder_data_and_find_cell_start": ([vp] * 5 + [u32, u32], i32),
        "xdem_b_calculate_p2p": ([vp] * 10 + [u32, u32], i32),
        "xdem_b_cube_hertzian_contact_force": ([vp] + [f64] * 5 + [vp] * 6 + [u32], i32),
        "xdem_b_shape_hertzian_contact_force":
            ([vp] * 5 + [f64] * 5 + [vp] * 7 + [u32] + [vp] * 4, i32),
        "xdem_b_update_polygons_position": ([vp, vp, vp, u32, vp], i32),
        "xdem_b_update_polygons_information": ([vp, vp, vp, u32], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res


def header_symbols(header_path: str) -> list[str]:
    """Names of the functions declared in a C header (for the export check)."""
    import re
    txt = open(header_path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(xdem_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)
