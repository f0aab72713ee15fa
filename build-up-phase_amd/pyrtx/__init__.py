"""pyrtx — Python binding of librtx.so (the C ABI in include/rtx.h) for tests and benchmarks.

The calls follow the Vulkan KHR ray-tracing flow of vulkan-raytraced-triangle/main.cpp:
    ctx.build_blas(geometries)          ~ createBLAS        (main.cpp:674-831)
    ctx.build_tlas(instance records)    ~ createTLAS        (main.cpp:833-949)
    ctx.trace_rays(tlas, sbt, camera)   ~ vkCmdTraceRaysKHR (main.cpp:1349-1355)
Device memory comes from torch (HIP) tensors; work is recorded on torch's current stream.  There is no CPU
fallback: if librtx.so is missing or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "librtx.so")

RTX_SUCCESS = 0
RTX_GEOMETRY_OPAQUE_BIT = 0x1
RTX_BUILD_ALLOW_UPDATE_BIT = 0x1
RTX_BUILD_ALLOW_COMPACTION_BIT = 0x2
RTX_BUILD_PREFER_FAST_TRACE_BIT = 0x4
RTX_BUILD_PREFER_FAST_BUILD_BIT = 0x8
RTX_BUILD_MORTON63_BIT = 0x10000
RTX_BUILD_NO_TREELET_BIT = 0x20000
RTX_BUILD_TREELET_ROUNDS_SHIFT = 18


def RTX_BUILD_TREELET_ROUNDS(r: int) -> int:
    return (r & 3) << RTX_BUILD_TREELET_ROUNDS_SHIFT


RTX_BUILD_MAX_WIDTH_SHIFT = 20


def RTX_BUILD_MAX_WIDTH(w: int) -> int:
    """Children per BVH node (2..8; 8 is the default)."""
    return (w & 7) << RTX_BUILD_MAX_WIDTH_SHIFT

RTX_INDEX_TYPE_UINT32 = 1
RTX_QUERY_CLOSEST_HIT = 0
RTX_QUERY_ANY_HIT = 1
RTX_RAY_FLAG_OPAQUE, RTX_RAY_FLAG_NO_OPAQUE = 0x1, 0x2
MISS = 0xFFFFFFFF

ERRORS = {-1: "OUT_OF_HOST_MEMORY", -2: "OUT_OF_DEVICE_MEMORY", -3: "INITIALIZATION_FAILED", -4: "DEVICE_LOST",
          -1000: "INVALID_ARGUMENT", -1001: "STACK_OVERFLOW", -1002: "INVALID_AS", -1003: "FORMAT"}


class RtxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rtx error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


# ---- C structures (include/rtx.h) -------------------------------------------------------------------------------
class BlasGeometry(ctypes.Structure):
    _fields_ = [("vertex_data_device", ctypes.c_void_p), ("vertex_stride", ctypes.c_uint64),
                ("max_vertex", ctypes.c_uint32), ("index_type", ctypes.c_uint32),
                ("index_data_device", ctypes.c_void_p), ("transform_data_device", ctypes.c_void_p),
                ("primitive_count", ctypes.c_uint32), ("primitive_offset", ctypes.c_uint32),
                ("first_vertex", ctypes.c_uint32), ("transform_offset", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("geometry_type", ctypes.c_uint32)]


class BuildSizes(ctypes.Structure):
    _fields_ = [("acceleration_structure_size", ctypes.c_uint64), ("build_scratch_size", ctypes.c_uint64),
                ("update_scratch_size", ctypes.c_uint64)]


class AsInfo(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("primitive_count", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
                ("slot_count", ctypes.c_uint32), ("max_depth", ctypes.c_uint32), ("geometry_count", ctypes.c_uint32),
                ("bounds", ctypes.c_float * 6), ("sah_cost", ctypes.c_double), ("content_flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


RTX_AS_CONTENT_NON_OPAQUE_BIT, RTX_AS_CONTENT_AABBS_BIT = 0x1, 0x2
RTX_RENDER_FORM_NONE, RTX_RENDER_FORM_WHOLE_PATH, RTX_RENDER_FORM_WAVEFRONT, RTX_RENDER_FORM_DEVICE_CHOSEN = 0, 1, 2, 3


class Sbt(ctypes.Structure):
    _fields_ = [("hit_records_device", ctypes.c_void_p), ("hit_record_count", ctypes.c_uint32),
                ("sbt_record_offset", ctypes.c_uint32), ("sbt_record_stride", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("miss_color", ctypes.c_float * 3),
                ("miss_alpha_reserved", ctypes.c_float)]


class Camera(ctypes.Structure):
    _fields_ = [("position", ctypes.c_float * 3), ("y_fov_degree", ctypes.c_float)]


class TraceParams(ctypes.Structure):
    _fields_ = [("ray_flags", ctypes.c_uint32), ("cull_mask", ctypes.c_uint32), ("tmin", ctypes.c_float),
                ("tmax", ctypes.c_float)]


class TraceStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_uint64), ("node_fetches", ctypes.c_uint64), ("triangle_tests", ctypes.c_uint64),
                ("instance_entries", ctypes.c_uint64)]


class RenderParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("rr_depth", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("sample_offset", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("light_triangles_device", ctypes.c_void_p), ("light_triangle_count", ctypes.c_uint32),
                ("ray_offset", ctypes.c_float)]


class RenderView(ctypes.Structure):
    _fields_ = [("camera", Camera), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("light_first", ctypes.c_uint32), ("light_count", ctypes.c_uint32),
                ("clip_min", ctypes.c_float * 3), ("clip_max", ctypes.c_float * 3)]


RTX_RENDER_GRAPH_BIT = 0x1
RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT, RTX_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER = 0x4, 0x8
RTX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES, RTX_RAY_FLAG_CULL_FRONT_FACING_TRIANGLES = 0x10, 0x20
RTX_RAY_FLAG_CULL_OPAQUE, RTX_RAY_FLAG_CULL_NO_OPAQUE, RTX_RAY_FLAG_SKIP_TRIANGLES = 0x40, 0x80, 0x100
RTX_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT, RTX_INSTANCE_TRIANGLE_FLIP_FACING_BIT = 0x1, 0x2
RTX_INSTANCE_FORCE_OPAQUE_BIT, RTX_INSTANCE_FORCE_NO_OPAQUE_BIT = 0x4, 0x8
RTX_PROGRAM_RECORD_COLOR, RTX_PROGRAM_BARYCENTRIC, RTX_PROGRAM_MATCH_BARYCENTRIC, RTX_PROGRAM_LAMBERTIAN = 0, 1, 2, 3
RTX_ANY_HIT_ACCEPT, RTX_ANY_HIT_CUTOUT = 0, 1
RTX_GEOMETRY_TYPE_TRIANGLES, RTX_GEOMETRY_TYPE_AABBS = 0, 1
RTX_INTERSECTION_BOX, RTX_INTERSECTION_SPHERE = 0, 1
RTX_RAY_FLAG_SKIP_AABBS = 0x200
RTX_FILM_SRGB_BIT, RTX_FILM_SWAP_RB_BIT, RTX_FILM_OPAQUE_BIT = 0x1, 0x2, 0x4
RTX_FILM_REFERENCE_BYTES = RTX_FILM_SWAP_RB_BIT
RTX_COPY_MODE_CLONE, RTX_COPY_MODE_COMPACT = 0, 1
RTX_ERROR_OUT_OF_DEVICE_MEMORY = -2
RTX_ERROR_INVALID_ARGUMENT = -1000
RTX_ERROR_FORMAT = -1003


class ObjMesh(ctypes.Structure):
    _fields_ = [("vertex_count", ctypes.c_uint32), ("triangle_count", ctypes.c_uint32),
                ("geometry_count", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("vertices", ctypes.POINTER(ctypes.c_float)), ("indices", ctypes.POINTER(ctypes.c_uint32)),
                ("geometry_first_triangle", ctypes.POINTER(ctypes.c_uint32))]


class BlasBuildInfo(ctypes.Structure):
    _fields_ = [("geometries", ctypes.c_void_p), ("geometry_count", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("dst_device", ctypes.c_void_p), ("dst_size", ctypes.c_uint64)]


class NrcConfig(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint32), ("batch_size", ctypes.c_uint32), ("learning_rate", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("epsilon", ctypes.c_float),
                ("scene_min", ctypes.c_float * 3), ("scene_max", ctypes.c_float * 3),
                ("cache_vertex", ctypes.c_uint32), ("train_suffix", ctypes.c_uint32)]


NRC_PARAM_COUNT = 23888


class TraceConfig(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("lds_stack_depth", ctypes.c_uint32), ("ray_chunk", ctypes.c_uint32),
                ("refill_lanes", ctypes.c_uint32), ("refill_interval", ctypes.c_uint32),
                ("leaf_lanes", ctypes.c_uint32), ("stack_limit", ctypes.c_uint32)]


RTX_TRACE_RAY_CHUNK_AUTO = 1
RTX_TRACE_STACK_LIMIT_DERIVED = 0xFFFFFFFF


# symbol -> (restype, argtypes); the exported surface of include/rtx.h
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
SIGNATURES = {
    "rtx_create_context": (ctypes.c_int, [ctypes.c_int, _P, ctypes.POINTER(_P)]),
    "rtx_destroy_context": (ctypes.c_int, [_P]),
    "rtx_last_error_string": (ctypes.c_char_p, [_P]),
    "rtx_sync": (ctypes.c_int, [_P]),
    "rtx_get_stream": (_P, [_P]),
    "rtx_get_proc_addr": (_P, [ctypes.c_char_p]),
    "rtx_get_blas_build_sizes": (ctypes.c_int, [_P, _P, _U32, _U32, ctypes.POINTER(BuildSizes)]),
    "rtx_get_tlas_build_sizes": (ctypes.c_int, [_P, _U32, _U32, ctypes.POINTER(BuildSizes)]),
    "rtx_cmd_build_blas": (ctypes.c_int, [_P, _P, _U32, _U32, _P, _U64, _P, _U64, ctypes.POINTER(_U64)]),
    "rtx_cmd_build_tlas": (ctypes.c_int, [_P, _P, _U32, _U32, _P, _U64, _P, _U64, ctypes.POINTER(_U64)]),
    "rtx_get_as_info": (ctypes.c_int, [_P, _U64, ctypes.POINTER(AsInfo)]),
    "rtx_cmd_trace_rays": (ctypes.c_int, [_P, _U64, ctypes.POINTER(Sbt), ctypes.POINTER(Camera),
                                          ctypes.POINTER(TraceParams), _U32, _U32, _P]),
    "rtx_cmd_trace_ray_buffer": (ctypes.c_int, [_P, _U64, _P, _U32, _U32, _U32, _U32, _P, _P]),
    "rtx_cmd_trace_ray_buffer_noalloc": (ctypes.c_int, [_P, _U64, _P, _U32, _U32, _U32, _U32, _P, _P]),
    "rtx_set_trace_stats": (ctypes.c_int, [_P, ctypes.c_int]),
    "rtx_get_trace_stats": (ctypes.c_int, [_P, ctypes.POINTER(TraceStats)]),
    "rtx_set_trace_config": (ctypes.c_int, [_P, ctypes.POINTER(TraceConfig)]),
    "rtx_get_trace_config": (ctypes.c_int, [_P, ctypes.POINTER(TraceConfig)]),
    "rtx_cmd_render": (ctypes.c_int, [_P, _U64, ctypes.POINTER(Sbt), ctypes.POINTER(Camera),
                                      ctypes.POINTER(RenderParams), _P, _P]),
    "rtx_get_render_ray_count": (ctypes.c_int, [_P, ctypes.POINTER(_U64)]),
    "rtx_get_render_form": (ctypes.c_int, [_P, ctypes.POINTER(_U32)]),
    "rtx_nrc_create": (ctypes.c_int, [_P, ctypes.POINTER(NrcConfig), ctypes.POINTER(_P)]),
    "rtx_nrc_destroy": (ctypes.c_int, [_P]),
    "rtx_nrc_cmd_infer": (ctypes.c_int, [_P, _P, _U32, _P]),
    "rtx_nrc_cmd_train": (ctypes.c_int, [_P, _P, _U32]),
    "rtx_nrc_get_params": (ctypes.c_int, [_P, _P]),
    "rtx_nrc_set_params": (ctypes.c_int, [_P, _P]),
    "rtx_nrc_get_status": (ctypes.c_int, [_P, ctypes.POINTER(_U32), ctypes.POINTER(ctypes.c_float)]),
    "rtx_cmd_render_nrc": (ctypes.c_int, [_P, _U64, ctypes.POINTER(Sbt), ctypes.POINTER(Camera),
                                          ctypes.POINTER(RenderParams), _P, _U32, _P, _P]),
    "rtx_nrc_debug_gradient": (ctypes.c_int, [_P, _P, _U32, _P, _P]),
    "rtx_debug_lbvh": (ctypes.c_int, [_P, _P, _U32, _U32, _P, _P, _P, _P]),
    "rtx_debug_sort": (ctypes.c_int, [_P, _P, _P, _U32, _U32]),
    "rtx_autotune_trace": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _U32, _U32, _P, _U32, ctypes.POINTER(TraceConfig),
                                          ctypes.POINTER(ctypes.c_float)]),
    "rtx_autotune_render": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _P, _P, _P, _U32, ctypes.POINTER(TraceConfig),
                                           ctypes.POINTER(ctypes.c_float)]),
    "rtx_get_blas_batch_build_sizes": (ctypes.c_int, [_P, _P, _U32, _U32, _P, ctypes.POINTER(_U64)]),
    "rtx_cmd_build_blas_batch": (ctypes.c_int, [_P, _P, _U32, _U32, _P, _U64, _P]),
    "rtx_cmd_render_batch": (ctypes.c_int, [_P, _U64, ctypes.POINTER(Sbt), _P, _U32, ctypes.POINTER(RenderParams),
                                            _P, _P]),
    "rtx_autotune_render_batch": (ctypes.c_int, [_P, _U64, ctypes.POINTER(Sbt), _P, _U32,
                                                 ctypes.POINTER(RenderParams), _P, _U32, ctypes.POINTER(TraceConfig),
                                                 ctypes.POINTER(ctypes.c_float)]),
    "rtx_cmd_film_to_rgba8": (ctypes.c_int, [_P, _P, _U32, _U32, _U32, _P]),
    "rtx_get_as_compacted_size": (ctypes.c_int, [_P, _U64, ctypes.POINTER(_U64)]),
    "rtx_cmd_copy_as": (ctypes.c_int, [_P, _U64, _U32, _P, _U64, ctypes.POINTER(_U64)]),
    "rtx_obj_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ObjMesh)]),
    "rtx_obj_parse": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ObjMesh)]),
    "rtx_obj_free": (None, [ctypes.POINTER(ObjMesh)]),
    "rtx_cmd_update_blas": (ctypes.c_int, [_P, _P, _U32, _U64, _P, _U64]),
    "rtx_debug_treelet": (ctypes.c_int, [_P, _P, _U32, _U32, _U32, _P, _P, _P, _P, _P]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load librtx.so and bind every C-ABI symbol.  Raises if the library is missing."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RtxError(-3, f"{path} not built: run __graft_entry__.build() or make -C {PKG_DIR}")
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if path == LIB_PATH:
        _lib = L
    return L


# ---- host mirror of the reference flow ---------------------------------------------------------------------------
@dataclass
class AccelerationStructure:
    storage: object     # torch uint8 tensor holding the AS
    address: int
    count: int          # triangles or instances
    keepalive: tuple = ()


def _torch():
    import torch
    return torch


def _dev_tensor(arr, device):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(arr)).to(device)


class Context:
    """One rtx_context on a HIP device, recording on torch's current stream."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RtxError(-3, "no GPU: librtx has no CPU path")
        self.lib = load_library()
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        h = ctypes.c_void_p()
        self._check(self.lib.rtx_create_context(device, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h)))
        self.handle = h

    def _check(self, code):
        if code != RTX_SUCCESS:
            msg = self.lib.rtx_last_error_string(self.handle).decode() if getattr(self, "handle", None) else ""
            raise RtxError(code, msg)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.rtx_destroy_context(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._check(self.lib.rtx_sync(self.handle))

    # -- acceleration structures ---------------------------------------------------------------------------------
    def _geometry_structs(self, geoms):
        """geoms: list of dicts with device tensors: vertices (float32 [V,3]), indices (uint32/int32 [T,3]),
        transform (float32 [3,4]) or None."""
        arr = (BlasGeometry * max(1, len(geoms)))()
        for i, g in enumerate(geoms):
            if "aabbs" in g:  # procedural geometry: float32 [N, 6] boxes
                b = g["aabbs"]
                arr[i].vertex_data_device = b.data_ptr()
                arr[i].vertex_stride = 24
                arr[i].primitive_count = b.numel() // 6
                arr[i].geometry_type = RTX_GEOMETRY_TYPE_AABBS
                arr[i].flags = RTX_GEOMETRY_OPAQUE_BIT if g.get("opaque", True) else 0
                continue
            v, ix, x = g["vertices"], g["indices"], g.get("transform")
            arr[i].vertex_data_device = v.data_ptr()
            arr[i].vertex_stride = 12
            arr[i].max_vertex = max(0, v.shape[0] - 1)
            arr[i].index_type = RTX_INDEX_TYPE_UINT32
            arr[i].index_data_device = ix.data_ptr()
            arr[i].transform_data_device = x.data_ptr() if x is not None else None
            arr[i].primitive_count = ix.numel() // 3
            arr[i].primitive_offset = 0
            arr[i].first_vertex = 0
            arr[i].transform_offset = 0
            arr[i].flags = RTX_GEOMETRY_OPAQUE_BIT if g.get("opaque", True) else 0
        return arr

    def blas_build_sizes(self, geoms, flags=RTX_BUILD_PREFER_FAST_TRACE_BIT):
        arr = self._geometry_structs(geoms)
        sz = BuildSizes()
        self._check(self.lib.rtx_get_blas_build_sizes(self.handle, ctypes.cast(arr, _P), len(geoms), flags,
                                                      ctypes.byref(sz)))
        return sz.acceleration_structure_size, sz.build_scratch_size

    def build_blas(self, geoms, flags=RTX_BUILD_PREFER_FAST_TRACE_BIT, scratch=None) -> AccelerationStructure:
        torch = _torch()
        arr = self._geometry_structs(geoms)
        sz = BuildSizes()
        self._check(self.lib.rtx_get_blas_build_sizes(self.handle, ctypes.cast(arr, _P), len(geoms), flags,
                                                      ctypes.byref(sz)))
        storage = torch.empty(int(sz.acceleration_structure_size), dtype=torch.uint8, device=self.device)
        if scratch is None or scratch.numel() < sz.build_scratch_size:
            scratch = torch.empty(int(sz.build_scratch_size), dtype=torch.uint8, device=self.device)
        addr = ctypes.c_uint64()
        self._check(self.lib.rtx_cmd_build_blas(self.handle, ctypes.cast(arr, _P), len(geoms), flags,
                                                storage.data_ptr(), storage.numel(), scratch.data_ptr(),
                                                scratch.numel(), ctypes.byref(addr)))
        n = sum((g["aabbs"].numel() // 6) if "aabbs" in g else (g["indices"].numel() // 3) for g in geoms)
        return AccelerationStructure(storage, addr.value, n, (scratch, tuple(geoms)))

    def update_blas(self, blas: AccelerationStructure, geoms):
        """rtx_cmd_update_blas: new vertex data (same geometry / triangle counts) for a BLAS built with
        RTX_BUILD_ALLOW_UPDATE_BIT."""
        torch = _torch()
        arr = self._geometry_structs(geoms)
        sz = BuildSizes()
        self._check(self.lib.rtx_get_blas_build_sizes(self.handle, ctypes.cast(arr, _P), len(geoms),
                                                      RTX_BUILD_ALLOW_UPDATE_BIT, ctypes.byref(sz)))
        scratch = torch.empty(max(256, int(sz.update_scratch_size)), dtype=torch.uint8, device=self.device)
        self._check(self.lib.rtx_cmd_update_blas(self.handle, ctypes.cast(arr, _P), len(geoms), blas.address,
                                                 scratch.data_ptr(), scratch.numel()))
        blas.keepalive = (blas.keepalive, scratch, tuple(geoms))
        return blas

    def build_blas_obj(self, vertices, indices, first, flags=RTX_BUILD_PREFER_FAST_TRACE_BIT):
        """One BLAS over an OBJ mesh (load_obj output): one geometry per OBJ group, all sharing the vertex and index
        buffers through primitive_offset."""
        torch = _torch()
        dv = _dev_tensor(np.ascontiguousarray(vertices, np.float32), self.device)
        di = _dev_tensor(np.ascontiguousarray(indices, np.uint32).view(np.int32), self.device)
        g = len(first) - 1
        arr = (BlasGeometry * max(1, g))()
        for k in range(g):
            arr[k].vertex_data_device = dv.data_ptr()
            arr[k].vertex_stride = 12
            arr[k].max_vertex = max(0, vertices.shape[0] - 1)
            arr[k].index_type = RTX_INDEX_TYPE_UINT32
            arr[k].index_data_device = di.data_ptr()
            arr[k].transform_data_device = None
            arr[k].primitive_count = int(first[k + 1] - first[k])
            arr[k].primitive_offset = int(first[k]) * 12
            arr[k].first_vertex = 0
            arr[k].transform_offset = 0
            arr[k].flags = RTX_GEOMETRY_OPAQUE_BIT
        sz = BuildSizes()
        self._check(self.lib.rtx_get_blas_build_sizes(self.handle, ctypes.cast(arr, _P), g, flags, ctypes.byref(sz)))
        storage = torch.empty(int(sz.acceleration_structure_size), dtype=torch.uint8, device=self.device)
        scratch = torch.empty(int(sz.build_scratch_size), dtype=torch.uint8, device=self.device)
        addr = ctypes.c_uint64()
        self._check(self.lib.rtx_cmd_build_blas(self.handle, ctypes.cast(arr, _P), g, flags, storage.data_ptr(),
                                                storage.numel(), scratch.data_ptr(), scratch.numel(),
                                                ctypes.byref(addr)))
        return AccelerationStructure(storage, addr.value, int(indices.shape[0]), (scratch, dv, di))

    def build_blas_batch(self, geom_lists, flags=RTX_BUILD_PREFER_FAST_TRACE_BIT, into=None, scratch=None):
        """rtx_cmd_build_blas_batch: one build pass over several BLASes (list of geometry lists as for
        build_blas).  Returns one AccelerationStructure per list.  into: existing BLASes (one per list) whose
        storage the build reuses (same addresses); scratch: a device uint8 tensor to build in."""
        torch = _torch()
        k = len(geom_lists)
        arrs = [self._geometry_structs(g) for g in geom_lists]
        infos = (BlasBuildInfo * k)()
        for b, (g, arr) in enumerate(zip(geom_lists, arrs)):
            infos[b].geometries = ctypes.cast(arr, _P)
            infos[b].geometry_count = len(g)
        as_sizes = (ctypes.c_uint64 * k)()
        scratch_size = ctypes.c_uint64()
        self._check(self.lib.rtx_get_blas_batch_build_sizes(self.handle, ctypes.cast(infos, _P), k, flags,
                                                            ctypes.cast(as_sizes, _P), ctypes.byref(scratch_size)))
        if into is not None:
            if len(into) != k or any(a.storage.numel() < as_sizes[b] for b, a in enumerate(into)):
                raise RtxError(RTX_ERROR_INVALID_ARGUMENT, "into: one BLAS per list, each at least its size query")
            storages = [a.storage for a in into]
        else:
            storages = [torch.empty(int(as_sizes[b]), dtype=torch.uint8, device=self.device) for b in range(k)]
        for b in range(k):
            infos[b].dst_device = storages[b].data_ptr()
            infos[b].dst_size = storages[b].numel()
        if scratch is None or scratch.numel() < scratch_size.value:
            scratch = torch.empty(int(scratch_size.value), dtype=torch.uint8, device=self.device)
        addrs = (ctypes.c_uint64 * k)()
        self._check(self.lib.rtx_cmd_build_blas_batch(self.handle, ctypes.cast(infos, _P), k, flags,
                                                      scratch.data_ptr(), scratch.numel(), ctypes.cast(addrs, _P)))
        out = []
        for b, g in enumerate(geom_lists):
            n = sum(x["indices"].numel() // 3 for x in g)
            out.append(AccelerationStructure(storages[b], addrs[b], n, (scratch, tuple(g))))
        return out

    def build_tlas(self, records, flags=RTX_BUILD_PREFER_FAST_TRACE_BIT, scratch=None,
                   into: AccelerationStructure | None = None) -> AccelerationStructure:
        """records: numpy structured array of 64-B instance records (oracle-compatible layout) or a device uint8
        tensor of n*64 bytes.  into: an existing TLAS whose storage the build reuses (same address)."""
        torch = _torch()
        if isinstance(records, np.ndarray):
            raw = np.ascontiguousarray(records).view(np.uint8).reshape(-1)
            inst = _dev_tensor(raw, self.device)
        else:
            inst = records
        n = inst.numel() // 64
        sz = BuildSizes()
        self._check(self.lib.rtx_get_tlas_build_sizes(self.handle, n, flags, ctypes.byref(sz)))
        if into is not None and into.storage.numel() >= sz.acceleration_structure_size:
            storage = into.storage
        else:
            storage = torch.empty(int(sz.acceleration_structure_size), dtype=torch.uint8, device=self.device)
        if scratch is None or scratch.numel() < sz.build_scratch_size:
            scratch = torch.empty(int(sz.build_scratch_size), dtype=torch.uint8, device=self.device)
        addr = ctypes.c_uint64()
        self._check(self.lib.rtx_cmd_build_tlas(self.handle, inst.data_ptr(), n, flags, storage.data_ptr(),
                                                storage.numel(), scratch.data_ptr(), scratch.numel(),
                                                ctypes.byref(addr)))
        return AccelerationStructure(storage, addr.value, n, (scratch, inst))

    def as_info(self, a: AccelerationStructure) -> AsInfo:
        info = AsInfo()
        self._check(self.lib.rtx_get_as_info(self.handle, a.address, ctypes.byref(info)))
        return info

    # -- tracing --------------------------------------------------------------------------------------------------
    def trace_rays(self, tlas: AccelerationStructure, records, camera, width, height, miss_color=(0.0, 0.0, 0.2),
                   tmin=0.0, tmax=100.0, cull_mask=0xFF, sbt_offset=0, sbt_stride=1, film=None,
                   ray_flags=RTX_RAY_FLAG_OPAQUE):
        """vkCmdTraceRaysKHR with the reference raygen (main.cpp:1033-1055) into an RGBA32F film [H, W, 4]."""
        torch = _torch()
        if isinstance(records, np.ndarray):
            rec = _dev_tensor(np.ascontiguousarray(records).view(np.uint8).reshape(-1), self.device)
        else:
            rec = records
        nrec = rec.numel() // 48
        if film is None:
            film = torch.empty((height, width, 4), dtype=torch.float32, device=self.device)
        sbt = Sbt(rec.data_ptr() if nrec else None, nrec, sbt_offset, sbt_stride, 0,
                  (ctypes.c_float * 3)(*miss_color), 0.0)
        cam = Camera((ctypes.c_float * 3)(*camera[:3]), float(camera[3]))
        params = TraceParams(ray_flags, cull_mask, tmin, tmax)
        self._check(self.lib.rtx_cmd_trace_rays(self.handle, tlas.address, ctypes.byref(sbt), ctypes.byref(cam),
                                                ctypes.byref(params), width, height, film.data_ptr()))
        self._keep = (rec,)
        return film

    def trace_ray_buffer(self, tlas: AccelerationStructure, rays, mode=RTX_QUERY_CLOSEST_HIT, cull_mask=0xFF,
                         out=None, ray_flags=0):
        """rays: device float32 tensor [N, 8].  Returns hits (uint32 view [N, 6]: t, u, v as float bits, prim,
        geom, inst) for closest-hit, or occluded uint32 [N] for any-hit."""
        torch = _torch()
        n = rays.shape[0]
        if mode == RTX_QUERY_CLOSEST_HIT:
            if out is None:
                out = torch.empty((n, 6), dtype=torch.int32, device=self.device)
            self._check(self.lib.rtx_cmd_trace_ray_buffer(self.handle, tlas.address, rays.data_ptr(), n, mode,
                                                          ray_flags, cull_mask, out.data_ptr(), None))
        else:
            if out is None:
                out = torch.empty((n,), dtype=torch.int32, device=self.device)
            self._check(self.lib.rtx_cmd_trace_ray_buffer(self.handle, tlas.address, rays.data_ptr(), n, mode,
                                                          ray_flags, cull_mask, None, out.data_ptr()))
        return out

    def _render_args(self, records, camera, lights, width, height, spp, max_depth, rr_depth, seed, sample_offset,
                     ray_offset):
        if isinstance(records, np.ndarray):
            rec = _dev_tensor(np.ascontiguousarray(records).view(np.uint8).reshape(-1), self.device)
        else:
            rec = records
        if isinstance(lights, np.ndarray):
            lt = _dev_tensor(np.ascontiguousarray(lights, np.float32).reshape(-1, 12), self.device)
        else:
            lt = lights
        nl = lt.shape[0] if lt is not None else 0
        nrec = rec.numel() // 48
        sbt = Sbt(rec.data_ptr() if nrec else None, nrec, 0, 1, 0, (ctypes.c_float * 3)(0, 0, 0), 0.0)
        cam = Camera((ctypes.c_float * 3)(*camera[:3]), float(camera[3]))
        p = RenderParams(width, height, spp, max_depth, rr_depth, seed, sample_offset, 0,
                         lt.data_ptr() if nl else None, nl, ray_offset)
        self._keep = (rec, lt)
        return sbt, cam, p

    def render(self, tlas: AccelerationStructure, records, camera, lights, width, height, spp=1, max_depth=8,
               rr_depth=3, seed=1234, sample_offset=0, ray_offset=1e-4, film=None, accum64=None):
        """Wavefront path tracer (rtx_cmd_render).  lights: numpy float32 [L, 12] (v0 v1 v2 emission) or a device
        tensor.  Returns the film [H, W, 4] (mean radiance of these samples)."""
        torch = _torch()
        sbt, cam, p = self._render_args(records, camera, lights, width, height, spp, max_depth, rr_depth, seed,
                                        sample_offset, ray_offset)
        if film is None:
            film = torch.empty((height, width, 4), dtype=torch.float32, device=self.device)
        self._check(self.lib.rtx_cmd_render(self.handle, tlas.address, ctypes.byref(sbt), ctypes.byref(cam),
                                            ctypes.byref(p), film.data_ptr(),
                                            accum64.data_ptr() if accum64 is not None else None))
        return film

    @staticmethod
    def _views(views):
        k = len(views)
        arr = (RenderView * k)()
        npix = 0
        for i, v in enumerate(views):
            arr[i].camera = Camera((ctypes.c_float * 3)(*v["camera"][:3]), float(v["camera"][3]))
            arr[i].width, arr[i].height = v["width"], v["height"]
            arr[i].light_first, arr[i].light_count = v["light_first"], v["light_count"]
            arr[i].clip_min = (ctypes.c_float * 3)(*v["clip_min"])
            arr[i].clip_max = (ctypes.c_float * 3)(*v["clip_max"])
            npix += v["width"] * v["height"]
        return arr, npix

    def render_batch(self, tlas: AccelerationStructure, records, views, lights, spp=1, max_depth=6, rr_depth=3,
                     seed=1234, sample_offset=0, ray_offset=1e-4, graph=False, films=None, accum64=None):
        """rtx_cmd_render_batch: views = list of dicts (camera, width, height, light_first, light_count, clip_min,
        clip_max).  Returns the concatenated films [sum(w*h), 4] (view order)."""
        torch = _torch()
        sbt, cam, p = self._render_args(records, (0, 0, 0, 60.0), lights, 1, 1, spp, max_depth, rr_depth, seed,
                                        sample_offset, ray_offset)
        p.flags = RTX_RENDER_GRAPH_BIT if graph else 0
        arr, npix = self._views(views)
        k = len(views)
        if films is None:
            films = torch.empty((npix, 4), dtype=torch.float32, device=self.device)
        self._check(self.lib.rtx_cmd_render_batch(self.handle, tlas.address, ctypes.byref(sbt),
                                                  ctypes.cast(arr, _P), k, ctypes.byref(p), films.data_ptr(),
                                                  accum64.data_ptr() if accum64 is not None else None))
        return films

    def autotune_render_batch(self, tlas: AccelerationStructure, records, views, lights, spp=1, max_depth=6,
                              rr_depth=3, seed=1234, ray_offset=1e-4, reps=2):
        """rtx_autotune_render_batch; the best configuration stays active.  Returns (config dict, ms per call)."""
        torch = _torch()
        sbt, cam, p = self._render_args(records, (0, 0, 0, 60.0), lights, 1, 1, spp, max_depth, rr_depth, seed,
                                        1 << 28, ray_offset)
        arr, npix = self._views(views)
        films = torch.empty((npix, 4), dtype=torch.float32, device=self.device)
        c = TraceConfig()
        ms = ctypes.c_float()
        self._check(self.lib.rtx_autotune_render_batch(self.handle, tlas.address, ctypes.byref(sbt),
                                                       ctypes.cast(arr, _P), len(views), ctypes.byref(p),
                                                       films.data_ptr(), reps, ctypes.byref(c), ctypes.byref(ms)))
        return self.config(), ms.value

    def autotune_render(self, tlas: AccelerationStructure, records, camera, lights, width, height, spp=1,
                        max_depth=8, rr_depth=3, seed=1234, ray_offset=1e-4, reps=2):
        """rtx_autotune_render: tune the traversal configuration on spp-sample renders of this workload; the best
        configuration stays active.  Returns (config dict, ms per render)."""
        torch = _torch()
        sbt, cam, p = self._render_args(records, camera, lights, width, height, spp, max_depth, rr_depth, seed,
                                        1 << 28, ray_offset)
        film = torch.empty((height, width, 4), dtype=torch.float32, device=self.device)
        c = TraceConfig()
        ms = ctypes.c_float()
        self._check(self.lib.rtx_autotune_render(self.handle, tlas.address, ctypes.byref(sbt), ctypes.byref(cam),
                                                 ctypes.byref(p), film.data_ptr(), reps, ctypes.byref(c),
                                                 ctypes.byref(ms)))
        return self.config(), ms.value

    def render_nrc(self, tlas: AccelerationStructure, records, camera, lights, nrc: "Nrc", width, height, spp=1,
                   train_paths=65536, max_depth=8, rr_depth=3, seed=1234, sample_offset=0, ray_offset=1e-4,
                   film=None, accum64=None, graph=False):
        """rtx_cmd_render_nrc: one training pass of train_paths paths + spp cache-terminated samples per pixel
        (graph: record the frame into a HIP graph once and replay it)."""
        torch = _torch()
        rec = _dev_tensor(np.ascontiguousarray(records).view(np.uint8).reshape(-1), self.device) \
            if isinstance(records, np.ndarray) else records
        lt = _dev_tensor(np.ascontiguousarray(lights, np.float32).reshape(-1, 12), self.device) \
            if isinstance(lights, np.ndarray) else lights
        if film is None:
            film = torch.empty((height, width, 4), dtype=torch.float32, device=self.device)
        nrec = rec.numel() // 48
        sbt = Sbt(rec.data_ptr(), nrec, 0, 1, 0, (ctypes.c_float * 3)(0, 0, 0), 0.0)
        cam = Camera((ctypes.c_float * 3)(*camera[:3]), float(camera[3]))
        p = RenderParams(width, height, spp, max_depth, rr_depth, seed, sample_offset,
                         RTX_RENDER_GRAPH_BIT if graph else 0, lt.data_ptr(), lt.shape[0], ray_offset)
        self._check(self.lib.rtx_cmd_render_nrc(self.handle, tlas.address, ctypes.byref(sbt), ctypes.byref(cam),
                                                ctypes.byref(p), nrc.handle, train_paths, film.data_ptr(),
                                                accum64.data_ptr() if accum64 is not None else None))
        self._keep = (rec, lt)
        return film

    def render_ray_count(self) -> int:
        n = ctypes.c_uint64()
        self._check(self.lib.rtx_get_render_ray_count(self.handle, ctypes.byref(n)))
        return n.value

    def render_form(self) -> int:
        """RTX_RENDER_FORM_* of the last render call."""
        f = ctypes.c_uint32()
        self._check(self.lib.rtx_get_render_form(self.handle, ctypes.byref(f)))
        return f.value

    def set_stats(self, enable: bool):
        self._check(self.lib.rtx_set_trace_stats(self.handle, int(enable)))

    def stats(self) -> dict:
        s = TraceStats()
        self._check(self.lib.rtx_get_trace_stats(self.handle, ctypes.byref(s)))
        return {"rays": s.rays, "node_fetches": s.node_fetches, "triangle_tests": s.triangle_tests,
                "instance_entries": s.instance_entries}

    def set_config(self, block_size=0, blocks_per_cu=0, lds_stack_depth=0, ray_chunk=0, refill_lanes=0,
                   refill_interval=0, leaf_lanes=0, stack_limit=0):
        c = TraceConfig(block_size, blocks_per_cu, lds_stack_depth, ray_chunk, refill_lanes, refill_interval,
                        leaf_lanes, stack_limit)
        self._check(self.lib.rtx_set_trace_config(self.handle, ctypes.byref(c)))

    def config(self) -> dict:
        c = TraceConfig()
        self._check(self.lib.rtx_get_trace_config(self.handle, ctypes.byref(c)))
        return {"block_size": c.block_size, "blocks_per_cu": c.blocks_per_cu, "lds_stack_depth": c.lds_stack_depth,
                "ray_chunk": c.ray_chunk, "refill_lanes": c.refill_lanes, "refill_interval": c.refill_interval,
                "leaf_lanes": c.leaf_lanes, "stack_limit": c.stack_limit}

    def autotune(self, tlas: AccelerationStructure, rays, mode=RTX_QUERY_CLOSEST_HIT, reps=2):
        """rtx_autotune_trace on a device ray tensor [N, 8] float32; the best configuration stays active.
        Returns (config dict, ms per launch)."""
        torch = _torch()
        n = rays.shape[0]
        res = torch.empty((n, 6) if mode == RTX_QUERY_CLOSEST_HIT else (n,), dtype=torch.int32, device=self.device)
        c = TraceConfig()
        ms = ctypes.c_float()
        self._check(self.lib.rtx_autotune_trace(self.handle, tlas.address, rays.data_ptr(), n, mode, res.data_ptr(),
                                                reps, ctypes.byref(c), ctypes.byref(ms)))
        return self.config(), ms.value

    def film_to_rgba8(self, film, flags=RTX_FILM_REFERENCE_BYTES):
        """rtx_cmd_film_to_rgba8: device float32 film [H, W, 4] -> device uint8 [H, W, 4]."""
        torch = _torch()
        h, w = film.shape[0], film.shape[1]
        out = torch.empty((h, w, 4), dtype=torch.uint8, device=self.device)
        self._check(self.lib.rtx_cmd_film_to_rgba8(self.handle, film.data_ptr(), w, h, flags, out.data_ptr()))
        return out

    def compacted_size(self, a: AccelerationStructure) -> int:
        n = ctypes.c_uint64()
        self._check(self.lib.rtx_get_as_compacted_size(self.handle, a.address, ctypes.byref(n)))
        return n.value

    def copy_as(self, a: AccelerationStructure, mode=RTX_COPY_MODE_COMPACT) -> AccelerationStructure:
        """rtx_cmd_copy_as into new storage (compacted size for COMPACT, the source's size for CLONE)."""
        torch = _torch()
        size = self.compacted_size(a) if mode == RTX_COPY_MODE_COMPACT else a.storage.numel()
        dst = torch.empty(size, dtype=torch.uint8, device=self.device)
        addr = ctypes.c_uint64()
        self._check(self.lib.rtx_cmd_copy_as(self.handle, a.address, mode, dst.data_ptr(), size,
                                             ctypes.byref(addr)))
        return AccelerationStructure(dst, addr.value, a.count, a.keepalive)

    # -- verification hooks ----------------------------------------------------------------------------------------
    def debug_lbvh(self, boxes, bits=30):
        """boxes: device float32 [n, 6].  Returns (sorted codes, order, left, right) as numpy arrays."""
        torch = _torch()
        n = boxes.shape[0]
        codes = torch.empty((max(n, 1),), dtype=torch.int64 if bits == 63 else torch.int32, device=self.device)
        order = torch.empty((max(n, 1),), dtype=torch.int32, device=self.device)
        left = torch.empty((max(n - 1, 1),), dtype=torch.int32, device=self.device)
        right = torch.empty((max(n - 1, 1),), dtype=torch.int32, device=self.device)
        self._check(self.lib.rtx_debug_lbvh(self.handle, boxes.data_ptr(), n, bits, codes.data_ptr(),
                                            order.data_ptr(), left.data_ptr(), right.data_ptr()))
        c = codes.cpu().numpy()[:n].view(np.uint64 if bits == 63 else np.uint32)
        return c, order.cpu().numpy()[:n].view(np.uint32), left.cpu().numpy()[:max(n - 1, 0)], \
            right.cpu().numpy()[:max(n - 1, 0)]

    def debug_treelet(self, boxes, rounds, bits=30):
        """boxes: device float32 [n, 6], n >= 2.  Returns (order, left, right, cost, height) after the build's
        refit + treelet refinement passes."""
        torch = _torch()
        n = boxes.shape[0]
        order = torch.empty((n,), dtype=torch.int32, device=self.device)
        left = torch.empty((n - 1,), dtype=torch.int32, device=self.device)
        right = torch.empty((n - 1,), dtype=torch.int32, device=self.device)
        cost = torch.empty((n - 1,), dtype=torch.float32, device=self.device)
        hgt = torch.empty((n - 1,), dtype=torch.int32, device=self.device)
        self._check(self.lib.rtx_debug_treelet(self.handle, boxes.data_ptr(), n, bits, rounds, order.data_ptr(),
                                               left.data_ptr(), right.data_ptr(), cost.data_ptr(), hgt.data_ptr()))
        return (order.cpu().numpy().view(np.uint32), left.cpu().numpy(), right.cpu().numpy(), cost.cpu().numpy(),
                hgt.cpu().numpy().view(np.uint32))

    def debug_sort(self, keys: np.ndarray):
        torch = _torch()
        kb = keys.dtype.itemsize
        n = keys.size
        k = torch.from_numpy(keys.view(np.int64 if kb == 8 else np.int32).copy()).to(self.device)
        v = torch.empty((max(n, 1),), dtype=torch.int32, device=self.device)
        self._check(self.lib.rtx_debug_sort(self.handle, k.data_ptr(), v.data_ptr(), n, kb))
        return k.cpu().numpy().view(keys.dtype), v.cpu().numpy()[:n].view(np.uint32)

    # -- scenes ---------------------------------------------------------------------------------------------------
    def upload_scene(self, scene, flags=RTX_BUILD_PREFER_FAST_TRACE_BIT, batched=False):
        """Build one BLAS per scene.blas entry (one batched build when batched=True) and the TLAS over scene
        instances.  Returns (blases, tlas)."""
        torch = _torch()
        lists = []
        from .scenes import AabbGeometry
        for geoms in scene.blas:
            dg = []
            for g in geoms:
                if isinstance(g, AabbGeometry):
                    dg.append({"aabbs": _dev_tensor(np.ascontiguousarray(g.boxes, np.float32).reshape(-1, 6),
                                                    self.device), "opaque": g.opaque})
                    continue
                d = {"vertices": _dev_tensor(g.vertices.astype(np.float32), self.device),
                     "indices": _dev_tensor(g.indices.astype(np.uint32).view(np.int32), self.device),
                     "transform": None if g.transform is None else
                     _dev_tensor(np.asarray(g.transform, np.float32).reshape(3, 4), self.device),
                     "opaque": g.opaque}
                dg.append(d)
            lists.append(dg)
        if batched:
            blases = self.build_blas_batch(lists, flags)
        else:
            blases = [self.build_blas(dg, flags) for dg in lists]
        rec = instance_records(scene, [b.address for b in blases])
        tlas = self.build_tlas(rec, flags)
        torch.cuda.current_stream(self.device)
        return blases, tlas


class Nrc:
    """Neural radiance cache bound to a Context (rtx_nrc_* functions)."""

    def __init__(self, ctx: "Context", scene_min, scene_max, seed=1, batch_size=16384, lr=1e-3, beta1=0.9,
                 beta2=0.99, eps=1e-8, cache_vertex=1, train_suffix=0):
        self.ctx = ctx
        cfg = NrcConfig(seed, batch_size, lr, beta1, beta2, eps, (ctypes.c_float * 3)(*scene_min),
                        (ctypes.c_float * 3)(*scene_max), cache_vertex, train_suffix)
        h = ctypes.c_void_p()
        ctx._check(ctx.lib.rtx_nrc_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.batch_size = batch_size

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.rtx_nrc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def infer(self, queries):
        """queries: device float32 [n, 12] -> device float32 [n, 3]."""
        torch = _torch()
        n = queries.shape[0]
        out = torch.empty((max(n, 1), 3), dtype=torch.float32, device=self.ctx.device)
        self.ctx._check(self.ctx.lib.rtx_nrc_cmd_infer(self.handle, queries.data_ptr(), n, out.data_ptr()))
        return out[:n]

    def train(self, samples):
        """samples: device float32 [n, 16] (query[12], target[3], pad)."""
        self.ctx._check(self.ctx.lib.rtx_nrc_cmd_train(self.handle, samples.data_ptr(), samples.shape[0]))

    def gradient(self, samples):
        torch = _torch()
        g = torch.empty(NRC_PARAM_COUNT, dtype=torch.float32, device=self.ctx.device)
        loss = torch.empty(4, dtype=torch.float32, device=self.ctx.device)
        self.ctx._check(self.ctx.lib.rtx_nrc_debug_gradient(self.handle, samples.data_ptr(), samples.shape[0],
                                                            g.data_ptr(), loss.data_ptr()))
        return g.cpu().numpy(), float(loss[0].item())

    def params(self) -> np.ndarray:
        p = np.zeros(NRC_PARAM_COUNT, np.float32)
        self.ctx._check(self.ctx.lib.rtx_nrc_get_params(self.handle, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self.ctx._check(self.ctx.lib.rtx_nrc_set_params(self.handle, p.ctypes.data))

    def status(self):
        st = ctypes.c_uint32()
        loss = ctypes.c_float()
        self.ctx._check(self.ctx.lib.rtx_nrc_get_status(self.handle, ctypes.byref(st), ctypes.byref(loss)))
        return st.value, loss.value


def load_obj(path=None, text=None):
    """rtx_obj_load / rtx_obj_parse (host only, no GPU needed).  Returns (vertices float32 [V, 3], indices uint32
    [T, 3], first uint32 [G + 1]) with geometry g = triangles first[g]:first[g + 1]."""
    lib = load_library()
    m = ObjMesh()
    if path is not None:
        code = lib.rtx_obj_load(os.fsencode(path), ctypes.byref(m))
    else:
        data = text.encode() if isinstance(text, str) else bytes(text)
        code = lib.rtx_obj_parse(data, len(data), ctypes.byref(m))
    if code != RTX_SUCCESS:
        raise RtxError(code, "OBJ " + (str(path) if path is not None else "text"))
    try:
        v = np.ctypeslib.as_array(m.vertices, (m.vertex_count * 3,)).reshape(-1, 3).copy() if m.vertex_count \
            else np.zeros((0, 3), np.float32)
        i = np.ctypeslib.as_array(m.indices, (m.triangle_count * 3,)).reshape(-1, 3).copy() if m.triangle_count \
            else np.zeros((0, 3), np.uint32)
        f = np.ctypeslib.as_array(m.geometry_first_triangle, (m.geometry_count + 1,)).copy()
    finally:
        lib.rtx_obj_free(ctypes.byref(m))
    return v.astype(np.float32), i.astype(np.uint32), f.astype(np.uint32)


INSTANCE_DTYPE = np.dtype([("xf", np.float32, (12,)), ("cm", np.uint32), ("sf", np.uint32), ("addr", np.uint64)])
assert INSTANCE_DTYPE.itemsize == 64


def instance_records(scene, blas_addresses) -> np.ndarray:
    """64-B VkAccelerationStructureInstanceKHR records (main.cpp:848-858) for scene's instances."""
    n = len(scene.blas_of_instance)
    rec = np.zeros(n, INSTANCE_DTYPE)
    rec["xf"] = np.asarray(scene.transforms, np.float32).reshape(n, 12)
    rec["cm"] = (np.asarray(scene.custom_index, np.uint32) & 0xFFFFFF) | (np.asarray(scene.mask, np.uint32) << 24)
    rec["sf"] = (np.asarray(scene.sbt_offset, np.uint32) & 0xFFFFFF) | (np.asarray(scene.flags, np.uint32) << 24)
    rec["addr"] = np.asarray([blas_addresses[b] for b in scene.blas_of_instance], np.uint64)
    return rec


def hits_to_numpy(h):
    """Split a [N, 6] int32 hit tensor into (t, u, v float32 arrays, prim, geom, inst uint32 arrays)."""
    a = h.cpu().numpy()
    f = a[:, :3].view(np.float32)
    u = a[:, 3:].view(np.uint32)
    return f[:, 0], f[:, 1], f[:, 2], u[:, 0], u[:, 1], u[:, 2]
