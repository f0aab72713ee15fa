"""FP64 CPU oracle for the ray-tracing hot path (ctypes wrapper over liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker.  The product library never imports it.  What each function restates is documented in rtx_oracle.cpp
with citations of vulkan-raytraced-triangle/main.cpp.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MISS = 0xFFFFFFFF
STATUS_MISS, STATUS_HIT, STATUS_TIE = 0, 1, 2


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.oracle_blas_create.restype = ctypes.c_int
        L.oracle_blas_create.argtypes = [ctypes.c_int, P, P, P, P, P]
        L.oracle_tlas_create.restype = ctypes.c_int
        L.oracle_tlas_create.argtypes = [ctypes.c_int, P]
        L.oracle_reset.restype = None
        L.oracle_blas_delta.restype = ctypes.c_double
        L.oracle_blas_delta.argtypes = [ctypes.c_int]
        L.oracle_trace.restype = None
        L.oracle_trace.argtypes = [ctypes.c_int, P, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, P,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, P, P, P, ctypes.c_int]
        L.oracle_trace_candidates.restype = None
        L.oracle_trace_candidates.argtypes = [ctypes.c_int, P, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int, P, P, P,
                                              P, P, ctypes.c_int]
        L.oracle_trace_brute.restype = None
        L.oracle_trace_brute.argtypes = [ctypes.c_int, P, ctypes.c_int64, ctypes.c_uint32, P, P, P]
        L.oracle_aspect.restype = None
        L.oracle_aspect.argtypes = [ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, P, P]
        L.oracle_raygen.restype = None
        L.oracle_raygen.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, P]
        L.oracle_shade.restype = None
        L.oracle_shade.argtypes = [ctypes.c_int, ctypes.c_int64, P, P, P, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint32, P, P]
        L.oracle_centroids.restype = None
        L.oracle_centroids.argtypes = [ctypes.c_int, P, P, P, P, P]
        L.oracle_morton.restype = None
        L.oracle_morton.argtypes = [P, ctypes.c_int64, ctypes.c_int, P]
        L.oracle_karras.restype = None
        L.oracle_karras.argtypes = [P, ctypes.c_int64, ctypes.c_int, P, P]
        L.oracle_treelet_dp.restype = None
        L.oracle_treelet_dp.argtypes = [P, P, P, P, P]
        L.oracle_treelet_refine.restype = None
        L.oracle_treelet_refine.argtypes = [ctypes.c_uint32, P, P, P, ctypes.c_uint32, P, P, P]
        L.oracle_pcg32.restype = None
        L.oracle_pcg32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P]
        L.oracle_pcg4d.restype = None
        L.oracle_pcg4d.argtypes = [P, ctypes.c_int64, P]
        L.oracle_nrc.restype = None
        L.oracle_nrc.argtypes = [P, P, ctypes.c_int, P, P, ctypes.c_int, P, P, P]
        L.oracle_adam.restype = None
        L.oracle_adam.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_int]
        L.oracle_render.restype = None
        L.oracle_render.argtypes = [ctypes.c_int, P, ctypes.c_uint32, P, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint32, P, ctypes.c_uint32, ctypes.c_float, P, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _geom_arrays(geoms):
    """geoms: list of (vertices float32 [V,3], indices uint32 [T,3], transform float32 [3,4] or None[, opaque]) or
    {"aabbs": float32 [N,6] boxes (min xyz, max xyz), "opaque": bool} for AABB geometry."""
    keep = []
    vp, ip, xp, pc, fl = [], [], [], [], []
    for g in geoms:
        if isinstance(g, dict):
            b = np.ascontiguousarray(g["aabbs"], dtype=np.float32).reshape(-1, 6)
            keep.append(b)
            vp.append(_ptr(b))
            ip.append(0)
            xp.append(0)
            pc.append(b.shape[0])
            fl.append((1 if g.get("opaque", True) else 0) | 0x100)
            continue
        v, i, x = g[:3]
        fl.append(1 if (len(g) < 4 or g[3]) else 0)
        v = np.ascontiguousarray(v, dtype=np.float32)
        i = np.ascontiguousarray(i, dtype=np.uint32)
        keep += [v, i]
        vp.append(_ptr(v))
        ip.append(_ptr(i))
        pc.append(i.shape[0] if i.ndim == 2 else i.size // 3)
        if x is None:
            xp.append(0)
        else:
            x = np.ascontiguousarray(x, dtype=np.float32).reshape(3, 4)
            keep.append(x)
            xp.append(_ptr(x))
    arrs = (np.array(vp, dtype=np.uint64), np.array(ip, dtype=np.uint64), np.array(pc, dtype=np.uint32),
            np.array(xp, dtype=np.uint64), np.array(fl, dtype=np.uint32))
    return keep, arrs


def reset() -> None:
    lib().oracle_reset()


def blas_create(geoms) -> int:
    keep, (vp, ip, pc, xp, fl) = _geom_arrays(geoms)
    return lib().oracle_blas_create(len(geoms), _ptr(vp), _ptr(ip), _ptr(pc), _ptr(xp), _ptr(fl))


def blas_delta(h: int) -> float:
    return lib().oracle_blas_delta(h)


def instance_records(transforms, blas_handles, custom_index=None, mask=None, sbt_offset=None, flags=None):
    """Pack 64-B instance records (main.cpp:848-858 layout).  blas_handles fill the address field."""
    n = len(blas_handles)
    rec = np.zeros(n, dtype=np.dtype([("xf", np.float32, (12,)), ("cm", np.uint32), ("sf", np.uint32),
                                      ("addr", np.uint64)]))
    rec["xf"] = np.asarray(transforms, dtype=np.float32).reshape(n, 12)
    ci = np.zeros(n, np.uint32) if custom_index is None else np.asarray(custom_index, np.uint32)
    mk = np.full(n, 0xFF, np.uint32) if mask is None else np.asarray(mask, np.uint32)
    so = np.zeros(n, np.uint32) if sbt_offset is None else np.asarray(sbt_offset, np.uint32)
    fl = np.zeros(n, np.uint32) if flags is None else np.asarray(flags, np.uint32)
    rec["cm"] = (ci & 0xFFFFFF) | (mk << 24)
    rec["sf"] = (so & 0xFFFFFF) | (fl << 24)
    rec["addr"] = np.asarray(blas_handles, dtype=np.uint64)
    return rec


def tlas_create(records: np.ndarray) -> int:
    records = np.ascontiguousarray(records)
    assert records.dtype.itemsize == 64
    return lib().oracle_tlas_create(len(records), _ptr(records))


def trace(tlas: int, rays: np.ndarray, any_hit: bool = False, cull_mask: int = 0xFF, threads: int = 0,
          ray_flags: int = 0, records=None, sbt_offset: int = 0, sbt_stride: int = 1):
    """rays: float32 [N,8].  Returns (t_uv float64 [N,3], ids uint32 [N,4] (prim, geom, inst, custom),
    status int32 [N]).  ray_flags: gl_RayFlags*EXT (facing and opacity culls, OPAQUE / NO_OPAQUE, SKIP_TRIANGLES).
    records: 48-B hit records whose any-hit programs run on non-opaque candidates (None: none run)."""
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    n = rays.shape[0]
    t_uv = np.zeros((n, 3), np.float64)
    ids = np.zeros((n, 4), np.uint32)
    st = np.zeros(n, np.int32)
    rec = None if records is None else np.ascontiguousarray(records)
    if rec is not None:
        assert rec.dtype.itemsize == 48
    lib().oracle_trace(tlas, _ptr(rays), n, int(any_hit), cull_mask, ray_flags, None if rec is None else _ptr(rec),
                       0 if rec is None else len(rec), sbt_offset, sbt_stride, _ptr(t_uv), _ptr(ids), _ptr(st),
                       threads)
    return t_uv, ids, st


def trace_candidates(tlas: int, rays: np.ndarray, cull_mask: int = 0xFF, max_c: int = 16, threads: int = 0):
    """Candidate sets of (epsilon-tie) rays: every triangle an FP32 traversal may report for each ray (FP64 inside,
    or within its FP32 uncertainty of an edge; t in the ray's window up to its uncertainty), ascending t.  Returns
    dict(count [n], ids [n, max_c, 3] (inst, geom, prim), t, tunc [n, max_c], inside, near [n, max_c] bool)."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    count = np.zeros(n, np.uint32)
    ids = np.zeros((n, max_c, 3), np.uint32)
    t = np.zeros((n, max_c), np.float64)
    tunc = np.zeros((n, max_c), np.float64)
    fl = np.zeros((n, max_c), np.uint32)
    lib().oracle_trace_candidates(tlas, _ptr(rays), n, cull_mask, max_c, _ptr(count), _ptr(ids), _ptr(t), _ptr(tunc),
                                  _ptr(fl), threads)
    return {"count": count, "ids": ids, "t": t, "tunc": tunc, "inside": (fl & 1) != 0, "near": (fl & 2) != 0,
            "max_c": max_c}


def check_ties(rays: np.ndarray, cand: dict, gpu_inst, gpu_geom, gpu_prim, gpu_t, occluded=None) -> dict:
    """What an FP32 traversal may answer on rays whose FP64 answer is not robust (epsilon-ties), given their candidate
    sets (trace_candidates):
      - closest hit: a hit must be a candidate that can be the closest one within the uncertainties (t - tunc at
        most the smallest t + tunc of the FP64-inside candidates in the window) with its t within tunc + 1e-5 t of
        the candidate's; a miss is allowed only when no candidate is robustly inside (inside, not near an edge, t
        clear of tmin / tmax by its uncertainty);
      - any hit (occluded given): 1 needs some candidate, 0 needs no robustly inside one.
    Returns counts: rays, hit_not_candidate, miss_not_allowed, t_bad, unchecked (candidate list truncated),
    occluded_bad."""
    rays = np.asarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    mc = cand["max_c"]
    cnt = cand["count"].astype(np.int64)
    valid = np.arange(mc)[None, :] < np.minimum(cnt, mc)[:, None]
    tmin = rays[:, 3].astype(np.float64)[:, None]
    tmax = rays[:, 7].astype(np.float64)[:, None]
    t, tu, ins, near = cand["t"], cand["tunc"], cand["inside"] & valid, cand["near"] & valid
    in_win = ins & (t >= tmin) & (t <= tmax)
    tstar = np.where(in_win, t + tu, np.inf).min(1)
    admissible = valid & (t - tu <= tstar[:, None])
    robust = ins & ~near & (t - tu >= tmin) & (t + tu <= tmax)
    must_hit = robust.any(1)
    truncated = cnt > mc
    out = {"rays": int(n), "unchecked": int(truncated.sum())}
    if occluded is not None:
        occ = np.asarray(occluded).astype(bool)
        bad = (occ & (cnt == 0)) | (~occ & must_hit)
        out["occluded_bad"] = int((bad & ~truncated).sum())
        return out
    miss = np.asarray(gpu_inst, np.uint32) == 0xFFFFFFFF
    same = admissible & (cand["ids"][:, :, 0] == np.asarray(gpu_inst, np.uint32)[:, None]) & \
        (cand["ids"][:, :, 1] == np.asarray(gpu_geom, np.uint32)[:, None]) & \
        (cand["ids"][:, :, 2] == np.asarray(gpu_prim, np.uint32)[:, None])
    tg = np.asarray(gpu_t, np.float64)[:, None]
    t_ok = same & (np.abs(tg - t) <= tu + 1e-5 * np.abs(t))
    out["miss_not_allowed"] = int((miss & must_hit & ~truncated).sum())
    out["hit_not_candidate"] = int((~miss & ~same.any(1) & ~truncated).sum())
    out["t_bad"] = int((~miss & same.any(1) & ~t_ok.any(1)).sum())
    return out


def trace_brute(tlas: int, rays: np.ndarray, cull_mask: int = 0xFF):
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    n = rays.shape[0]
    t_uv = np.zeros((n, 3), np.float64)
    ids = np.zeros((n, 4), np.uint32)
    st = np.zeros(n, np.int32)
    lib().oracle_trace_brute(tlas, _ptr(rays), n, cull_mask, _ptr(t_uv), _ptr(ids), _ptr(st))
    return t_uv, ids, st


def aspect(fov_deg: float, w: int, h: int):
    ax = ctypes.c_float()
    ay = ctypes.c_float()
    lib().oracle_aspect(fov_deg, w, h, ctypes.addressof(ax), ctypes.addressof(ay))
    return ax.value, ay.value


def raygen(camera, w: int, h: int, tmin: float = 0.0, tmax: float = 100.0) -> np.ndarray:
    cam = np.ascontiguousarray(camera, dtype=np.float32).reshape(4)
    rays = np.zeros((h * w, 8), np.float32)
    lib().oracle_raygen(_ptr(cam), w, h, tmin, tmax, _ptr(rays))
    return rays


def shade(tlas: int, t_uv, ids, records: np.ndarray, sbt_offset: int, sbt_stride: int, miss_color) -> np.ndarray:
    n = ids.shape[0]
    film = np.zeros((n, 4), np.float32)
    rec = np.ascontiguousarray(records)
    assert rec.dtype.itemsize == 48
    miss = np.ascontiguousarray(miss_color, dtype=np.float32)
    t_uv = np.ascontiguousarray(t_uv, dtype=np.float64)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    lib().oracle_shade(tlas, n, _ptr(t_uv), _ptr(ids), _ptr(rec), len(rec), sbt_offset, sbt_stride, _ptr(miss),
                       _ptr(film))
    return film


def trace_rays(tlas: int, records, camera, w, h, miss_color=(0.0, 0.0, 0.2), tmin=0.0, tmax=100.0, cull_mask=0xFF,
               sbt_offset=0, sbt_stride=1, threads=0, ray_flags=1):
    """Oracle of vkCmdTraceRaysKHR with the reference raygen/miss/chit (ray flags default gl_RayFlagsOpaqueEXT,
    main.cpp:1049): returns (film [H,W,4], status [H,W])."""
    rays = raygen(camera, w, h, tmin, tmax)
    t_uv, ids, st = trace(tlas, rays, False, cull_mask, threads, ray_flags, records, sbt_offset, sbt_stride)
    film = shade(tlas, t_uv, ids, records, sbt_offset, sbt_stride, miss_color)
    return film.reshape(h, w, 4), st.reshape(h, w)


def centroids(geoms) -> np.ndarray:
    keep, (vp, ip, pc, xp, _) = _geom_arrays(geoms)
    n = int(pc.sum())
    cen = np.zeros((n, 3), np.float32)
    lib().oracle_centroids(len(geoms), _ptr(vp), _ptr(ip), _ptr(pc), _ptr(xp), _ptr(cen))
    return cen


def morton(cen: np.ndarray, bits: int = 30) -> np.ndarray:
    cen = np.ascontiguousarray(cen, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(cen.shape[0], np.uint32 if bits == 30 else np.uint64)
    lib().oracle_morton(_ptr(cen), cen.shape[0], bits, _ptr(out))
    return out


def sort_pairs(keys: np.ndarray):
    """Stable key sort: returns (sorted keys, original indices)."""
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    return keys[order], order


def karras(sorted_keys: np.ndarray, key_bits: int):
    keys = np.ascontiguousarray(sorted_keys.astype(np.uint64))
    n = keys.size
    left = np.zeros(max(n - 1, 0), np.int32)
    right = np.zeros(max(n - 1, 0), np.int32)
    lib().oracle_karras(_ptr(keys), n, key_bits, _ptr(left), _ptr(right))
    return left, right


def treelet_dp(boxes: np.ndarray, leaf_cost: np.ndarray):
    """Subset DP of the treelet refinement over 7 leaves: boxes float32 [7, 6], leaf subtree costs [7].
    Returns (area[128], cost[128], part[128]) indexed by leaf subset bitmask."""
    b = np.ascontiguousarray(boxes, np.float32).reshape(7, 6)
    c = np.ascontiguousarray(leaf_cost, np.float32).reshape(7)
    area = np.zeros(128, np.float32)
    cost = np.zeros(128, np.float32)
    part = np.zeros(128, np.uint32)
    lib().oracle_treelet_dp(_ptr(b), _ptr(c), _ptr(area), _ptr(cost), _ptr(part))
    return area, cost, part


def treelet_refine(left: np.ndarray, right: np.ndarray, leaf_boxes: np.ndarray, rounds: int):
    """Refit + SAH treelet refinement of a binary radix tree (Karras child codes), leaf boxes in sorted order.
    Returns (left, right, cost, height, sah) with cost/height per internal node."""
    L = np.ascontiguousarray(left, np.int32).copy()
    R = np.ascontiguousarray(right, np.int32).copy()
    n = L.size + 1
    b = np.ascontiguousarray(leaf_boxes, np.float32).reshape(n, 6)
    cost = np.zeros(n - 1, np.float32)
    hgt = np.zeros(n - 1, np.uint32)
    sah = ctypes.c_double()
    lib().oracle_treelet_refine(n, _ptr(L), _ptr(R), _ptr(b), rounds, _ptr(cost), _ptr(hgt), ctypes.byref(sah))
    return L, R, cost, hgt, sah.value


def pcg32(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().oracle_pcg32(initstate, initseq, n, _ptr(out))
    return out


def pcg4d(v: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.uint32).reshape(-1, 4)
    out = np.zeros_like(v)
    lib().oracle_pcg4d(_ptr(v), v.shape[0], _ptr(out))
    return out


def render(tlas: int, records: np.ndarray, camera, lights: np.ndarray, w: int, h: int, spp: int,
           sample_offset: int = 0, max_depth: int = 8, rr_depth: int = 3, seed: int = 1234,
           ray_offset: float = 1e-4, threads: int = 0, accum: np.ndarray | None = None) -> np.ndarray:
    """FP64 restatement of rtx_cmd_render: returns the per-pixel radiance sums [h, w, 3] (float64)."""
    rec = np.ascontiguousarray(records)
    assert rec.dtype.itemsize == 48
    cam = np.ascontiguousarray(camera, dtype=np.float32).reshape(4)
    lt = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 12)
    if accum is None:
        accum = np.zeros((h, w, 3), np.float64)
    lib().oracle_render(tlas, _ptr(rec), len(rec), _ptr(cam), w, h, spp, sample_offset, max_depth, rr_depth, seed,
                        _ptr(lt), lt.shape[0], ray_offset, _ptr(accum), threads)
    return accum


NRC_PARAM_COUNT = 23888


def nrc(params: np.ndarray, samples: np.ndarray, lo, hi, bf16: bool = True, want_grad: bool = True):
    """FP64 cache MLP over samples (float32 [n, 16]: query[12], target[3], pad).  Returns (pred [n,3], grad
    [23888] or None, mean loss)."""
    p = np.ascontiguousarray(params, np.float32)
    smp = np.ascontiguousarray(samples, np.float32).reshape(-1, 16)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    n = smp.shape[0]
    out = np.zeros((n, 3), np.float64)
    grad = np.zeros(NRC_PARAM_COUNT, np.float64) if want_grad else None
    loss = np.zeros(1, np.float64)
    lib().oracle_nrc(_ptr(p), _ptr(smp), n, _ptr(lo), _ptr(hi), int(bf16), _ptr(out),
                     _ptr(grad) if want_grad else None, _ptr(loss))
    return out, grad, float(loss[0])


def adam(params, m, v, grad, lr, b1, b2, eps, step):
    """In-place Adam step on float32 arrays."""
    lib().oracle_adam(_ptr(params), _ptr(m), _ptr(v), _ptr(np.ascontiguousarray(grad, np.float64)), params.size,
                      lr, b1, b2, eps, step)


# ---- CPU baseline (libcpubase.so): the reference's ray tracing on CPU cores, FP32, timed by bench.py ----------------
_CB_PATH = os.path.join(_HERE, "libcpubase.so")
_cb = None


def cpubase():
    global _cb
    if _cb is None:
        if not os.path.exists(_CB_PATH):
            build()
        L = ctypes.CDLL(_CB_PATH)
        P, I, U32, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_int64
        L.cb_reset.restype = None
        L.cb_blas_create.restype = I
        L.cb_blas_create.argtypes = [I, P, P, P, P]
        L.cb_blas_create_many.restype = None
        L.cb_blas_create_many.argtypes = [I, P, P, P, P, P, P]
        L.cb_tlas_create.restype = I
        L.cb_tlas_create.argtypes = [I, P]
        L.cb_trace.restype = None
        L.cb_trace.argtypes = [I, P, I64, I, P, P, I]
        L.cb_trace_rays.restype = None
        L.cb_trace_rays.argtypes = [I, P, U32, U32, U32, P, P, U32, U32, ctypes.c_float, ctypes.c_float, P, I]
        L.cb_render.restype = ctypes.c_uint64
        L.cb_render.argtypes = [I, P, U32, P, U32, U32, U32, U32, U32, U32, U32, P, U32, ctypes.c_float, P, I]
        _cb = L
    return _cb


class CpuBaseline:
    """A scene in the CPU baseline (FP32 binned-SAH BVH2 per BLAS + one over the instances): the reference's ray
    tracing as a CPU driver would run it.  BLASes are built in parallel, one thread each."""

    def __init__(self, scene):
        L = cpubase()
        L.cb_reset()
        keep, vp, ip, pc, xp, first = [], [], [], [], [], [0]
        for geoms in scene.blas:
            for g in geoms:
                v = np.ascontiguousarray(g.vertices, np.float32)
                i = np.ascontiguousarray(g.indices, np.uint32)
                keep += [v, i]
                vp.append(v.ctypes.data)
                ip.append(i.ctypes.data)
                pc.append(i.shape[0])
                if g.transform is None:
                    xp.append(0)
                else:
                    x = np.ascontiguousarray(g.transform, np.float32).reshape(3, 4)
                    keep.append(x)
                    xp.append(x.ctypes.data)
            first.append(len(vp))
        arr = [np.array(a, t) for a, t in ((first, np.int32), (vp, np.uint64), (ip, np.uint64), (pc, np.uint32),
                                           (xp, np.uint64))]
        handles = np.zeros(len(scene.blas), np.int32)
        L.cb_blas_create_many(len(scene.blas), *[a.ctypes.data for a in arr], handles.ctypes.data)
        rec = instance_records(scene.transforms, [int(handles[b]) for b in scene.blas_of_instance],
                               scene.custom_index, scene.mask, scene.sbt_offset, scene.flags)
        self.tlas = L.cb_tlas_create(len(rec), rec.ctypes.data)
        self.L = L

    def trace(self, rays, any_hit=False, threads=0):
        """Closest hit: uint32 [n, 6] (t, u, v float bits, prim, geom, inst); any hit: uint32 [n] occluded."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        if any_hit:
            occ = np.zeros(n, np.uint32)
            self.L.cb_trace(self.tlas, rays.ctypes.data, n, 1, None, occ.ctypes.data, threads)
            return occ
        hits = np.zeros((n, 6), np.uint32)
        self.L.cb_trace(self.tlas, rays.ctypes.data, n, 0, hits.ctypes.data, None, threads)
        return hits

    def trace_rays(self, records, camera, w, h, miss_color=(0.0, 0.0, 0.2), tmin=0.0, tmax=100.0, sbt_offset=0,
                   sbt_stride=1, threads=0):
        rec = np.ascontiguousarray(records)
        film = np.zeros((h, w, 4), np.float32)
        miss = np.asarray(miss_color, np.float32)
        cam = np.asarray(camera, np.float32).reshape(4)
        self.L.cb_trace_rays(self.tlas, rec.ctypes.data, len(rec), sbt_offset, sbt_stride, miss.ctypes.data,
                             cam.ctypes.data, w, h, tmin, tmax, film.ctypes.data, threads)
        return film

    def render(self, records, camera, lights, w, h, spp, sample_offset=0, max_depth=8, rr_depth=3, seed=1234,
               ray_offset=1e-4, threads=0):
        """Radiance sums [h, w, 3] (float32) of spp samples per pixel, and the rays traced."""
        rec = np.ascontiguousarray(records)
        cam = np.asarray(camera, np.float32).reshape(4)
        lt = np.ascontiguousarray(lights, np.float32).reshape(-1, 12)
        acc = np.zeros((h, w, 3), np.float32)
        n = self.L.cb_render(self.tlas, rec.ctypes.data, len(rec), cam.ctypes.data, w, h, spp, sample_offset,
                             max_depth, rr_depth, seed, lt.ctypes.data, lt.shape[0], ray_offset, acc.ctypes.data,
                             threads)
        return acc, int(n)
