"""Shared helpers for the parity tests: build the same scene in the oracle and compare results."""
from __future__ import annotations

import numpy as np

import oracle as O


def oracle_scene(scene):
    """Build scene in the FP64 oracle.  Returns the oracle TLAS handle."""
    from pyrtx.scenes import AabbGeometry

    def geom(g):
        if isinstance(g, AabbGeometry):
            return {"aabbs": g.boxes, "opaque": g.opaque}
        return (g.vertices, g.indices, g.transform, g.opaque)

    handles = [O.blas_create([geom(g) for g in geoms]) for geoms in scene.blas]
    rec = O.instance_records(scene.transforms, [handles[b] for b in scene.blas_of_instance], scene.custom_index,
                             scene.mask, scene.sbt_offset, scene.flags)
    return O.tlas_create(rec)


def compare_closest(gpu_hits, o_tuv, o_ids, o_st, rel_tol=1e-5, tlas=None, rays=None):
    """gpu_hits: (t, u, v, prim, geom, inst) numpy arrays.  Returns a dict of mismatch counts on robust rays; with
    the oracle TLAS and the rays, the epsilon-tie rays are checked against their candidate sets too (tie_* keys:
    a hit outside the set, a miss where a candidate is robustly inside, a t off its candidate's)."""
    t, u, v, prim, geom, inst = gpu_hits
    robust_hit = o_st == 1
    robust_miss = o_st == 0
    bad_miss = robust_miss & (inst != 0xFFFFFFFF)
    id_bad = robust_hit & ((inst != o_ids[:, 2]) | (geom != o_ids[:, 1]) | (prim != o_ids[:, 0]))
    ok = robust_hit & ~id_bad
    t_err = np.abs(t[ok].astype(np.float64) - o_tuv[ok, 0]) / np.maximum(np.abs(o_tuv[ok, 0]), 1e-30)
    u_err = np.abs(u[ok].astype(np.float64) - o_tuv[ok, 1])
    v_err = np.abs(v[ok].astype(np.float64) - o_tuv[ok, 2])
    tie_rep = {}
    if tlas is not None and (o_st == 2).any():
        tie = o_st == 2
        r = np.asarray(rays, np.float32).reshape(-1, 8)[tie]
        c = O.trace_candidates(tlas, r)
        tie_rep = {"tie_" + k: v for k, v in O.check_ties(r, c, inst[tie], geom[tie], prim[tie], t[tie]).items()}
    return {
        **tie_rep,
        "robust_hits": int(robust_hit.sum()),
        "robust_misses": int(robust_miss.sum()),
        "ties": int((o_st == 2).sum()),
        "miss_mismatch": int(bad_miss.sum()),
        "id_mismatch": int(id_bad.sum()),
        "t_bad": int((t_err > rel_tol).sum()),
        "uv_bad": int(((u_err > rel_tol) | (v_err > rel_tol)).sum()),
        "max_t_rel": float(t_err.max()) if t_err.size else 0.0,
        "max_uv_abs": float(max(u_err.max(), v_err.max())) if u_err.size else 0.0,
    }


def assert_ties_ok(rep):
    """No epsilon-tie ray answered outside its candidate set (when compare_closest checked them)."""
    for k in ("tie_hit_not_candidate", "tie_miss_not_allowed", "tie_t_bad"):
        assert rep.get(k, 0) == 0, rep


def small_mixed_scene(seed=7, instances=24, freq=12, cells=16):
    """A few rotated/scaled instances of a displaced icosphere and a terrain (oracle-friendly sizes)."""
    from pyrtx import scenes as S
    rng = np.random.default_rng(seed)
    sph = S.icosphere(freq, radius=1.0, seed=seed, amplitude=0.08)
    ter = S.terrain(cells, size=2.0, height=0.3, seed=seed + 1)
    R = S.random_rotation(rng, instances)
    sc = rng.uniform(0.5, 2.0, instances)
    pos = rng.uniform(-6, 6, (instances, 3))
    xf = np.zeros((instances, 3, 4))
    xf[:, :, :3] = R * sc[:, None, None]
    xf[:, :, 3] = pos
    return S.Scene(blas=[[sph], [ter]], transforms=xf.astype(np.float32),
                   blas_of_instance=np.arange(instances) % 2, custom_index=np.arange(instances, dtype=np.uint32) + 7,
                   mask=np.full(instances, 0xFF, np.uint32), sbt_offset=np.zeros(instances, np.uint32),
                   flags=np.zeros(instances, np.uint32), camera=(0.0, 0.0, 20.0, 60.0), width=96, height=64,
                   meta={"name": "mixed"})
