"""The bench's full-size parity count (bench.hit_parity) on synthetic hit records: planted ID, miss, t and u/v
errors are each counted once, epsilon-ties are excluded, and it agrees with the parity tests' own comparison."""
import os
import sys

import numpy as np

from helpers import compare_closest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
import bench  # noqa: E402


def _records(tuv, ids):
    h = np.zeros((len(tuv), 6), np.int32)
    h[:, :3] = tuv.astype(np.float32).view(np.int32)
    h[:, 3:] = ids.view(np.int32)
    return h


def _split(h):
    f = np.ascontiguousarray(h[:, :3]).view(np.float32)
    u = np.ascontiguousarray(h[:, 3:]).view(np.uint32)
    return f[:, 0], f[:, 1], f[:, 2], u[:, 0], u[:, 1], u[:, 2]


def test_hit_parity_counts_planted_errors():
    rng = np.random.default_rng(3)
    n = 4000
    st = rng.choice([0, 1, 2], size=n, p=[0.5, 0.45, 0.05]).astype(np.int32)
    tuv = np.stack([rng.uniform(1, 100, n), rng.uniform(0, 0.5, n), rng.uniform(0, 0.5, n)], axis=1)
    o_ids = rng.integers(0, 1000, (n, 4)).astype(np.uint32)
    ids = o_ids[:, :3].copy()
    ids[st == 0] = 0xFFFFFFFF
    h = _records(tuv, ids)
    clean = bench.hit_parity(h, tuv, o_ids, st)
    assert clean["rays"] == n and clean["ties"] == int((st == 2).sum()) and clean["robust_hits"] == int((st == 1).sum())
    assert clean["id_mismatch"] == clean["miss_mismatch"] == clean["t_over_tol"] == clean["uv_over_tol"] == 0
    assert clean["max_t_rel"] < 1e-7 and clean["max_uv_abs"] < 1e-7

    hits, misses, ties = np.nonzero(st == 1)[0], np.nonzero(st == 0)[0], np.nonzero(st == 2)[0]
    bad = ids.copy()
    bad[hits[0], 0] ^= 1  # wrong primitive
    bad[hits[1], 2] = 0xFFFFFFFF  # GPU miss where the oracle hits
    bad[misses[0], :] = (1, 2, 3)  # GPU hit where the oracle misses
    bad[ties[0], :] = (7, 7, 7)  # a tie: not counted
    tb = tuv.copy()
    tb[hits[2], 0] *= 1 + 1e-4  # t off by 1e-4 relative
    tb[hits[3], 2] += 1e-3  # v off by 1e-3
    hb = _records(tb, bad)
    rep = bench.hit_parity(hb, tuv, o_ids, st)
    assert rep["id_mismatch"] == 2 and rep["miss_mismatch"] == 1
    assert rep["t_over_tol"] == 1 and rep["uv_over_tol"] == 1

    ref = compare_closest(_split(hb), tuv, o_ids, st)
    assert (ref["id_mismatch"], ref["miss_mismatch"], ref["t_bad"], ref["uv_bad"]) == (
        rep["id_mismatch"], rep["miss_mismatch"], rep["t_over_tol"], rep["uv_over_tol"])


def test_tie_check_counts_planted_errors():
    """Epsilon-tie rays are checked against their candidate sets: the oracle's own answers pass, and a hit outside
    the set, a miss on a ray with a robustly inside candidate and a wrong t are each caught (by bench.hit_parity
    too)."""
    import oracle as O
    from helpers import oracle_scene, small_mixed_scene
    from pyrtx import scenes as S
    O.reset()
    sc = small_mixed_scene(seed=3, instances=24, freq=14, cells=20)
    t = oracle_scene(sc)
    rays = np.concatenate([S.camera_rays(sc.camera, 160, 120), S.incoherent_rays(sc, 40000, seed=3)])
    tuv, ids, st = O.trace(t, rays)
    tie = st == 2
    assert tie.sum() > 20
    r = rays[tie]
    c = O.trace_candidates(t, r)
    inst, geom, prim, tt = ids[tie, 2].copy(), ids[tie, 1].copy(), ids[tie, 0].copy(), tuv[tie, 0].copy()
    ok = O.check_ties(r, c, inst, geom, prim, tt)
    assert ok["hit_not_candidate"] == 0 and ok["miss_not_allowed"] == 0 and ok["t_bad"] == 0, ok
    hit = inst != 0xFFFFFFFF
    k = int(np.nonzero(hit)[0][0])
    bad = prim.copy()
    bad[k] += 100000
    assert O.check_ties(r, c, inst, geom, bad, tt)["hit_not_candidate"] == 1
    bt = tt.copy()
    bt[k] *= 1.01
    assert O.check_ties(r, c, inst, geom, prim, bt)["t_bad"] == 1
    # a ray whose candidate set has a robustly inside member must not miss
    valid = np.arange(c["max_c"])[None, :] < np.minimum(c["count"], c["max_c"])[:, None]
    robust = (c["inside"] & ~c["near"] & valid & (c["t"] - c["tunc"] >= r[:, 3:4]) &
              (c["t"] + c["tunc"] <= r[:, 7:8])).any(1)
    if robust.any():
        j = int(np.nonzero(robust)[0][0])
        mi = inst.copy()
        mi[j] = 0xFFFFFFFF
        assert O.check_ties(r, c, mi, geom, prim, tt)["miss_not_allowed"] == 1
    # the same through the bench's full-size parity count
    h = _records(tuv, ids[:, :3].copy())
    h[:, 5] = ids[:, 2].view(np.int32)
    rep = bench.hit_parity(h, tuv, ids, st, tlas=t, rays=rays)
    assert rep["ties_checked"]["hit_not_candidate"] == 0 and rep["ties_checked"]["miss_not_allowed"] == 0
    h2 = h.copy()
    h2[np.nonzero(tie)[0][k], 3] += 100000
    assert bench.hit_parity(h2, tuv, ids, st, tlas=t, rays=rays)["ties_checked"]["hit_not_candidate"] == 1
    O.reset()
