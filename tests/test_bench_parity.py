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
