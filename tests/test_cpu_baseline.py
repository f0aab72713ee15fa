"""The CPU baseline (oracle/cpu_baseline.cpp, timed by bench.py beside the GPU) answers like the FP64 oracle: a
baseline that traced wrong rays would be no baseline.  Its FP32 ray transform is not rebased onto the BLAS box
the way the GPU's is, so a few robust oracle rays near edges may differ; the bar is 99.9% of them."""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_scene, small_mixed_scene
from pyrtx import scenes as S


@pytest.fixture(autouse=True)
def _reset():
    O.reset()
    yield
    O.reset()


def _agree(hits, tuv, ids, st):
    inst = hits[:, 5]
    robust_hit, robust_miss = st == 1, st == 0
    same = robust_hit & (inst == ids[:, 2]) & (hits[:, 4] == ids[:, 1]) & (hits[:, 3] == ids[:, 0])
    t = hits[:, 0].view(np.float32).astype(np.float64)
    t_ok = np.abs(t - tuv[:, 0]) <= 1e-4 * np.abs(tuv[:, 0])
    return (same.sum() / max(1, robust_hit.sum()), (robust_miss & (inst == 0xFFFFFFFF)).sum() / max(1, robust_miss.sum()),
            (same & t_ok).sum() / max(1, same.sum()))


@pytest.mark.parametrize("seed", [3, 11])
def test_cpu_baseline_rays_match_oracle(seed):
    sc = small_mixed_scene(seed=seed, instances=24, freq=14, cells=20)
    cb = O.CpuBaseline(sc)
    t = oracle_scene(sc)
    for rays in (S.camera_rays(sc.camera, 160, 120), S.incoherent_rays(sc, 20000, seed=seed)):
        h = cb.trace(rays)
        tuv, ids, st = O.trace(t, rays)
        hit_ok, miss_ok, t_ok = _agree(h, tuv, ids, st)
        assert (st == 1).sum() > 100
        assert hit_ok >= 0.999 and miss_ok >= 0.999 and t_ok >= 0.999, (hit_ok, miss_ok, t_ok)
        occ = cb.trace(rays, any_hit=True)
        _, _, sta = O.trace(t, rays, any_hit=True)
        assert (occ[sta == 1] == 1).mean() >= 0.999 and (occ[sta == 0] == 0).mean() >= 0.999


def test_cpu_baseline_reference_film():
    """raygen / closest-hit / miss of main.cpp:1033-1091 on the reference scene: the oracle's film off epsilon-ties."""
    sc = S.scene_r0()
    sc.width, sc.height = 300, 200
    cb = O.CpuBaseline(sc)
    f = cb.trace_rays(sc.records, sc.camera, 300, 200, sc.miss_color)
    t = oracle_scene(sc)
    of, st = O.trace_rays(t, sc.records, sc.camera, 300, 200, sc.miss_color)
    robust = st != 2
    assert (st == 1).sum() > 1000
    assert np.abs(f - of)[robust].max() < 1e-5
    assert np.all(f[..., 3] == 0.0)


def test_cpu_baseline_render_matches_oracle_estimator():
    """The FP32 CPU path tracer is the same estimator with the same random numbers as the oracle's FP64 one."""
    W, H, spp = 48, 27, 8
    sc, lights = S.cornell_scene(3, width=W, height=H)
    cb = O.CpuBaseline(sc)
    acc, n = cb.render(sc.records, sc.camera, lights, W, H, spp)
    t = oracle_scene(sc)
    o = O.render(t, sc.records, sc.camera, lights, W, H, spp) / spp
    g = acc.astype(np.float64) / spp
    err = float(np.mean((g - o) ** 2 / (o ** 2 + 0.01)))
    close = (np.abs(g - o) <= 1e-3 * np.maximum(1.0, np.abs(o))).all(-1).mean()
    assert n > W * H * spp * 2
    assert err < 1e-6 and close > 0.99, (err, close)
