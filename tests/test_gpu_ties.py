"""GPU checks of epsilon-tie rays: rays whose FP64 answer an FP32 traversal cannot reproduce bit for bit are not
dropped but checked against their candidate sets (oracle.trace_candidates / oracle.check_ties).

Watertightness: rays aimed exactly at shared edges and shared vertices of closed or edge-sharing triangle sets (an
8-triangle fan, the reference scene's quad diagonals, main.cpp:676-682 indices (0, 1, 3), (1, 2, 3), and a closed
icosphere) must hit - never slip through a crack - report one of the triangles meeting there, with its t, and be
occluded for any-hit queries."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import oracle_scene
from pyrtx import scenes as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import pyrtx
    c = pyrtx.Context(0)
    yield c
    c.sync()
    c.close()


@pytest.fixture(autouse=True)
def _reset_oracle():
    O.reset()
    yield
    O.reset()


def _rays_to(origins, targets, tmax=1e30):
    o = np.asarray(origins, np.float64)
    d = np.asarray(targets, np.float64) - o
    rays = np.zeros((len(d), 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays[:, 7] = tmax
    return rays


def _world_vertices(sc, inst):
    """FP64 world positions of the vertices of instance `inst`'s geometries (geometry and instance transforms)."""
    out = []
    M = np.vstack([np.asarray(sc.transforms[inst], np.float64).reshape(3, 4), [0, 0, 0, 1]])
    for g in sc.blas[sc.blas_of_instance[inst]]:
        v = np.c_[np.asarray(g.vertices, np.float64), np.ones(len(g.vertices))]
        if g.transform is not None:
            v = v @ np.vstack([np.asarray(g.transform, np.float64).reshape(3, 4), [0, 0, 0, 1]]).T
        out.append((v @ M.T)[:, :3])
    return out


def _check_watertight(ctx, sc, rays, label):
    import pyrtx
    blases, tlas = ctx.upload_scene(sc)
    r = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    h = pyrtx.hits_to_numpy(ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_CLOSEST_HIT))
    occ = ctx.trace_ray_buffer(tlas, r, pyrtx.RTX_QUERY_ANY_HIT).cpu().numpy()
    ctx.sync()
    t, u, v, prim, geom, inst = h
    ot = oracle_scene(sc)
    cand = O.trace_candidates(ot, rays)
    rep = O.check_ties(rays, cand, inst, geom, prim, t)
    misses = int((inst == 0xFFFFFFFF).sum())
    _, _, st = O.trace(ot, rays)
    _say(f"  {label}: {len(rays)} rays, {int((st == 2).sum())} epsilon-ties, GPU misses {misses}, "
         f"not occluded {int((occ == 0).sum())}, {rep}")
    assert np.all(cand["count"] > 0)           # the oracle has candidates for every ray
    assert misses == 0 and np.all(occ == 1)     # no crack
    assert rep["hit_not_candidate"] == 0 and rep["t_bad"] == 0 and rep["unchecked"] == 0
    return st


def test_watertight_fan(ctx):
    _say("watertight: 8-triangle fan")
    k = 8
    ang = np.linspace(0, 2 * np.pi, k, endpoint=False)
    verts = np.concatenate([[[0, 0, 0]], np.stack([np.cos(ang), np.sin(ang), np.zeros(k)], 1)]).astype(np.float32)
    idx = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.uint32)
    sc = S.Scene(blas=[[S.Geometry(verts, idx)]], transforms=S.translation()[None], blas_of_instance=np.array([0]),
                 custom_index=np.zeros(1, np.uint32), mask=np.array([0xFF], np.uint32),
                 sbt_offset=np.zeros(1, np.uint32), flags=np.zeros(1, np.uint32))
    v = verts.astype(np.float64)
    # targets: the shared centre vertex, points along every spoke (shared edge), every outer vertex
    s = np.linspace(0.03, 0.97, 33)
    spokes = (v[1:][:, None, :] * s[None, :, None]).reshape(-1, 3)
    targets = np.concatenate([v[:1], spokes])
    rng = np.random.default_rng(4)
    origins = np.concatenate([np.array([[0.0, 0.0, 1.0]]).repeat(len(targets) // 2, 0),
                              rng.uniform(-2, 2, (len(targets) - len(targets) // 2, 3)) * [1, 1, 0] + [0, 0, 3]])
    rays = np.concatenate([_rays_to(origins, targets), _rays_to(origins[::-1] * [1, 1, -1], targets)])
    st = _check_watertight(ctx, sc, rays, "fan")
    assert (st == 2).mean() > 0.5  # they are edge cases for the oracle


def test_watertight_reference_quad_diagonals(ctx):
    _say("watertight: R0 quad diagonals")
    sc = S.scene_r0()
    targets = []
    a = np.linspace(0.01, 0.99, 99)
    for inst in range(2):
        for wv in _world_vertices(sc, inst):  # quads (0, 1, 3), (1, 2, 3): shared edge 1-3
            targets.append(wv[1][None, :] + a[:, None] * (wv[3] - wv[1])[None, :])
    targets = np.concatenate(targets)
    cam = np.array(sc.camera[:3], np.float64)
    rays = _rays_to(np.broadcast_to(cam, targets.shape), targets, tmax=100.0)
    _check_watertight(ctx, sc, rays, "R0 diagonals")


@pytest.mark.parametrize("freq", [8, 40])
def test_watertight_closed_icosphere(ctx, freq):
    """Rays from outside toward every vertex and edge midpoint of a closed, displaced icosphere (each shared by 2-6
    triangles) under a rotated, scaled instance."""
    _say(f"watertight: closed icosphere, frequency {freq}")
    g = S.icosphere(freq, radius=1.0, seed=5, amplitude=0.05)
    rng = np.random.default_rng(freq)
    R = S.random_rotation(rng, 1)[0]
    xf = np.zeros((1, 3, 4))
    xf[0, :, :3] = R * 1.7
    xf[0, :, 3] = (0.3, -2.0, 5.0)
    sc = S.Scene(blas=[[g]], transforms=xf.astype(np.float32), blas_of_instance=np.array([0]),
                 custom_index=np.zeros(1, np.uint32), mask=np.array([0xFF], np.uint32),
                 sbt_offset=np.zeros(1, np.uint32), flags=np.zeros(1, np.uint32))
    wv = _world_vertices(sc, 0)[0]
    tri = g.indices.astype(np.int64)
    edges = np.unique(np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), 1), axis=0)
    nv = min(len(wv), 4000)
    ne = min(len(edges), 6000)
    pick_v = rng.choice(len(wv), nv, replace=False)
    pick_e = edges[rng.choice(len(edges), ne, replace=False)]
    targets = np.concatenate([wv[pick_v], 0.5 * (wv[pick_e[:, 0]] + wv[pick_e[:, 1]])])
    centre = np.asarray(xf[0, :, 3], np.float64)
    out = targets - centre
    # origins outside the sphere, on the side of the target (the ray meets the surface there first)
    origins = centre + out * 3.0 + rng.normal(0, 0.3, out.shape)
    rays = _rays_to(origins, targets)
    _check_watertight(ctx, sc, rays, f"icosphere f{freq}")
