"""CPU tests that pin the FP64 oracle before it is trusted as the checker.

Each test checks the oracle against something it does not share code with: the reference scene's closed-form
image (vulkan-raytraced-triangle/main.cpp constants), textbook definitions (bit interleaving, binary radix
trees), brute force, published known answers (PCG32), and invariants."""
import math

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


def r0_closed_form(W=1200, H=800):
    """The R0 picture derived by hand: every hit has t = 10 (origin z = 10, dir.z = -1); hit point
    (X, Y) = (10 ax ndc.x, -10 ay ndc.y); quads and colours from main.cpp:676-695, 835-858, 1080-1091, 1307-1317."""
    ay = math.tan(math.radians(60.0) / 2)
    ax = ay * W / H
    x = np.arange(W)
    y = np.arange(H)
    nx = (x + 0.5) / W * 2 - 1
    ny = (y + 0.5) / H * 2 - 1
    X = np.broadcast_to(10 * ax * nx[None, :], (H, W))
    Y = np.broadcast_to(-10 * ay * ny[:, None], (H, W))
    img = np.zeros((H, W, 4))
    img[..., 2] = 0.2
    cols = {(0, 0): (0.6, 0.1, 0.2), (0, 1): (0.1, 0.8, 0.4), (1, 0): (0.9, 0.7, 0.1), (1, 1): (0.3, 0.6, 0.9)}
    for inst, (y0, y1) in enumerate([(1, 3), (-3, -1)]):
        for g, (x0, x1) in enumerate([(-3, -1), (1, 3)]):
            m = (X >= x0) & (X <= x1) & (Y >= y0) & (Y <= y1)
            img[m, :3] = cols[(inst, g)]
            if (inst, g) == (1, 1):
                px, py = X - 2, Y + 2
                m2 = m & (px + py > 0)
                img[m2, 0] = ((1 - py) / 2)[m2]
                img[m2, 1] = ((px + py) / 2)[m2]
                img[m2, 2] = ((1 - px) / 2)[m2]
    # distance (in world units) to the nearest quad edge or tri-1 diagonal, to exclude edge pixels
    d = np.full((H, W), np.inf)
    for cx in (-2, 2):
        for cy in (2, -2):
            lx, ly = X - cx, Y - cy
            inside = (np.abs(lx) <= 1) & (np.abs(ly) <= 1)
            edge = np.minimum(np.abs(np.abs(lx) - 1), np.abs(np.abs(ly) - 1))
            diag = np.abs(lx + ly) / math.sqrt(2)
            d = np.where(inside | (edge < 1e-3), np.minimum(d, np.minimum(edge, diag)), d)
    return img, d


def test_r0_matches_closed_form():
    sc = S.scene_r0()
    t = oracle_scene(sc)
    film, st = O.trace_rays(t, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    img, d = r0_closed_form()
    assert film.shape == (800, 1200, 4)
    assert np.all(film[..., 3] == 0.0)  # imageStore(vec4(hit, 0.0)), main.cpp:1054
    clear = d > 1e-4
    assert np.abs(film[clear] - img[clear]).max() < 2e-6
    # every hit is at t = 10 exactly
    rays = O.raygen(sc.camera, sc.width, sc.height)
    tuv, ids, st2 = O.trace(t, rays)
    assert np.all(tuv[st2 == 1, 0] == 10.0)
    # ties only near edges/diagonals
    assert np.all(d.ravel()[st2 == 2] < 1e-3)
    # hit counts: 4 quads of 2x2 world units
    assert (st2 == 1).sum() > 70000


def test_c1_literal_all_miss_and_mirror_closed_form():
    sc = S.scene_c1(-2.5)
    t = oracle_scene(sc)
    film, st = O.trace_rays(t, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    assert np.all(st == 0)
    assert np.all(film[..., :3] == np.array([0, 0, 0.2], np.float32))
    O.reset()
    sc = S.scene_c1(2.5)
    t = oracle_scene(sc)
    film, st = O.trace_rays(t, sc.records, sc.camera, sc.width, sc.height, sc.miss_color)
    W, H = sc.width, sc.height
    ay = math.tan(math.radians(30.0))
    ax = ay * W / H
    x = np.arange(W)
    y = np.arange(H)
    X = np.broadcast_to(2.5 * ax * ((x + 0.5) / W * 2 - 1)[None, :], (H, W))
    Y = np.broadcast_to(-2.5 * ay * ((y + 0.5) / H * 2 - 1)[:, None], (H, W))
    v0, v1, v2 = np.array([1, 1.0]), np.array([-1, 1.0]), np.array([0, -1.0])
    den = (v1[1] - v2[1]) * (v0[0] - v2[0]) + (v2[0] - v1[0]) * (v0[1] - v2[1])
    b0 = ((v1[1] - v2[1]) * (X - v2[0]) + (v2[0] - v1[0]) * (Y - v2[1])) / den
    b1 = ((v2[1] - v0[1]) * (X - v2[0]) + (v0[0] - v2[0]) * (Y - v2[1])) / den
    b2 = 1 - b0 - b1
    inside = (b0 > 1e-4) & (b1 > 1e-4) & (b2 > 1e-4)
    outside = (b0 < -1e-4) | (b1 < -1e-4) | (b2 < -1e-4)
    assert np.all(st[inside] == 1) and np.all(st[outside] == 0)
    exp = np.stack([b0, b1, b2], -1)
    assert np.abs(film[inside][:, :3] - exp[inside]).max() < 2e-6
    rays = O.raygen(sc.camera, W, H)
    tuv, ids, st2 = O.trace(t, rays)
    assert np.allclose(tuv[st2 == 1, 0], 2.5, rtol=0, atol=1e-12)


def test_bvh_equals_brute_force():
    sc = small_mixed_scene(seed=3, instances=10, freq=6, cells=8)
    t = oracle_scene(sc)
    rng = np.random.default_rng(0)
    rays = S.camera_rays(sc.camera, 64, 48)
    inc = S.incoherent_rays(sc, 3000, seed=5)
    for rr in (rays, inc):
        a = O.trace(t, rr)
        b = O.trace_brute(t, rr)
        hit = b[2] == 1
        assert np.array_equal(a[1][hit], b[1][hit])
        assert np.array_equal(a[2] == 0, b[2] == 0)
        assert np.abs(a[0][hit] - b[0][hit]).max() == 0.0
    del rng


def test_known_answers_axis_aligned():
    verts = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    b = O.blas_create([(verts, idx, None)])
    t = O.tlas_create(O.instance_records(S.translation(z=5.0)[None], [b]))
    # ray from (0.5, 0.25, 10) straight down -z hits at z=5 -> t = 5, P local (0.5, 0.25): u = 0.25, v = 0.125
    rays = np.array([[0.5, 0.25, 10, 0, 0, 0, -1, 100], [0.5, 0.25, 10, 0, 0, 0, -2, 100],
                     [3, 3, 10, 0, 0, 0, -1, 100], [0.5, 0.25, 10, 0, 0, 0, -1, 4.9]], np.float32)
    tuv, ids, st = O.trace(t, rays)
    assert list(st) == [1, 1, 0, 0]
    assert tuv[0, 0] == 5.0 and abs(tuv[0, 1] - 0.25) < 1e-15 and abs(tuv[0, 2] - 0.125) < 1e-15
    assert tuv[1, 0] == 2.5  # unnormalised direction: t is parametric
    assert ids[0, 2] == 0 and ids[2, 2] == O.MISS


def degenerate_geometry():
    """Five triangles side by side at z = 0 along x (triangle p covers x in [3p, 3p + 2], y in [0, 2]): 0 and 4 are
    ordinary, 1 has a NaN X component (inactive under Vulkan's rule), 2 an infinite Y component, 3 is collinear
    (zero area)."""
    v = []
    for p in range(5):
        x0 = 3.0 * p
        v += [[x0, 0, 0], [x0 + 2, 0, 0], [x0, 2, 0]]
    v = np.array(v, np.float32)
    v[3 * 1 + 1, 0] = np.nan
    v[3 * 2 + 2, 1] = np.inf
    v[3 * 3 + 2] = [9.0 + 1.0, 0.0, 0.0]  # on the edge v0 v1: collinear
    return v, np.arange(15, dtype=np.uint32).reshape(5, 3)


def test_inactive_and_degenerate_triangles_known_answers():
    """Triangles with a NaN or infinite vertex component are inactive (never hit) and keep their primitive index;
    a zero-area triangle is never hit; rays with an empty [tmin, tmax] miss; axis-parallel rays hit exactly."""
    v, idx = degenerate_geometry()
    b = O.blas_create([(v, idx, None)])
    t = O.tlas_create(O.instance_records(S.translation()[None], [b]))
    # straight down -z through (3p + 0.5, 0.5): hits triangle p at t = 1 where it is active and not degenerate
    rays = np.array([[3 * p + 0.5, 0.5, 1, 0, 0, 0, -1, 10] for p in range(5)] +
                    [[10.0, 0.0, 1, 0, 0, 0, -1, 10],        # onto the degenerate triangle's segment
                     [0.5, 0.5, 1, 2.0, 0, 0, -1, 1.5],      # tmin > tmax
                     [0.5, 0.5, 1, 1.0, 0, 0, -1, 1.0],      # tmin = tmax = the hit distance
                     [0.5, 0.5, 0.0, 0, 0, 0, 1, 10],        # origin on the triangle's plane, going up: t = 0
                     [-1.0, 0.5, 0.0, 0, 1, 0, 0, 100]],     # in the triangle's plane (parallel): no hit
                    np.float32)
    tuv, ids, st = O.trace(t, rays)
    # status 1 robust hit, 0 robust miss, 2 epsilon-tie: a hit exactly at tmin or tmax (rays 7 and 8) is a tie,
    # which an FP32 traversal may report either way
    assert list(st) == [1, 0, 0, 0, 1, 0, 0, 2, 2, 0], list(st)
    assert ids[0, 0] == 0 and ids[4, 0] == 4  # primitive indices keep the inactive triangles' positions
    assert tuv[0, 0] == 1.0 and tuv[4, 0] == 1.0 and tuv[7, 0] == 1.0 and tuv[8, 0] == 0.0


def cutout_scene(front_opaque=False, inst_flags=0, n=4):
    """A right triangle v0 (0,0,0), v1 (1,0,0), v2 (0,1,0) at z = 0, so that a ray down -z through (x, y) meets it
    at barycentrics (u, v) = (x, y) exactly, plus an opaque backstop triangle at z = -1 (geometry 1).  Hit records:
    geometry 0 carries the cutout any-hit program with n cells per edge."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    back = np.array([[-5, -5, -1], [10, -5, -1], [-5, 10, -1]], np.float32)  # one triangle: no edge over [0,1]^2
    b = O.blas_create([(tri, np.array([[0, 1, 2]], np.uint32), None, front_opaque),
                       (back, np.array([[0, 1, 2]], np.uint32), None, True)])
    t = O.tlas_create(O.instance_records(S.translation()[None], [b], flags=[inst_flags]))
    rec = S.hit_records([0, 0], [(1, 0, 0), (0, 1, 0)], any_hit=[S.ANY_HIT_CUTOUT | (n << 8), 0])
    return t, rec


def grid_rays(k=64):
    # cell-centred sample points, never on a cell boundary of n = 4 (multiples of 1/4) nor on the diagonal
    x = (np.arange(k) + 0.37) / k
    X, Y = np.meshgrid(x, x)
    rays = np.zeros((k * k, 8), np.float32)
    rays[:, 0] = X.ravel()
    rays[:, 1] = Y.ravel()
    rays[:, 2] = 1
    rays[:, 6] = -1
    rays[:, 7] = 10
    return rays


def cutout_expected(rays, n=4, cut=True):
    """Closed form: geometry 0 when inside the triangle and (not cut or the checker cell is kept), else 1."""
    x, y = rays[:, 0].astype(np.float64), rays[:, 1].astype(np.float64)
    inside = (x > 0) & (y > 0) & (x + y < 1)
    kept = ((np.floor(n * x) + np.floor(n * y)) % 2) == 0
    return np.where(inside & (kept | (not cut)), 0, 1)


def test_any_hit_cutout_and_opacity_closed_form():
    rays = grid_rays()
    x, y = rays[:, 0].astype(np.float64), rays[:, 1].astype(np.float64)
    inside = (x > 0) & (y > 0) & (x + y < 1)
    away = np.abs(x + y - 1) > 1e-3  # the diagonal edge is a tie band
    # non-opaque triangle with the cutout program: holes show the backstop
    t, rec = cutout_scene()
    tuv, ids, st = O.trace(t, rays, records=rec)
    assert np.all(st[away] == 1)
    assert np.array_equal(ids[away, 1], cutout_expected(rays)[away])
    assert np.all(tuv[ids[:, 1] == 1, 0] == 2.0) and np.all(tuv[(ids[:, 1] == 0) & away, 0] == 1.0)
    # (u, v) of a kept candidate are the hit point's x, y
    k = away & (ids[:, 1] == 0)
    assert np.abs(tuv[k, 1] - x[k]).max() < 1e-12 and np.abs(tuv[k, 2] - y[k]).max() < 1e-12
    # without records no any-hit program runs (a hit group without an any-hit shader accepts)
    _, ids0, _ = O.trace(t, rays)
    assert np.array_equal(ids0[away, 1], cutout_expected(rays, cut=False)[away])
    # ray OPAQUE overrides the geometry: no program
    _, ids1, _ = O.trace(t, rays, records=rec, ray_flags=0x1)
    assert np.array_equal(ids1[away, 1], cutout_expected(rays, cut=False)[away])
    # CULL_NO_OPAQUE drops the triangle everywhere; CULL_OPAQUE drops the backstop
    _, ids2, st2 = O.trace(t, rays, records=rec, ray_flags=0x80)
    assert np.all(ids2[:, 1] == 1) and np.all(st2 == 1)
    _, ids3, st3 = O.trace(t, rays, records=rec, ray_flags=0x40)
    exp3 = cutout_expected(rays)
    assert np.array_equal(st3[away] == 1, exp3[away] == 0)
    # any-hit queries see only accepted candidates: a hole in front of nothing is a miss
    _, _, st4 = O.trace(t, rays, any_hit=True, records=rec, ray_flags=0x40)
    assert np.array_equal(st4[away] == 1, exp3[away] == 0)
    # the instance's FORCE_OPAQUE wins over the geometry, the ray's NO_OPAQUE over the instance
    t5, rec5 = cutout_scene(inst_flags=0x4)
    _, ids5, _ = O.trace(t5, rays, records=rec5)
    assert np.array_equal(ids5[away, 1], cutout_expected(rays, cut=False)[away])
    _, ids6, _ = O.trace(t5, rays, records=rec5, ray_flags=0x2)
    assert np.array_equal(ids6[away, 1], cutout_expected(rays)[away])
    # an opaque geometry made non-opaque by the instance's FORCE_NO_OPAQUE runs its program
    t7, rec7 = cutout_scene(front_opaque=True, inst_flags=0x8)
    _, ids7, _ = O.trace(t7, rays, records=rec7)
    assert np.array_equal(ids7[away, 1], cutout_expected(rays)[away])
    assert inside.sum() > 1500


def test_any_hit_cutout_boundaries_are_ties():
    # rays exactly on the cell boundaries u = 1/4, 1/2 (x) of a n = 4 cutout are epsilon-ties, never robust
    t, rec = cutout_scene()
    ys = (np.arange(40) + 0.5) / 160
    rays = np.zeros((3 * 40, 8), np.float32)
    for i, xb in enumerate([0.25, 0.5, 0.125 + 1e-9]):
        rays[40 * i:40 * (i + 1), 0] = xb
        rays[40 * i:40 * (i + 1), 1] = ys
    rays[:, 2] = 1
    rays[:, 6] = -1
    rays[:, 7] = 10
    _, _, st = O.trace(t, rays, records=rec)
    assert np.all(st[:80] == 2)
    assert np.all(st[80:] == 1)  # mid-cell: robust


def _proc_tlas(boxes, opaque=True, tri=None):
    geoms = [{"aabbs": np.asarray(boxes, np.float32), "opaque": opaque}]
    if tri is not None:
        geoms.append(tri)
    b = O.blas_create(geoms)
    return O.tlas_create(O.instance_records(S.translation()[None], [b]))


def _down_rays(x, y, z0=10.0, tmin=0.0):
    rays = np.zeros((len(x), 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2], rays[:, 3] = x, y, z0, tmin
    rays[:, 6] = -1
    rays[:, 7] = 100
    return rays


def test_intersection_programs_closed_form():
    rng = np.random.default_rng(3)
    x = rng.uniform(-1.4, 1.4, 3000).astype(np.float32)
    y = rng.uniform(-1.4, 1.4, 3000).astype(np.float32)
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    rays = _down_rays(x, y)
    r2 = xd ** 2 + yd ** 2
    # sphere program: the unit sphere inscribed in [-1, 1]^3; t = 10 - sqrt(1 - x^2 - y^2), (u, v) = ((x+1)/2, (y+1)/2)
    t = _proc_tlas([[-1, -1, -1, 1, 1, 1]])
    rec = S.hit_records([S.hit_group(0, S.INTERSECTION_SPHERE)], [(1, 0, 0)])
    tuv, ids, st = O.trace(t, rays, records=rec)
    inside, outside = r2 < 1 - 1e-3, r2 > 1 + 1e-3
    assert np.all(st[inside] == 1) and np.all(st[outside] == 0)
    assert np.abs(tuv[inside, 0] - (10 - np.sqrt(1 - r2[inside]))).max() < 1e-12
    assert np.abs(tuv[inside, 1] - (xd[inside] + 1) / 2).max() < 1e-12
    assert np.abs(tuv[inside, 2] - (yd[inside] + 1) / 2).max() < 1e-12
    # grazing rays (|r| = 1) are ties, never robust
    g = _down_rays(np.float32([1.0, 0.0]), np.float32([0.0, -1.0]))
    assert np.all(O.trace(t, g, records=rec)[2] == 2)
    # from the centre the sphere is hit where the ray leaves it (t = 1), not at the entry behind the origin
    tuv, _, st = O.trace(t, _down_rays(np.float32([0.2]), np.float32([0.1]), z0=0.0), records=rec)
    x0, y0 = float(np.float32(0.2)), float(np.float32(0.1))
    assert st[0] == 1 and abs(tuv[0, 0] - np.sqrt(1 - x0 * x0 - y0 * y0)) < 1e-12
    # box program (the default without records): box [0, 2] x [0, 1] x [-1, 0], top face z = 0 at t = 10,
    # (u, v) along x, y (cyclic after z); from inside, the bottom face
    t = _proc_tlas([[0, 0, -1, 2, 1, 0]])
    xb, yb = rng.uniform(0.01, 1.99, 500).astype(np.float32), rng.uniform(0.01, 0.99, 500).astype(np.float32)
    tuv, ids, st = O.trace(t, _down_rays(xb, yb))
    assert np.all(st == 1) and np.all(tuv[:, 0] == 10.0)
    assert np.abs(tuv[:, 1] - xb / 2.0).max() < 1e-12 and np.abs(tuv[:, 2] - yb.astype(np.float64)).max() < 1e-12
    tuv, _, st = O.trace(t, _down_rays(np.float32([0.5]), np.float32([0.5]), z0=-0.25))
    assert st[0] == 1 and tuv[0, 0] == 0.75
    assert np.all(O.trace(t, _down_rays(np.float32([2.5]), np.float32([0.5])))[2] == 0)
    # an edge of the hit face is a tie; SKIP_AABBS removes the box, SKIP_TRIANGLES does not
    assert O.trace(t, _down_rays(np.float32([0.0]), np.float32([0.5])))[2][0] == 2
    assert np.all(O.trace(t, _down_rays(xb, yb), ray_flags=0x200)[2] == 0)
    assert np.all(O.trace(t, _down_rays(xb, yb), ray_flags=0x100)[2] == 1)
    # inactive boxes (min > max, NaN) are not primitives
    t = _proc_tlas([[0, 0, -1, -2, 1, 0], [np.nan, 0, 0, 1, 1, 1]])
    assert np.all(O.trace(t, _down_rays(xb, yb))[2] == 0)


def test_watertight_fan_and_ties():
    # 8-triangle fan around the origin: interior rays hit exactly one triangle, rays through the shared
    # vertex or along shared edges are epsilon-ties, never misses.
    k = 8
    ang = np.linspace(0, 2 * np.pi, k, endpoint=False)
    verts = np.concatenate([[[0, 0, 0]], np.stack([np.cos(ang), np.sin(ang), np.zeros(k)], 1)]).astype(np.float32)
    idx = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.uint32)
    b = O.blas_create([(verts, idx, None)])
    t = O.tlas_create(O.instance_records(S.translation()[None], [b]))
    rng = np.random.default_rng(1)
    r = rng.uniform(0.05, 0.85, 2000)
    a = rng.uniform(0, 2 * np.pi, 2000)
    rays = np.zeros((2000, 8), np.float32)
    rays[:, 0] = r * np.cos(a)
    rays[:, 1] = r * np.sin(a)
    rays[:, 2] = 1
    rays[:, 6] = -1
    rays[:, 7] = 10
    tuv, ids, st = O.trace(t, rays)
    assert np.all(st != 0)
    sector = (np.floor(np.mod(a, 2 * np.pi) / (2 * np.pi / k)).astype(int)) % k
    ok = st == 1
    assert np.all(ids[ok, 0] == sector[ok])
    # through the shared vertex and exactly along edges
    edge = np.zeros((k + 1, 8), np.float32)
    edge[:k, 0] = 0.5 * np.cos(ang)
    edge[:k, 1] = 0.5 * np.sin(ang)
    edge[:, 2] = 1
    edge[:, 6] = -1
    edge[:, 7] = 10
    _, _, st = O.trace(t, edge)
    assert np.all(st == 2)


def _interleave(x, y, z, bits):
    code = 0
    for i in range(bits):
        code |= ((x >> i) & 1) << (3 * i + 2)
        code |= ((y >> i) & 1) << (3 * i + 1)
        code |= ((z >> i) & 1) << (3 * i)
    return code


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_textbook(bits):
    rng = np.random.default_rng(bits)
    n = 4000
    cen = rng.uniform(-3, 5, (n, 3)).astype(np.float32)
    cen[:50] = cen[50]  # duplicates
    codes = O.morton(cen, bits)
    k = bits // 3
    cells = np.float32(2 ** k)
    cmin = cen.min(0)
    cmax = cen.max(0)
    scale = (cells / (cmax - cmin)).astype(np.float32)
    f = ((cen - cmin).astype(np.float32) * scale).astype(np.float32)
    q = np.where(f > 0, np.minimum(f, np.float32(2 ** k - 1)), 0).astype(np.uint64)
    for i in range(0, n, 7):
        assert int(codes[i]) == _interleave(int(q[i, 0]), int(q[i, 1]), int(q[i, 2]), k)
    # exhaustive per-axis spread of 10-bit values
    if bits == 30:
        for axis in range(3):
            c = np.zeros((1024, 3), np.float32)
            c[:, axis] = np.arange(1024, dtype=np.float32) + 0.5
            c[0, :] = 0.0
            c[1023, :] = 1024.0
            codes = O.morton(c, 30)
            vals = np.arange(1024)
            for v in vals[1:-1]:
                qq = [0, 0, 0]
                qq[axis] = int(v)
                assert int(codes[v]) == _interleave(qq[0], qq[1], qq[2], 10)


def _radix_tree(keys, bits):
    """Textbook binary radix tree (top-down split at the first differing bit, index tie-break), labelled as in
    Karras 2012: a range [a, b] is internal node a if it is a right child, b if it is a left child; root = 0."""
    n = len(keys)
    ext = [(int(k) << 32) | i for i, k in enumerate(keys)]
    left = np.zeros(n - 1, np.int64)
    right = np.zeros(n - 1, np.int64)
    stack = [(0, n - 1, 0)]
    while stack:
        a, b, label = stack.pop()
        x = ext[a] ^ ext[b]
        hb = x.bit_length() - 1
        # split: last position whose bit hb equals ext[a]'s
        lo, hi = a, b
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if (ext[mid] >> hb) == (ext[a] >> hb):
                lo = mid
            else:
                hi = mid
        g = lo
        left[label] = g if g > a else -1 - g
        right[label] = g + 1 if g + 1 < b else -1 - (g + 1)
        if g > a:
            stack.append((a, g, g))
        if g + 1 < b:
            stack.append((g + 1, b, g + 1))
    return left, right


@pytest.mark.parametrize("bits,n", [(30, 1000), (30, 2), (63, 777)])
def test_karras_equals_textbook_radix_tree(bits, n):
    rng = np.random.default_rng(n)
    kb = 32 if bits == 30 else 64
    hi = 2 ** 30 if bits == 30 else 2 ** 62
    keys = np.sort(rng.integers(0, hi, n, dtype=np.uint64)) if n > 2 else np.array([5, 5], np.uint64)
    if n > 10:
        keys[10:30] = keys[10]  # duplicate keys exercise the index tie-break
        keys = np.sort(keys)
    left, right = O.karras(keys, kb)
    el, er = _radix_tree(keys, kb)
    assert np.array_equal(left, el) and np.array_equal(right, er)
    # invariants: every leaf once, every internal node except the root once
    kids = np.concatenate([left, right])
    leaves = -1 - kids[kids < 0]
    internal = kids[kids >= 0]
    assert np.array_equal(np.sort(leaves), np.arange(n))
    assert np.array_equal(np.sort(internal), np.arange(1, n - 1))


def test_pcg32_known_answers():
    # pcg32-demo (O'Neill, pcg-c-basic): seed 42, sequence 54
    got = O.pcg32(42, 54, 6)
    assert [hex(int(x)) for x in got] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                          "0xcbed606e"]


def test_pcg4d_uniformity():
    v = np.stack(np.meshgrid(np.arange(64), np.arange(64), np.arange(8), np.arange(2), indexing="ij"), -1)
    out = O.pcg4d(v.reshape(-1, 4).astype(np.uint32))
    assert out.shape == (64 * 64 * 8 * 2, 4)
    for c in range(4):
        h = np.bincount((out[:, c] >> 28).astype(np.int64), minlength=16)
        e = out.shape[0] / 16
        chi2 = ((h - e) ** 2 / e).sum()
        assert chi2 < 50  # 15 dof
    assert len(np.unique(out[:, 0])) > 0.99 * out.shape[0]


def test_tie_rule_is_rare_on_query_scene_sample():
    sc = S.query_scene(scale_tris=0.002, instances=64)
    t = oracle_scene(sc)
    rays = S.camera_rays(sc.camera, 96, 96)
    tuv, ids, st = O.trace(t, rays)
    assert (st == 2).mean() < 0.02
    inc = S.incoherent_rays(sc, 4000)
    tuv, ids, st = O.trace(t, inc)
    assert (st == 2).mean() < 0.02 and (st == 1).sum() > 20


def test_nrc_oracle_backprop_matches_finite_differences():
    # The relative L2 loss treats its denominator (y^2 + 0.01) as a constant (stop-gradient), so the FP64
    # backprop is checked against central differences of the loss with denominators frozen at the current
    # prediction.
    rng = np.random.default_rng(0)
    p = (rng.standard_normal(O.NRC_PARAM_COUNT) * 0.1).astype(np.float32)
    smp = np.zeros((8, 16), np.float32)
    smp[:, :12] = rng.uniform(0, 1, (8, 12))
    smp[:, 12:15] = rng.uniform(0, 2, (8, 3))
    lo, hi = np.zeros(3, np.float32), np.ones(3, np.float32)
    y0, g, _ = O.nrc(p, smp, lo, hi, bf16=False)
    den = y0 ** 2 + 0.01
    t = smp[:, 12:15].astype(np.float64)

    def surrogate(q):
        y = O.nrc(q, smp, lo, hi, bf16=False, want_grad=False)[0]
        return ((y - t) ** 2 / den).sum() / len(smp)

    idx = rng.choice(O.NRC_PARAM_COUNT, 60, replace=False)
    for i in idx:
        q1 = p.copy()
        q1[i] += np.float32(1e-3)
        q2 = p.copy()
        q2[i] -= np.float32(1e-3)
        fd = (surrogate(q1) - surrogate(q2)) / (float(q1[i]) - float(q2[i]))
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i])



# ---- SAH treelet refinement -------------------------------------------------------------------------------------
def _all_trees(k):
    """Every full binary tree over leaves 0..k-1 (leaf k inserted above every subtree of every smaller tree):
    (2k-3)!! trees."""
    def insert(t, leaf):
        yield (t, leaf)
        if isinstance(t, tuple):
            a, b = t
            for x in insert(a, leaf):
                yield (x, b)
            for x in insert(b, leaf):
                yield (a, x)
    trees = [0]
    for leaf in range(1, k):
        trees = [x for t in trees for x in insert(t, leaf)]
    return trees


def _union_area64(boxes, mask):
    sel = boxes[[j for j in range(7) if mask >> j & 1]].astype(np.float64)
    lo, hi = sel[:, :3].min(0), sel[:, 3:].max(0)
    d = hi - lo
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0]


def _tree_cost64(t, boxes, leaf_cost):
    if not isinstance(t, tuple):
        return 1 << t, float(leaf_cost[t])
    ma, ca = _tree_cost64(t[0], boxes, leaf_cost)
    mb, cb = _tree_cost64(t[1], boxes, leaf_cost)
    m = ma | mb
    return m, _union_area64(boxes, m) + ca + cb


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_treelet_dp_equals_brute_force_over_all_10395_trees(seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-1, 1, (7, 3)).astype(np.float32)
    ext = rng.uniform(0, 1, (7, 3)) ** (3 if seed % 2 else 1)
    boxes = np.concatenate([lo, (lo + ext).astype(np.float32)], 1)
    leaf_cost = np.where(rng.random(7) < 0.5, 0.0, rng.uniform(0, 20, 7)).astype(np.float32)
    area, cost, part = O.treelet_dp(boxes, leaf_cost)
    trees = _all_trees(7)
    assert len(trees) == 10395
    brute = min(_tree_cost64(t, boxes, leaf_cost)[1] for t in trees)
    assert abs(cost[127] - brute) <= 1e-5 * brute
    # the tree the split table encodes has that cost
    def rebuild(S):
        if bin(S).count("1") == 1:
            return S.bit_length() - 1
        return (rebuild(int(part[S])), rebuild(S ^ int(part[S])))
    assert abs(_tree_cost64(rebuild(127), boxes, leaf_cost)[1] - brute) <= 1e-5 * brute
    for S in range(1, 128):
        assert abs(area[S] - _union_area64(boxes, S)) <= 1e-6 * max(1.0, area[S])
        if bin(S).count("1") > 1:
            P = int(part[S])
            assert P & S == P and 0 < P < S and P & (S & -S)


def _tree_check(left, right, leaf_boxes):
    """Validate a binary tree (Karras codes) and return (float64 SAH from scratch, height)."""
    n = left.size + 1
    kids = np.concatenate([left, right])
    assert np.array_equal(np.sort(-1 - kids[kids < 0]), np.arange(n))
    assert np.array_equal(np.sort(kids[kids >= 0]), np.arange(1, n - 1))
    b64 = leaf_boxes.astype(np.float64)
    def walk(c):
        if c < 0:
            return b64[-1 - c], 0.0, 0
        bl, cl, hl = walk(int(left[c]))
        br, cr, hr = walk(int(right[c]))
        b = np.concatenate([np.minimum(bl[:3], br[:3]), np.maximum(bl[3:], br[3:])])
        d = b[3:] - b[:3]
        return b, d[0] * d[1] + d[1] * d[2] + d[2] * d[0] + cl + cr, 1 + max(hl, hr)
    import sys
    sys.setrecursionlimit(10000)
    b, c, h = walk(0)
    d = b[3:] - b[:3]
    return c / (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]), h


@pytest.mark.parametrize("mesh", ["sphere", "random"])
def test_treelet_refinement_invariants(mesh):
    if mesh == "sphere":
        g = S.icosphere(12, radius=1.0, seed=3, amplitude=0.05)
        tri = g.vertices[g.indices]
    else:
        rng = np.random.default_rng(11)
        c = rng.uniform(-5, 5, (3000, 1, 3))
        tri = (c + rng.normal(0, 0.2, (3000, 3, 3))).astype(np.float32)
    boxes = np.concatenate([tri.min(1), tri.max(1)], 1).astype(np.float32)
    cen = ((boxes[:, :3] + boxes[:, 3:]) * np.float32(0.5)).astype(np.float32)
    keys, order = O.sort_pairs(O.morton(cen, 30))
    kl, kr = O.karras(keys, 32)
    lb = boxes[order]
    res = {r: O.treelet_refine(kl, kr, lb, r) for r in (0, 1, 2, 3)}
    # rounds = 0 is a plain refit: the Karras tree is unchanged
    assert np.array_equal(res[0][0], kl) and np.array_equal(res[0][1], kr)
    prev = None
    for r in (0, 1, 2, 3):
        L, R, cost, hgt, sah = res[r]
        sah64, h = _tree_check(L, R, lb)
        assert abs(sah - sah64) <= 1e-4 * sah64          # reported cost matches the tree
        assert hgt[0] == h and h <= res[0][3][0]          # height never grows
        if prev is not None:
            assert sah <= prev * (1 + 1e-6)              # cost never increases
        prev = sah
    assert res[1][4] < 0.97 * res[0][4]                  # refinement does reduce the SAH cost


# ---- path tracer restatement -------------------------------------------------------------------------------------
@pytest.mark.parametrize("rr_depth", [8, 3])
def test_oracle_render_white_furnace(rr_depth):
    """Closed cube whose walls all emit 1 and reflect a: with NEE over the same walls the estimator's expectation
    is 1 + a + ... + a^8 at max depth 8 (the white-furnace closed form), with or without Russian roulette."""
    a = 0.5
    sc, lights = S.furnace_scene(a, 1.0, 16, 16)
    t = oracle_scene(sc)
    img = O.render(t, sc.records, sc.camera, lights, 16, 16, 128, max_depth=8, rr_depth=rr_depth) / 128
    expected = (1 - a ** 9) / (1 - a)
    assert abs(img.mean() - expected) / expected < 0.01, (img.mean(), expected)


def test_oracle_render_error_falls_as_one_over_spp():
    """Monte Carlo: the squared error against a high-sample estimate falls ~1/spp (independent sample sets)."""
    sc, lights = S.cornell_scene(2, width=12, height=8)
    t = oracle_scene(sc)
    ref = O.render(t, sc.records, sc.camera, lights, 12, 8, 2048, sample_offset=1 << 20) / 2048
    e = []
    for spp in (8, 32):
        x = O.render(t, sc.records, sc.camera, lights, 12, 8, spp) / spp
        e.append(float(np.mean((x - ref) ** 2)))
    assert 2.0 < e[0] / e[1] < 8.0, e
