"""Seeded synthetic scenes and ray sets (input generation, numpy only).

Scenes:
  * R0  — the scene hard-coded in vulkan-raytraced-triangle/main.cpp: one quad (2 triangles, main.cpp:676-682)
          used by 2 geometries with x -/+ 2 transforms (main.cpp:684-695, 743), 2 instances at y +/- 2 with
          custom index 100 and SBT offsets 0 / 2 (main.cpp:835-858), camera (0,0,10) 60 deg (main.cpp:1015),
          1200x800 (main.cpp:13-14), hit records main.cpp:1307-1317, miss (0,0,0.2) main.cpp:1065.
  * C1 / C1' — one triangle (1,1,0),(-1,1,0),(0,-1,0), identity instance, camera at z = -2.5 / +2.5, 800x600.
  * geodesic icospheres (frequency n: 20 n^2 triangles) displaced by seeded value noise, height-field terrains,
    a Cornell-style box, and the 512-instance / 20M-triangle ray-query scene with coherent and incoherent rays.

Every mesh is watertight: a vertex on an edge shared by two icosahedron faces is computed from the same two
corners with the same weights in the same order (corners summed by ascending global index), so both faces
produce bit-identical coordinates.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------------------------------------------
# scene containers (host numpy; pyrtx uploads them)
# ---------------------------------------------------------------------------------------------------------------


@dataclass
class Geometry:
    vertices: np.ndarray            # float32 [V, 3]
    indices: np.ndarray             # uint32 [T, 3]
    transform: np.ndarray | None = None  # float32 [3, 4] row-major, applied at build time
    opaque: bool = True             # RTX_GEOMETRY_OPAQUE_BIT; False runs the hit record's any-hit program


@dataclass
class AabbGeometry:
    """Procedural geometry (RTX_GEOMETRY_TYPE_AABBS): one box per primitive, min xyz / max xyz; its surface comes
    from the hit record's intersection program."""
    boxes: np.ndarray               # float32 [N, 6]
    opaque: bool = True
    transform: None = None

    @property
    def indices(self) -> np.ndarray:  # no triangles
        return np.zeros((0, 3), np.uint32)

    @property
    def vertices(self) -> np.ndarray:  # the boxes' min and max corners (for bounds)
        b = np.asarray(self.boxes, np.float32).reshape(-1, 6)
        return np.concatenate([b[:, :3], b[:, 3:]], 0)


@dataclass
class Scene:
    blas: list                      # list of list[Geometry]
    transforms: np.ndarray          # float32 [I, 3, 4] object -> world
    blas_of_instance: np.ndarray    # int [I]
    custom_index: np.ndarray        # uint32 [I]
    mask: np.ndarray                # uint32 [I]
    sbt_offset: np.ndarray          # uint32 [I]
    flags: np.ndarray               # uint32 [I]
    records: np.ndarray | None = None   # HIT_RECORD_DTYPE [R]
    miss_color: tuple = (0.0, 0.0, 0.2)
    camera: tuple = (0.0, 0.0, 10.0, 60.0)
    width: int = 1200
    height: int = 800
    meta: dict = field(default_factory=dict)

    @property
    def triangle_count(self) -> int:
        return int(sum(sum(g.indices.shape[0] for g in self.blas[b]) for b in self.blas_of_instance))

    @property
    def unique_triangle_count(self) -> int:
        return int(sum(sum(g.indices.shape[0] for g in geoms) for geoms in self.blas))


# rtx_hit_record (48 B): program, color[3], emission[3], match_primitive/instance/custom_index/geometry, any_hit
HIT_RECORD_DTYPE = np.dtype([("program", np.uint32), ("color", np.float32, (3,)), ("emission", np.float32, (3,)),
                             ("match", np.int32, (4,)), ("any_hit", np.uint32)])
assert HIT_RECORD_DTYPE.itemsize == 48

PROGRAM_RECORD_COLOR, PROGRAM_BARYCENTRIC, PROGRAM_MATCH_BARYCENTRIC, PROGRAM_LAMBERTIAN = 0, 1, 2, 3
INSTANCE_TRIANGLE_FACING_CULL_DISABLE = 0x1
ANY_HIT_ACCEPT = 0      # rtx_hit_record.any_hit programs (RTX_ANY_HIT_*); parameter in bits 8-31
ANY_HIT_CUTOUT = 1
INTERSECTION_BOX = 0    # intersection programs of AABB geometry: bits 8-15 of rtx_hit_record.program
INTERSECTION_SPHERE = 1


def hit_group(closest: int, intersection: int = INTERSECTION_BOX) -> int:
    """rtx_hit_record.program word (RTX_HIT_GROUP)."""
    return closest | (intersection << 8)


def hit_records(programs, colors, emissions=None, match=None, any_hit=None) -> np.ndarray:
    """any_hit: RTX_ANY_HIT_PROGRAM words (program | parameter << 8) per record, default accept."""
    n = len(programs)
    r = np.zeros(n, HIT_RECORD_DTYPE)
    if any_hit is not None:
        r["any_hit"] = np.asarray(any_hit, np.uint32).reshape(n)
    r["program"] = programs
    r["color"] = np.asarray(colors, np.float32).reshape(n, 3)
    if emissions is not None:
        r["emission"] = np.asarray(emissions, np.float32).reshape(n, 3)
    if match is not None:
        r["match"] = np.asarray(match, np.int32).reshape(n, 4)
    else:
        r["match"] = -1
    return r


def translation(x=0.0, y=0.0, z=0.0) -> np.ndarray:
    m = np.zeros((3, 4), np.float32)
    m[0, 0] = m[1, 1] = m[2, 2] = 1.0
    m[:, 3] = (x, y, z)
    return m


# ---------------------------------------------------------------------------------------------------------------
# reference scenes
# ---------------------------------------------------------------------------------------------------------------

def scene_r0() -> Scene:
    """The reference's own scene (main.cpp:674-949, 1001-1017, 1264-1320)."""
    verts = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    idx = np.array([[0, 1, 3], [1, 2, 3]], np.uint32)
    geoms = [Geometry(verts, idx, translation(x=-2.0)), Geometry(verts, idx, translation(x=2.0))]
    transforms = np.stack([translation(y=2.0), translation(y=-2.0)])
    colors = [(0.6, 0.1, 0.2), (0.1, 0.8, 0.4), (0.9, 0.7, 0.1), (0.3, 0.6, 0.9)]
    rec = hit_records([PROGRAM_MATCH_BARYCENTRIC] * 4, colors, match=[(1, 1, 100, 1)] * 4)
    return Scene(blas=[geoms], transforms=transforms, blas_of_instance=np.array([0, 0]),
                 custom_index=np.array([100, 100], np.uint32), mask=np.array([0xFF, 0xFF], np.uint32),
                 sbt_offset=np.array([0, 2], np.uint32),
                 flags=np.array([INSTANCE_TRIANGLE_FACING_CULL_DISABLE] * 2, np.uint32),
                 records=rec, miss_color=(0.0, 0.0, 0.2), camera=(0.0, 0.0, 10.0, 60.0), width=1200, height=800,
                 meta={"name": "R0"})


def scene_c1(camera_z: float = -2.5) -> Scene:
    """Single-triangle scene: (1,1,0), (-1,1,0), (0,-1,0), identity instance, barycentric closest-hit."""
    verts = np.array([[1, 1, 0], [-1, 1, 0], [0, -1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    rec = hit_records([PROGRAM_BARYCENTRIC], [(0.0, 0.0, 0.0)])
    return Scene(blas=[[Geometry(verts, idx)]], transforms=translation()[None], blas_of_instance=np.array([0]),
                 custom_index=np.zeros(1, np.uint32), mask=np.array([0xFF], np.uint32),
                 sbt_offset=np.zeros(1, np.uint32), flags=np.zeros(1, np.uint32), records=rec,
                 miss_color=(0.0, 0.0, 0.2), camera=(0.0, 0.0, camera_z, 60.0), width=800, height=600,
                 meta={"name": "C1" if camera_z < 0 else "C1'"})


# ---------------------------------------------------------------------------------------------------------------
# procedural meshes
# ---------------------------------------------------------------------------------------------------------------

def _hash3(ix, iy, iz, seed):
    """Integer lattice hash (uint32 arithmetic) -> float64 in [-1, 1)."""
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iy.astype(np.uint32) * np.uint32(0xD8163841)) ^ \
            (iz.astype(np.uint32) * np.uint32(0xCB1AB31F)) ^ np.uint32(seed * 0x9E3779B9 & 0xFFFFFFFF)
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(0x7FEB352D)
        h = h ^ (h >> np.uint32(15))
        h = h * np.uint32(0x846CA68B)
        h = h ^ (h >> np.uint32(16))
    return h.astype(np.float64) * (2.0 / 4294967296.0) - 1.0


def value_noise(p: np.ndarray, seed: int, octaves: int = 3, freq: float = 4.0) -> np.ndarray:
    """Seeded 3D value noise (trilinear, smoothstep), sum of octaves with amplitude 1/2^k, range ~[-1, 1]."""
    p = np.asarray(p, np.float64)
    out = np.zeros(p.shape[0], np.float64)
    amp, f, norm = 1.0, freq, 0.0
    for k in range(octaves):
        q = p * f
        i0 = np.floor(q)
        fr = q - i0
        w = fr * fr * (3.0 - 2.0 * fr)
        i0 = i0.astype(np.int64)
        acc = np.zeros(p.shape[0], np.float64)
        for dx in (0, 1):
            wx = w[:, 0] if dx else 1.0 - w[:, 0]
            for dy in (0, 1):
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                for dz in (0, 1):
                    wz = w[:, 2] if dz else 1.0 - w[:, 2]
                    acc += wx * wy * wz * _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + 101 * k)
        out += amp * acc
        norm += amp
        amp *= 0.5
        f *= 2.0
    return out / norm


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def icosphere(freq: int, radius: float = 1.0, seed: int | None = None, amplitude: float = 0.0,
              noise_freq: float = 4.0, center=(0.0, 0.0, 0.0)) -> Geometry:
    """Geodesic icosphere with 20*freq^2 triangles, optionally displaced along the normal by value noise."""
    n = int(freq)
    cv, cf = _icosahedron()
    ii, jj = [], []
    for i in range(n + 1):
        for j in range(n + 1 - i):
            ii.append(i)
            jj.append(j)
    ii = np.array(ii, np.int64)
    jj = np.array(jj, np.int64)
    npts = ii.size
    w = np.stack([n - ii - jj, ii, jj], axis=1).astype(np.float64)  # weights of corners (a, b, c)
    pts = np.empty((20, npts, 3), np.float64)
    for f in range(20):
        corners = cf[f]
        order = np.argsort(corners)  # sum corners by ascending global index
        acc = w[:, order[0], None] * cv[corners[order[0]]]
        acc = acc + w[:, order[1], None] * cv[corners[order[1]]]
        acc = acc + w[:, order[2], None] * cv[corners[order[2]]]
        pts[f] = acc / n
    pts = pts.reshape(-1, 3)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    if seed is not None and amplitude != 0.0:
        r = radius + amplitude * radius * value_noise(pts, seed, 3, noise_freq)
    else:
        r = np.full(pts.shape[0], radius)
    pts = pts * r[:, None] + np.asarray(center, np.float64)
    # triangles of the per-face grid
    lut = -np.ones((n + 1, n + 1), np.int64)
    lut[ii, jj] = np.arange(npts)
    tri = []
    for i in range(n):
        j = np.arange(n - i)
        tri.append(np.stack([lut[i, j], lut[i + 1, j], lut[i, j + 1]], 1))
        if n - i - 1 > 0:
            j2 = np.arange(n - i - 1)
            tri.append(np.stack([lut[i + 1, j2], lut[i + 1, j2 + 1], lut[i, j2 + 1]], 1))
    tri = np.concatenate(tri, 0)
    faces = (tri[None, :, :] + (np.arange(20) * npts)[:, None, None]).reshape(-1, 3)
    return Geometry(pts.astype(np.float32), faces.astype(np.uint32))


def terrain(cells: int = 1024, size: float = 2.0, height: float = 0.25, seed: int = 0) -> Geometry:
    """Height field of cells x cells quads (2 triangles each) over [-size/2, size/2]^2 in xz."""
    n = cells
    xs = np.linspace(-size / 2, size / 2, n + 1)
    gx, gz = np.meshgrid(xs, xs, indexing="xy")
    p = np.stack([gx.ravel(), np.zeros(gx.size), gz.ravel()], 1)
    h = height * value_noise(p * (1.0 / size) + 0.5, seed, 3, 6.0)
    verts = np.stack([gx.ravel(), h, gz.ravel()], 1).astype(np.float32)
    r = np.arange(n)
    a = (r[None, :] + (n + 1) * r[:, None]).ravel()  # lower-left corner of each cell
    b, c, d = a + 1, a + n + 1, a + n + 2
    idx = np.concatenate([np.stack([a, b, c], 1), np.stack([b, d, c], 1)], 0).astype(np.uint32)
    return Geometry(verts, idx)


def random_rotation(rng, n):
    """Uniform random rotations (Shoemake 1992 quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)], 1)
    return R


# ---------------------------------------------------------------------------------------------------------------
# the 512-instance ray-query scene
# ---------------------------------------------------------------------------------------------------------------

QUERY_SCENE_FREQS = (160, 236, 290, 332, 447)  # 20 n^2 = 512000 .. 3996180 triangles


def query_scene(scale_tris: float = 1.0, instances: int = 512, seed: int = 4242, terrain_cells: int = 1024) -> Scene:
    """512 instances of 10 base meshes (5 displaced icospheres, 5 terrains) in a 100^3 box.

    scale_tris < 1 shrinks every base mesh (frequency * sqrt(scale), cells * sqrt(scale)) for parity tests at
    oracle-friendly sizes; 1.0 is the full 19,994,340-triangle configuration."""
    s = math.sqrt(scale_tris)
    meshes = []
    for k, f in enumerate(QUERY_SCENE_FREQS):
        fk = max(2, int(round(f * s)))
        meshes.append([icosphere(fk, radius=1.0, seed=42 + k, amplitude=0.05, noise_freq=3.0)])
    for k in range(5):
        ck = max(2, int(round(terrain_cells * s)))
        meshes.append([terrain(ck, size=2.0, height=0.3, seed=142 + k)])
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-50.0, 50.0, size=(instances, 3))
    R = random_rotation(rng, instances)
    sc = rng.uniform(0.5, 2.0, size=instances)
    xf = np.zeros((instances, 3, 4), np.float64)
    xf[:, :, :3] = R * sc[:, None, None]
    xf[:, :, 3] = pos
    return Scene(blas=meshes, transforms=xf.astype(np.float32), blas_of_instance=np.arange(instances) % 10,
                 custom_index=np.arange(instances, dtype=np.uint32), mask=np.full(instances, 0xFF, np.uint32),
                 sbt_offset=np.zeros(instances, np.uint32), flags=np.zeros(instances, np.uint32),
                 camera=(0.0, 0.0, 150.0, 60.0), width=4096, height=4096,
                 meta={"name": "query", "scale_tris": scale_tris, "seed": seed})


def camera_rays(camera, width: int, height: int, tmin: float = 0.0, tmax: float = np.inf) -> np.ndarray:
    """Rays of the reference raygen program (main.cpp:1033-1052) as 32-B records [N, 8], row-major pixels.

    FP32 steps in the raygen order: s = (id + 0.5) / size; ndc = s * 2 - 1; dir = (ndc.x*ax, -(ndc.y*ay), -1),
    aspect_y = float(tan(radians(fov)/2)), aspect_x = aspect_y * float(W) / float(H)."""
    cam = np.asarray(camera, np.float32)
    ay = np.float32(math.tan(float(cam[3]) * (math.pi / 180.0) * 0.5))
    ax = np.float32(np.float32(ay * np.float32(width)) / np.float32(height))
    x = np.arange(width, dtype=np.float32)
    y = np.arange(height, dtype=np.float32)
    sx = (x + np.float32(0.5)) / np.float32(width)
    sy = (y + np.float32(0.5)) / np.float32(height)
    nx = sx * np.float32(2.0) - np.float32(1.0)
    ny = sy * np.float32(2.0) - np.float32(1.0)
    rays = np.empty((height, width, 8), np.float32)
    rays[..., 0] = cam[0]
    rays[..., 1] = cam[1]
    rays[..., 2] = cam[2]
    rays[..., 3] = tmin
    rays[..., 4] = (nx * ax)[None, :]
    rays[..., 5] = (-(ny * ay))[:, None]
    rays[..., 6] = -1.0
    rays[..., 7] = tmax
    return rays.reshape(-1, 8)


def scene_bounds(scene: Scene):
    """World AABB of all instances (from transformed mesh bounding-box corners)."""
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    mesh_box = []
    for geoms in scene.blas:
        v = np.concatenate([g.vertices for g in geoms], 0).astype(np.float64)
        mesh_box.append((v.min(0), v.max(0)))
    for i, b in enumerate(scene.blas_of_instance):
        mlo, mhi = mesh_box[b]
        corners = np.array([[x, y, z] for x in (mlo[0], mhi[0]) for y in (mlo[1], mhi[1]) for z in (mlo[2], mhi[2])])
        M = scene.transforms[i].astype(np.float64)
        w = corners @ M[:, :3].T + M[:, 3]
        lo = np.minimum(lo, w.min(0))
        hi = np.maximum(hi, w.max(0))
    return lo, hi


def incoherent_rays(scene: Scene, n: int, seed: int = 1234, tmin: float = 0.0, tmax: float = np.inf) -> np.ndarray:
    """Origins uniform in the scene AABB, directions uniform on the sphere (z = 1 - 2u1, phi = 2 pi u2)."""
    lo, hi = scene_bounds(scene)
    rng = np.random.default_rng(seed)
    o = lo + (hi - lo) * rng.random((n, 3))
    u1, u2 = rng.random(n), rng.random(n)
    z = 1.0 - 2.0 * u1
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * u2
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    rays = np.empty((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = tmin
    rays[:, 4:7] = d
    rays[:, 7] = tmax
    return rays


# ---------------------------------------------------------------------------------------------------------------
# path-tracing scenes
# ---------------------------------------------------------------------------------------------------------------

def _quad(a, b, c, d):
    """Two triangles (a, b, c), (a, c, d); their geometric normal follows (b - a) x (c - a)."""
    v = np.array([a, b, c, d], np.float32)
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _merge(parts):
    vs, ix, off = [], [], 0
    for v, i in parts:
        vs.append(v)
        ix.append(i + off)
        off += len(v)
    return np.concatenate(vs).astype(np.float32), np.concatenate(ix).astype(np.uint32)


CORNELL_WHITE = (0.75, 0.75, 0.75)
CORNELL_RED = (0.75, 0.15, 0.15)
CORNELL_GREEN = (0.15, 0.75, 0.15)
CORNELL_LIGHT_RADIANCE = 15.0


def cornell_scene(sphere_level: int = 8, seed: int = 1234, width: int = 1920, height: int = 1080,
                  closed: bool = True):
    """Cornell-style unit cube with a displaced icosphere.

    closed (default): the 6 walls of the unit cube (12 triangles; the camera is inside), an open-bottom block
    (5 faces, 10) and a 0.25 x 0.25 light quad (2) just under the ceiling facing down with radiance 15: 24
    triangles.  Camera: the reference raygen model at (0.5, 0.5, 0.99) looking down -z, 75 deg vertical field of
    view.  Icosphere centre (0.58, 0.31, 0.36).
    closed=False: the classic open-front box: 5 walls (10), a closed block (12), the light (2); camera outside at
    (0.5, 0.5, 1.9), 40 deg; icosphere centre (0.6, 0.32, 0.45).
    Albedos: walls and block 0.75 white, left wall (x = 0) red (0.75, 0.15, 0.15), right wall (x = 1) green
    (0.15, 0.75, 0.15).  Icosphere: subdivision level `sphere_level` (20 * 4^level triangles; level 8 =
    1,310,720), radius 0.3, displaced along the normal by 3-octave value noise of amplitude 0.03 (seed).
    Returns (Scene, lights [2, 12])."""
    white = [_quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)),      # floor
             _quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),      # ceiling
             _quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0))]      # back
    if closed:
        white.append(_quad((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)))  # front wall behind the camera
    x0, x1, z0, z1, h = 0.08, 0.30, 0.62, 0.84, 0.30
    block = [_quad((x0, h, z0), (x0, h, z1), (x1, h, z1), (x1, h, z0))] + \
        ([] if closed else [_quad((x0, 0, z0), (x1, 0, z0), (x1, 0, z1), (x0, 0, z1))]) + [
             _quad((x0, 0, z0), (x1, 0, z0), (x1, h, z0), (x0, h, z0)),
             _quad((x0, 0, z1), (x0, h, z1), (x1, h, z1), (x1, 0, z1)),
             _quad((x0, 0, z0), (x0, h, z0), (x0, h, z1), (x0, 0, z1)),
             _quad((x1, 0, z0), (x1, 0, z1), (x1, h, z1), (x1, h, z0))]
    red = [_quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0))]
    green = [_quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1))]
    ly = 1.0 - 1e-3
    lv, li = _quad((0.375, ly, 0.375), (0.625, ly, 0.375), (0.625, ly, 0.625), (0.375, ly, 0.625))
    geoms = [Geometry(*_merge(white + block)), Geometry(*_merge(red)), Geometry(*_merge(green)), Geometry(lv, li)]
    assert sum(g.indices.shape[0] for g in geoms) == 24
    centre = (0.58, 0.31, 0.36) if closed else (0.6, 0.32, 0.45)
    sph = icosphere(2 ** sphere_level, radius=0.3, seed=seed, amplitude=0.1, noise_freq=4.0, center=centre)
    recs = hit_records([PROGRAM_LAMBERTIAN] * 5,
                       [CORNELL_WHITE, CORNELL_RED, CORNELL_GREEN, (0.0, 0.0, 0.0), CORNELL_WHITE],
                       emissions=[(0, 0, 0), (0, 0, 0), (0, 0, 0), (CORNELL_LIGHT_RADIANCE,) * 3, (0, 0, 0)])
    lights = np.zeros((2, 12), np.float32)
    for k in range(2):
        lights[k, :9] = lv[li[k]].reshape(-1)
        lights[k, 9:] = CORNELL_LIGHT_RADIANCE
    sc = Scene(blas=[geoms, [sph]], transforms=np.stack([translation(), translation()]),
               blas_of_instance=np.array([0, 1]), custom_index=np.zeros(2, np.uint32),
               mask=np.full(2, 0xFF, np.uint32), sbt_offset=np.array([0, 4], np.uint32), flags=np.zeros(2, np.uint32),
               records=recs, camera=(0.5, 0.5, 0.99, 75.0) if closed else (0.5, 0.5, 1.9, 40.0), width=width,
               height=height, meta={"name": "cornell", "sphere_level": sphere_level, "closed": closed})
    return sc, lights


def furnace_scene(albedo: float = 0.5, emission: float = 1.0, width: int = 64, height: int = 64):
    """Closed unit cube whose 12 inward-facing triangles all emit `emission` and reflect `albedo`.  With NEE over
    the same triangles the expected radiance is emission * (1 + a + ... + a^max_depth) (white furnace)."""
    faces = [_quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)),   # floor, normal +y
             _quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)),   # ceiling, normal -y
             _quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)),   # back z=0, normal +z
             _quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)),   # front z=1, normal -z
             _quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)),   # left x=0, normal +x
             _quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0))]   # right x=1, normal -x
    v, i = _merge(faces)
    recs = hit_records([PROGRAM_LAMBERTIAN], [(albedo,) * 3], emissions=[(emission,) * 3])
    lights = np.zeros((12, 12), np.float32)
    lights[:, :9] = v[i].reshape(12, 9)
    lights[:, 9:] = emission
    sc = Scene(blas=[[Geometry(v, i)]], transforms=translation()[None], blas_of_instance=np.array([0]),
               custom_index=np.zeros(1, np.uint32), mask=np.full(1, 0xFF, np.uint32),
               sbt_offset=np.zeros(1, np.uint32), flags=np.zeros(1, np.uint32), records=recs,
               camera=(0.5, 0.5, 0.9, 60.0), width=width, height=height, meta={"name": "furnace"})
    return sc, lights


BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1024, 768))


def batch_scenes(count: int = 64, seed: int = 7, tri_lo: float = 1e3, tri_hi: float = 2e6, spacing: float = 3.0,
                 tri_scale: float = 1.0):
    """Uneven scenes for batched rendering, laid out as a scene atlas in one TLAS.

    Scene k: a displaced icosphere (radius 0.3 at (0.5, 0.35, 0.5); triangle count log-uniform in [tri_lo, tri_hi]
    x tri_scale, i.e. 20 f^2 with f = round(sqrt(T / 20))), a ground quad (y = 0, x and z in [-0.5, 1.5]) and one
    area light (0.3 x 0.3 quad at y = 1.2, facing down, radiance 10); camera at (0.5, 0.55, 2.0) looking down -z
    with a 45 deg vertical field of view; film size drawn from 256^2, 512^2, 1024x768.  Scene k's instance is
    translated to (spacing * (k % 8), 0, spacing * (k // 8)); its clip box is the local box
    [-0.6, 1.6] x [-0.05, 1.3] x [-0.6, 2.1] (a margin around the geometry) moved the same way.  Seeds: sizes and film sizes from `seed`, the
    noise of icosphere k from seed * 1000 + k.
    Returns (Scene, views: list of dicts for rtx_render_view, lights [2 * count, 12])."""
    rng = np.random.default_rng(seed)
    tris = np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi), count)) * tri_scale
    res = [BATCH_RESOLUTIONS[i] for i in rng.integers(0, len(BATCH_RESOLUTIONS), count)]
    albedo = rng.uniform(0.3, 0.8, (count, 3))
    blas, transforms, recs, views = [], [], [], []
    lights = np.zeros((2 * count, 12), np.float32)
    for k in range(count):
        f = max(1, int(round(np.sqrt(tris[k] / 20.0))))
        sph = icosphere(f, radius=0.3, seed=seed * 1000 + k, amplitude=0.03, noise_freq=4.0, center=(0.5, 0.35, 0.5))
        gv, gi = _quad((-0.5, 0, -0.5), (-0.5, 0, 1.5), (1.5, 0, 1.5), (1.5, 0, -0.5))      # normal +y
        lv, li = _quad((0.35, 1.2, 0.35), (0.65, 1.2, 0.35), (0.65, 1.2, 0.65), (0.35, 1.2, 0.65))  # normal -y
        blas.append([sph, Geometry(gv, gi), Geometry(lv, li)])
        ox, oz = spacing * (k % 8), spacing * (k // 8)
        transforms.append(translation(ox, 0.0, oz))
        recs.append(hit_records([PROGRAM_LAMBERTIAN] * 3, [tuple(albedo[k]), (0.5, 0.5, 0.5), (0.0, 0.0, 0.0)],
                                emissions=[(0, 0, 0), (0, 0, 0), (10.0, 10.0, 10.0)]))
        off = np.array([ox, 0.0, oz], np.float32)
        for t in range(2):
            lights[2 * k + t, :9] = (lv[li[t]] + off).reshape(-1)
            lights[2 * k + t, 9:] = 10.0
        w, h = res[k]
        views.append({"camera": (0.5 + ox, 0.55, 2.0 + oz, 45.0), "width": w, "height": h,
                      "light_first": 2 * k, "light_count": 2,
                      "clip_min": (-0.6 + ox, -0.05, -0.6 + oz), "clip_max": (1.6 + ox, 1.3, 2.1 + oz),
                      "triangles": int(sph.indices.shape[0]) + 4})
    sc = Scene(blas=blas, transforms=np.stack(transforms), blas_of_instance=np.arange(count),
               custom_index=np.arange(count, dtype=np.uint32), mask=np.full(count, 0xFF, np.uint32),
               sbt_offset=(3 * np.arange(count)).astype(np.uint32), flags=np.zeros(count, np.uint32),
               records=np.concatenate(recs), camera=views[0]["camera"], width=res[0][0], height=res[0][1],
               meta={"name": "batch", "count": count, "seed": seed})
    return sc, views, lights
