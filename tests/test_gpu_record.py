"""GPU tests of recording: builds, traces and renders captured into a HIP graph from the context's stream and
replayed - the command-buffer analogue of the reference, which records its builds and its trace into command
buffers and submits them (vulkan-raytraced-triangle/main.cpp:789-812, 930-942, 1340-1356).  Every replay runs with
other GPU work and synchronisations in between and is checked against direct calls on the same data."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rec():
    """A context on a stream of its own (a capture cannot be taken on the null stream)."""
    import pyrtx
    s = torch.cuda.Stream()
    c = pyrtx.Context(0, stream=s)
    yield c, s
    c.sync()
    c.close()


def _dev_geoms(geoms):
    """Device geometry dicts (as Context.upload_scene makes them) for a list of triangle Geometry."""
    out = []
    for g in geoms:
        out.append({"vertices": torch.from_numpy(g.vertices.astype(np.float32)).cuda(),
                    "indices": torch.from_numpy(g.indices.astype(np.uint32).view(np.int32)).cuda(),
                    "transform": None if g.transform is None else
                    torch.from_numpy(np.asarray(g.transform, np.float32).reshape(3, 4)).cuda(),
                    "opaque": g.opaque})
    return out


def _dev(arr, dtype=np.uint8):
    return torch.from_numpy(np.ascontiguousarray(arr).view(dtype).reshape(-1)).cuda()


def _device_scene(count=5, seed=3):
    """Uneven batch scenes with device-resident geometry, SBT records and lights."""
    from pyrtx import scenes as S
    sc, views, lights = S.batch_scenes(count=count, seed=seed, tri_lo=200, tri_hi=20000)
    for k, v in enumerate(views):
        v["width"], v["height"] = [(40, 30), (64, 48), (33, 17)][k % 3]
    lists = [_dev_geoms(geoms) for geoms in sc.blas]
    lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).reshape(-1, 12)).cuda()
    return sc, views, lists, _dev(sc.records), lt


def test_recorded_batched_build_tlas_and_render(rec):
    """One graph holds the batched BLAS build, the TLAS build and the batched render.  Its replays equal the direct
    calls, also after the vertex data changed (a replay rebuilds from the data as it is then); the render of the
    TLAS built inside the recording is recorded in both integrator forms and the device picks one."""
    import pyrtx
    c, s = rec
    _say("recorded batched build + TLAS build + batched render")
    sc, views, lists, records, lt = _device_scene()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        blases = c.build_blas_batch(lists)
        inst = _dev(pyrtx.instance_records(sc, [b.address for b in blases]))
        tlas = c.build_tlas(inst)
        direct = c.render_batch(tlas, records, views, lt, spp=4, max_depth=6).clone()
        c.sync()
        assert c.render_form() == pyrtx.RTX_RENDER_FORM_WHOLE_PATH
        bscratch = blases[0].keepalive[0]
        tscratch = tlas.keepalive[0]
        films = torch.empty_like(direct)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        c.build_blas_batch(lists, into=blases, scratch=bscratch)
        c.build_tlas(inst, into=tlas, scratch=tscratch)
        # RTX_RENDER_GRAPH_BIT inside a caller's capture: recorded as a direct call
        c.render_batch(tlas, records, views, lt, spp=4, max_depth=6, films=films, graph=True)
    assert c.render_form() == pyrtx.RTX_RENDER_FORM_DEVICE_CHOSEN
    for rep in range(3):
        films.fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        c.sync()
        assert torch.equal(films, direct), rep
        x = torch.randn(1 << 20, device="cuda").cumsum(0)  # unrelated work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(x).all()
    # new vertex data: shrink the first scene's mesh towards its centroid, then the direct calls into fresh storage
    v0 = lists[0][0]["vertices"]
    v0.copy_(v0.mean(0, keepdim=True) + 0.8 * (v0 - v0.mean(0, keepdim=True)))
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        blases2 = c.build_blas_batch(lists)
        inst2 = _dev(pyrtx.instance_records(sc, [b.address for b in blases2]))
        tlas2 = c.build_tlas(inst2)
        direct2 = c.render_batch(tlas2, records, views, lt, spp=4, max_depth=6).clone()
        c.sync()
    assert not torch.equal(direct2, direct)
    for rep in range(2):
        films.fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        c.sync()
        assert torch.equal(films, direct2), rep
    _say(f"  5 replays equal the direct calls ({films.shape[0]} pixels over {len(views)} views)")
    del g


def test_recorded_traces_on_a_prebuilt_tlas(rec):
    """Ray-buffer queries (also through the allocation-free entry point) and the reference raygen recorded on a TLAS
    built before the capture, a compacting copy of that TLAS traced in the same recording, and the film's
    conversion to display bytes replay to the direct calls' hits, film and bytes."""
    import pyrtx
    from pyrtx import scenes as S
    c, s = rec
    _say("recorded trace_ray_buffer + trace_rays on a prebuilt TLAS")
    sc, _ = S.cornell_scene(5, width=96, height=64)
    with torch.cuda.stream(s):
        blases, tlas = c.upload_scene(sc)
        rays = torch.from_numpy(S.incoherent_rays(sc, 1 << 16, seed=9)).cuda()
        recs = _dev(sc.records)
        d_hits = torch.empty((rays.shape[0], 6), dtype=torch.int32, device="cuda")
        d_occ = torch.empty((rays.shape[0],), dtype=torch.int32, device="cuda")
        c.trace_ray_buffer(tlas, rays, pyrtx.RTX_QUERY_CLOSEST_HIT, 0xFF, d_hits)
        c.trace_ray_buffer(tlas, rays, pyrtx.RTX_QUERY_ANY_HIT, 0xFF, d_occ)
        d_film = c.trace_rays(tlas, recs, sc.camera, 96, 64).clone()
        c.sync()
        d_rgba = c.film_to_rgba8(d_film).clone()
        size = c.compacted_size(tlas)  # a query: made before the capture
        dst = torch.empty(size, dtype=torch.uint8, device="cuda")
        c.sync()
        hits = torch.empty_like(d_hits)
        hits2 = torch.empty_like(d_hits)
        hits3 = torch.empty_like(d_hits)
        occ = torch.empty_like(d_occ)
        film = torch.empty_like(d_film)
        rgba = torch.empty_like(d_rgba)
    copy_t = pyrtx.AccelerationStructure(dst, dst.data_ptr(), tlas.count, (tlas,))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        c.trace_ray_buffer(tlas, rays, pyrtx.RTX_QUERY_CLOSEST_HIT, 0xFF, hits)
        c.trace_ray_buffer(tlas, rays, pyrtx.RTX_QUERY_ANY_HIT, 0xFF, occ)
        c.trace_rays(tlas, recs, sc.camera, 96, 64, film=film)
        c._check(c.lib.rtx_cmd_copy_as(c.handle, tlas.address, pyrtx.RTX_COPY_MODE_COMPACT, dst.data_ptr(), size,
                                       None))
        c.trace_ray_buffer(copy_t, rays, pyrtx.RTX_QUERY_CLOSEST_HIT, 0xFF, hits2)
        c._check(c.lib.rtx_cmd_trace_ray_buffer_noalloc(c.handle, tlas.address, rays.data_ptr(), rays.shape[0],
                                                        pyrtx.RTX_QUERY_CLOSEST_HIT, 0, 0xFF, hits3.data_ptr(), None))
        c._check(c.lib.rtx_cmd_film_to_rgba8(c.handle, film.data_ptr(), 96, 64, pyrtx.RTX_FILM_REFERENCE_BYTES,
                                             rgba.data_ptr()))
    for rep in range(3):
        hits.fill_(7)
        hits2.fill_(7)
        hits3.fill_(7)
        occ.fill_(7)
        film.fill_(-1.0)
        rgba.fill_(3)
        dst.fill_(0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        c.sync()
        assert torch.equal(hits, d_hits) and torch.equal(occ, d_occ) and torch.equal(film, d_film), rep
        assert torch.equal(hits2, d_hits) and torch.equal(hits3, d_hits) and torch.equal(rgba, d_rgba), rep
    assert int((d_hits[:, 5] != -1).sum()) > 100
    del g


def test_recording_refuses_what_it_cannot_record(rec):
    """Inside a capture, a call that would have to grow a context buffer and a BLAS update (which reads the
    structure's header) fail with their error codes instead of breaking the capture; the graph of what was
    recorded still replays."""
    import pyrtx
    from pyrtx import scenes as S
    c, s = rec
    _say("recording refusals")
    sc, lights = S.cornell_scene(3, width=32, height=32)
    with torch.cuda.stream(s):
        wall_geoms = _dev_geoms(sc.blas[0])
        walls = c.build_blas(wall_geoms)
        geoms = _dev_geoms(sc.blas[1])
        blas = c.build_blas(geoms, pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT | pyrtx.RTX_BUILD_ALLOW_UPDATE_BIT)
        tlas = c.build_tlas(_dev(pyrtx.instance_records(sc, [walls.address, blas.address])))
        recs = _dev(sc.records)
        lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).reshape(-1, 12)).cuda()
        c.render(tlas, recs, sc.camera, lt, 32, 32, spp=1)
        c.sync()
        big = torch.empty((2048, 2048, 4), dtype=torch.float32, device="cuda")
        small = torch.empty((32, 32, 4), dtype=torch.float32, device="cuda")
    errs = {}
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        c.render(tlas, recs, sc.camera, lt, 32, 32, spp=1, film=small)
        try:
            c.render(tlas, recs, sc.camera, lt, 2048, 2048, spp=1, film=big)
        except pyrtx.RtxError as e:
            errs["grow"] = (e.code, str(e))
        try:
            c.update_blas(blas, geoms)
        except pyrtx.RtxError as e:
            errs["update"] = (e.code, str(e))
        try:  # 1,200 geometry descriptors (77 KB) exceed what a recorded build carries
            c.build_blas(wall_geoms * 300, scratch=big.view(torch.uint8))
        except pyrtx.RtxError as e:
            errs["descriptors"] = (e.code, str(e))
    # the recorded render ran on a TLAS whose content was read back before the capture: whole-path launches
    assert c.render_form() == pyrtx.RTX_RENDER_FORM_WHOLE_PATH
    assert errs["grow"][0] == pyrtx.RTX_ERROR_OUT_OF_DEVICE_MEMORY and "before recording" in errs["grow"][1]
    assert errs["update"][0] == pyrtx.RTX_ERROR_INVALID_ARGUMENT and "cannot be recorded" in errs["update"][1]
    assert errs["descriptors"][0] == pyrtx.RTX_ERROR_INVALID_ARGUMENT and "at most 65536 bytes" in errs["descriptors"][1]
    with torch.cuda.stream(s):
        ref = c.render(tlas, recs, sc.camera, lt, 32, 32, spp=1).clone()
        c.sync()
    small.fill_(-1.0)
    g.replay()
    torch.cuda.synchronize()
    c.sync()
    assert torch.equal(small, ref)
    del g


def test_recorded_nrc_frames_train_on(rec):
    """NRC frames recorded by a capture of the context stream (the training wavefront's side stream joins it
    through the frame's fork / join events): each replay trains the cache again, and film, parameters and step
    count equal direct frames on an identical cache, replay after replay."""
    import pyrtx
    from pyrtx import scenes as S
    c, s = rec
    _say("recorded NRC frames")
    W, H = 80, 60
    sc, lights = S.cornell_scene(3, width=W, height=H)
    lo, hi = S.scene_bounds(sc)
    with torch.cuda.stream(s):
        blases, tlas = c.upload_scene(sc)
        recs = _dev(sc.records)
        lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).reshape(-1, 12)).cuda()
        direct = pyrtx.Nrc(c, lo, hi, seed=2, batch_size=2048, cache_vertex=1)
        graph = pyrtx.Nrc(c, lo, hi, seed=2, batch_size=2048, cache_vertex=1)
        fd = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        fg = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        for n in (direct, graph):  # a first frame on each cache outside the capture sizes its buffers
            c.render_nrc(tlas, recs, sc.camera, lt, n, W, H, spp=2, train_paths=4096, sample_offset=0,
                         film=fd if n is direct else fg)
        c.sync()
    try:
        assert torch.equal(fd, fg)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
            c.render_nrc(tlas, recs, sc.camera, lt, graph, W, H, spp=2, train_paths=4096, sample_offset=1, film=fg)
        for fr in range(3):
            with torch.cuda.stream(s):
                c.render_nrc(tlas, recs, sc.camera, lt, direct, W, H, spp=2, train_paths=4096, sample_offset=1,
                             film=fd)
                c.sync()
            fg.fill_(-1.0)
            g.replay()
            torch.cuda.synchronize()
            c.sync()
            assert torch.equal(fd, fg), fr
            assert np.array_equal(direct.params(), graph.params()), fr
            assert direct.status() == graph.status(), fr
            x = torch.randn(1 << 20, device="cuda").cumsum(0)  # unrelated work between replays
            torch.cuda.synchronize()
            assert torch.isfinite(x).all()
        assert graph.status()[0] >= 4 and float(fg[..., :3].mean()) > 0.01
        del g
    finally:
        direct.close()
        graph.close()


def test_device_picks_the_traversal_of_a_tlas_built_in_the_recording(rec):
    """A TLAS rebuilt inside the recording from an instance buffer that the caller switches between replays -
    opaque BLASes, then a cutout (any-hit) BLAS, then opaque again: each replay's render and closest-hit queries
    equal the direct calls on that content, so the guarded launches follow the content the device reads."""
    import pyrtx
    from pyrtx import scenes as S
    c, s = rec
    _say("device-chosen traversal for a TLAS built in the recording")
    W, H = 48, 32
    opaque, lights = S.cornell_scene(3, width=W, height=H)
    cut, _ = S.cornell_scene(3, width=W, height=H)
    cut.blas[1][0].opaque = False
    cut.records["any_hit"][4] = S.ANY_HIT_CUTOUT | (8 << 8)
    rays = torch.from_numpy(S.incoherent_rays(opaque, 1 << 14, seed=5)).cuda()
    with torch.cuda.stream(s):
        recs = _dev(cut.records)  # the cutout program only runs on non-opaque geometry
        lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).reshape(-1, 12)).cuda()
        b_op = [c.build_blas(_dev_geoms(g)) for g in opaque.blas]
        b_cut = [c.build_blas(_dev_geoms(g)) for g in cut.blas]
        inst = {k: _dev(pyrtx.instance_records(sc, [b.address for b in bl]))
                for k, sc, bl in (("opaque", opaque, b_op), ("cutout", cut, b_cut))}
        ref = {}
        for k in ("opaque", "cutout"):
            t = c.build_tlas(inst[k])
            c.sync()
            film = c.render(t, recs, opaque.camera, lt, W, H, spp=2).clone()
            hits = c.trace_ray_buffer(t, rays, pyrtx.RTX_QUERY_CLOSEST_HIT, 0xFF).clone()
            c.sync()
            ref[k] = (film, hits, c.render_form())
        assert ref["opaque"][2] == pyrtx.RTX_RENDER_FORM_WHOLE_PATH
        assert ref["cutout"][2] == pyrtx.RTX_RENDER_FORM_WAVEFRONT
        assert not torch.equal(ref["opaque"][0], ref["cutout"][0])
        buf = inst["opaque"].clone()
        tlas = c.build_tlas(buf)
        film = torch.empty_like(ref["opaque"][0])
        hits = torch.empty_like(ref["opaque"][1])
        c.sync()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        c.build_tlas(buf, into=tlas, scratch=tlas.keepalive[0])
        c.render(tlas, recs, opaque.camera, lt, W, H, spp=2, film=film)
        c.trace_ray_buffer(tlas, rays, pyrtx.RTX_QUERY_CLOSEST_HIT, 0xFF, hits)
    assert c.render_form() == pyrtx.RTX_RENDER_FORM_DEVICE_CHOSEN
    for k in ("opaque", "cutout", "opaque", "cutout"):
        buf.copy_(inst[k])
        film.fill_(-1.0)
        hits.fill_(7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        c.sync()
        assert torch.equal(film, ref[k][0]), k
        assert torch.equal(hits, ref[k][1]), k
    del g
