"""rtx_cmd_trace_ray_buffer_noalloc (working memory in a file-scope device array, no allocation per call) against
rtx_cmd_trace_ray_buffer through the C ABI: identical results on two calls in a row, the second with fewer rays;
launch shapes whose stack spill does not fit the array are rejected with RTX_ERROR_OUT_OF_DEVICE_MEMORY; stack
overflow in the array is reported by rtx_sync."""
import ctypes
import sys

import numpy as np
import pytest

from helpers import small_mixed_scene
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


def _both(ctx, tlas, rays, mode):
    """(rtx_cmd_trace_ray_buffer result, rtx_cmd_trace_ray_buffer_noalloc result) as numpy arrays."""
    import pyrtx
    n = rays.shape[0]
    shape = (n, 6) if mode == pyrtx.RTX_QUERY_CLOSEST_HIT else (n,)
    ref = torch.full(shape, -7, dtype=torch.int32, device=ctx.device)
    got = torch.full(shape, -9, dtype=torch.int32, device=ctx.device)
    closest = mode == pyrtx.RTX_QUERY_CLOSEST_HIT
    ctx._check(ctx.lib.rtx_cmd_trace_ray_buffer(ctx.handle, tlas.address, rays.data_ptr(), n, mode, 0, 0xFF,
                                                 ref.data_ptr() if closest else None,
                                                 None if closest else ref.data_ptr()))
    ctx._check(ctx.lib.rtx_cmd_trace_ray_buffer_noalloc(ctx.handle, tlas.address, rays.data_ptr(), n, mode, 0, 0xFF,
                                                        got.data_ptr() if closest else None,
                                                        None if closest else got.data_ptr()))
    ctx.sync()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_matches_trace_ray_buffer(ctx):
    import pyrtx
    _say("no-alloc ray-buffer queries vs rtx_cmd_trace_ray_buffer")
    for sc, n1, n2 in ((small_mixed_scene(seed=5, instances=24, freq=14, cells=20), 40000, 9000),
                       (S.query_scene(scale_tris=0.005), 60000, 20000)):
        blases, tlas = ctx.upload_scene(sc)
        rays = np.concatenate([S.camera_rays(sc.camera, 160, 120), S.incoherent_rays(sc, n1, seed=9)])
        for n in (len(rays), n2):  # two calls in a row, the second with fewer rays
            r = torch.from_numpy(np.ascontiguousarray(rays[:n])).cuda()
            for mode in (pyrtx.RTX_QUERY_CLOSEST_HIT, pyrtx.RTX_QUERY_ANY_HIT):
                ref, got = _both(ctx, tlas, r, mode)
                assert np.array_equal(ref, got), (sc.meta["name"], n, mode)
                if mode == pyrtx.RTX_QUERY_CLOSEST_HIT:
                    hits = int((ref[:, 5].view(np.uint32) != 0xFFFFFFFF).sum())
                    _say(f"  {sc.meta['name']}: {n} rays, {hits} hits, identical")
                    assert hits > 100


def test_noalloc_rejects_launches_beyond_its_array_and_reports_overflow(ctx):
    import pyrtx
    _say("no-alloc limits and errors")
    sc = small_mixed_scene(seed=6, instances=16, freq=12, cells=16)
    blases, tlas = ctx.upload_scene(sc)
    rays = torch.from_numpy(np.ascontiguousarray(S.incoherent_rays(sc, 20000, seed=2))).cuda()
    out = torch.empty((rays.shape[0], 6), dtype=torch.int32, device=ctx.device)
    base = ctx.config()
    try:
        # 4096 stack entries per lane, 1 in LDS: ~8.6 GB of spill area at the default shape
        ctx.set_config(stack_limit=4096, lds_stack_depth=1)
        code = ctx.lib.rtx_cmd_trace_ray_buffer_noalloc(ctx.handle, tlas.address, rays.data_ptr(), rays.shape[0],
                                                        pyrtx.RTX_QUERY_CLOSEST_HIT, 0, 0xFF, out.data_ptr(), None)
        msg = ctx.lib.rtx_last_error_string(ctx.handle).decode()
        _say(f"  rejected: {code} {msg}")
        assert code == -2 and "bytes of scratch" in msg
        # a stack of one entry overflows on these rays: rtx_sync reports it from the array's flags
        ctx.set_config(stack_limit=1, lds_stack_depth=12)
        ctx._check(ctx.lib.rtx_cmd_trace_ray_buffer_noalloc(ctx.handle, tlas.address, rays.data_ptr(),
                                                            rays.shape[0], pyrtx.RTX_QUERY_CLOSEST_HIT, 0, 0xFF,
                                                            out.data_ptr(), None))
        with pytest.raises(pyrtx.RtxError) as e:
            ctx.sync()
        assert e.value.code == -1001
        ctx.sync()  # cleared
    finally:
        ctx.set_config(**dict(base, stack_limit=pyrtx.RTX_TRACE_STACK_LIMIT_DERIVED))
    ref, got = _both(ctx, tlas, rays, pyrtx.RTX_QUERY_CLOSEST_HIT)
    assert np.array_equal(ref, got)
