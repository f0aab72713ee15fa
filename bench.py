"""Ray-query throughput on the 20M-triangle scene (one MI355X).

Workload: 512 instances of 10 base meshes (5 seeded displaced geodesic icospheres of 0.51M-4.0M triangles and
5 noise terrains of 2.1M triangles; 19,994,340 unique triangles) placed uniformly in a 100^3 box with uniform
random rotations and scales in [0.5, 2] (seed 4242).  One step traces, as closest-hit queries, 16,777,216
coherent primary rays of a 4096x4096 pinhole camera at (0, 0, 150) looking down -z (60 deg, the reference
raygen model) and 16,777,216 incoherent rays (origins uniform in the scene box, directions uniform on the
sphere, tmin 0, tmax inf; seed 1234).  Rays, hits and the acceleration structures are resident in HBM before
the timed region.

The second half of the metric runs the path tracer on the closed unit-cube Cornell box with a displaced
level-8 icosphere (1,310,720 triangles + 24 box/light triangles): 1920x1080, depth 8, Russian roulette after
depth 3, seed 1234.  It times a plain 64-spp render, builds a 4096-spp FP64-accumulated reference, and measures
the time the neural-radiance-cache renderer (one training pass + one cache-terminated sample per frame) needs to
reach the plain render's relMSE = mean((x - ref)^2 / (ref^2 + 0.01)).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "build-up-phase_amd"))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s closest-hit (coherent/incoherent, 20M tris); equal-relMSE time plain vs NRC"
DEFAULT_CFG = {"block_size": 256, "blocks_per_cu": 4, "lds_stack_depth": 12, "ray_chunk": 1, "refill_lanes": 48,
               "refill_interval": 4, "leaf_lanes": 8}
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=float, default=1.0, help="triangle-count scale of the scene (1.0 = 20M)")
    ap.add_argument("--cpu-sample", type=int, default=1 << 22,
                    help="rays per set for the CPU baseline, every (W*H / n)th ray (0: skip the CPU baseline)")
    ap.add_argument("--parity-sample", type=int, default=1 << 20,
                    help="rays per set checked against the FP64 oracle (0: skip)")
    ap.add_argument("--no-anyhit", action="store_true")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracing / NRC equal-relMSE part")
    ap.add_argument("--no-autotune", action="store_true", help="use the default traversal configuration")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-scene part")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the ray queries on the dense path-tracing scene (every ray enters a BLAS)")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched part")
    ap.add_argument("--no-query", action="store_true",
                    help="skip the 20M-triangle ray-query part (profiling the path-tracing part alone)")
    ap.add_argument("--ref-spp", type=int, default=4096, help="samples of the FP64 reference render")
    ap.add_argument("--nrc-frames", type=int, default=160, help="NRC frames rendered (one sample each)")
    ap.add_argument("--nrc-kernels", action="store_true",
                    help="only the NRC network kernels on random inputs (inference of 2,073,600 queries, training "
                         "steps of 16,384 samples; for profiling)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pyrtx
    from pyrtx import scenes as S

    if args.nrc_kernels:
        ctx = pyrtx.Context(0)
        print(json.dumps({"nrc_kernels": nrc_kernel_bench(ctx, torch, pyrtx)}), flush=True)
        ctx.close()
        return

    if args.no_query:
        ctx = pyrtx.Context(0)
        render = None if args.no_render else render_bench(ctx, args, np, torch, pyrtx, S)
        ctx.set_config(**DEFAULT_CFG)
        batch = None if args.no_batch else batch_bench(ctx, args, np, torch, pyrtx, S)
        print(json.dumps({"metric": METRIC, "value": None, "unit": "Mrays/s", "n_gpus": 1, "path_tracing": render,
                          "batched_scenes": batch}), flush=True)
        ctx.close()
        return

    t0 = time.time()
    sc = S.query_scene(args.scale)
    n_tris = sc.unique_triangle_count
    log(f"scene generated in {time.time() - t0:.1f}s: {n_tris} unique triangles, {len(sc.blas_of_instance)} instances")
    W = H = 4096
    coh = S.camera_rays(sc.camera, W, H)
    inc = S.incoherent_rays(sc, W * H, seed=1234)
    log("rays generated")

    ctx = pyrtx.Context(0)
    dev = ctx.device
    torch.cuda.synchronize()
    tb = time.time()
    blases, tlas = ctx.upload_scene(sc)
    ctx.sync()
    build_s = time.time() - tb
    info = [ctx.as_info(b) for b in blases]
    tinfo = ctx.as_info(tlas)
    log(f"BLAS+TLAS built in {build_s * 1e3:.0f} ms (incl. upload); nodes {sum(i.node_count for i in info)}, "
        f"slots {sum(i.slot_count for i in info)}, max depth {tinfo.max_depth}")
    # the same builds again from the device-resident inputs (GPU time only): the BLASes one by one, then the TLAS
    eb0, eb1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    eb0.record()
    rebuilt = [ctx.build_blas(list(b.keepalive[1])) for b in blases]
    rebuilt.append(ctx.build_tlas(tlas.keepalive[1]))
    eb1.record()
    torch.cuda.synchronize()
    build_dev_ms = eb0.elapsed_time(eb1)
    del rebuilt
    log(f"BLAS+TLAS rebuilt from device-resident inputs in {build_dev_ms:.1f} ms")
    # time the BLAS builds alone (inputs already on the device)
    rc = torch.from_numpy(coh).to(dev)
    ri = torch.from_numpy(inc).to(dev)
    del coh
    hits = torch.empty((W * H, 6), dtype=torch.int32, device=dev)
    occl = torch.empty((W * H,), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def trace(rays, mode=pyrtx.RTX_QUERY_CLOSEST_HIT):
        return ctx.trace_ray_buffer(tlas, rays, mode, 0xFF, hits if mode == pyrtx.RTX_QUERY_CLOSEST_HIT else occl)

    default_cfg = ctx.config()
    # project autotuner: one traversal configuration per ray set and query type (untimed setup step)
    tuned = {}
    if not args.no_autotune:
        for name, r in (("coherent", rc), ("incoherent", ri)):
            for qn, mode in (("closest", pyrtx.RTX_QUERY_CLOSEST_HIT), ("any", pyrtx.RTX_QUERY_ANY_HIT)):
                if qn == "any" and args.no_anyhit:
                    continue
                ctx.set_config(**default_cfg)
                cfg, tms = ctx.autotune(tlas, r, mode, reps=2)
                tuned[f"{qn}/{name}"] = cfg
                log(f"autotuned {qn}/{name}: {cfg} ({tms:.3f} ms)")
    cfg_of = lambda key: tuned.get(key, default_cfg)

    def trace_cfg(key, rays, mode=pyrtx.RTX_QUERY_CLOSEST_HIT):
        ctx.set_config(**cfg_of(key))
        return trace(rays, mode)

    # traversal counters (separate, untimed pass, same configurations as the timed launches)
    stats = {}
    for name, r in (("coherent", rc), ("incoherent", ri)):
        ctx.set_config(**cfg_of(f"closest/{name}"))
        ctx.set_stats(True)
        trace(r)
        ctx.sync()
        s = ctx.stats()
        ctx.set_stats(False)
        n = max(1, s["rays"])
        stats[name] = {k: s[k] / n for k in ("node_fetches", "triangle_tests", "instance_entries")}
        hn = hits.cpu().numpy()
        stats[name]["hit_fraction"] = float((hn[:, 5].view(np.uint32) != 0xFFFFFFFF).mean())
    log(f"per-ray counters: {json.dumps(stats)}")

    def bytes_per_ray(s):
        # algorithmic bytes: ray 32 + hit 24 + 64 per node, 48 per triangle, 64 per instance entry
        return 32 + 24 + 64 * s["node_fetches"] + 48 * s["triangle_tests"] + 64 * s["instance_entries"]

    if args.calib_copy:
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("calibration copy: 268435456 bytes read + 268435456 bytes written")
    for _ in range(args.warmup):
        trace_cfg("closest/coherent", rc)
        trace_cfg("closest/incoherent", ri)
    ctx.sync()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(4 * args.steps)]
    log(f"timing {args.steps} steps of 2 x {W * H} rays")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    k = 0
    for step in range(args.steps):
        for key, r in (("closest/coherent", rc), ("closest/incoherent", ri)):
            ctx.set_config(**cfg_of(key))
            ev[k][0].record()
            trace(r)
            ev[k][1].record()
            k += 1
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ctx.sync()
    ms = [a.elapsed_time(b) for a, b in ev[:k]]
    ms_coh = float(np.mean(ms[0::2]))
    ms_inc = float(np.mean(ms[1::2]))
    rays_per_step = 2 * W * H
    value = rays_per_step * args.steps / t_total / 1e6
    mr_coh = W * H / ms_coh / 1e3
    mr_inc = W * H / ms_inc / 1e3
    log(f"closest-hit: coherent {mr_coh:.1f} Mrays/s ({ms_coh:.2f} ms), incoherent {mr_inc:.1f} Mrays/s "
        f"({ms_inc:.2f} ms), whole-step {value:.1f} Mrays/s")

    any_res = {}
    if not args.no_anyhit:
        for name, r in (("coherent", rc), ("incoherent", ri)):
            ctx.set_config(**cfg_of(f"any/{name}"))
            trace(r, pyrtx.RTX_QUERY_ANY_HIT)
            ctx.sync()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(3):
                trace(r, pyrtx.RTX_QUERY_ANY_HIT)
            b.record()
            torch.cuda.synchronize()
            any_res[name] = W * H * 3 / a.elapsed_time(b) / 1e3
        log(f"any-hit: {any_res}")

    # tuned shapes of 5 waves per SIMD run the 96-VGPR kernel variant, which spills a few registers to scratch: time
    # each such shape against the same configuration at 4 waves per SIMD (the spill-free 128-VGPR variant)
    variant_check = {}
    for key, cfg in tuned.items():
        if (cfg["block_size"] // 64) * cfg["blocks_per_cu"] // 4 < 5:
            continue
        r = rc if key.endswith("coherent") and not key.endswith("incoherent") else ri
        mode = pyrtx.RTX_QUERY_CLOSEST_HIT if key.startswith("closest") else pyrtx.RTX_QUERY_ANY_HIT
        res = {}
        for tag, c in (("w5", cfg), ("w4", dict(cfg, blocks_per_cu=16 // (cfg["block_size"] // 64)))):
            ctx.set_config(**c)
            trace(r, mode)
            ctx.sync()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(3):
                trace(r, mode)
            b.record()
            torch.cuda.synchronize()
            res[f"{tag}_ms"] = round(a.elapsed_time(b) / 3, 3)
            res[f"{tag}_blocks_per_cu"] = c["blocks_per_cu"]
        variant_check[key] = res
    if variant_check:
        log(f"96- vs 128-VGPR variant at the tuned shapes: {json.dumps(variant_check)}")

    # the box's own streaming rate: a 1 GiB device-to-device copy (read + write), best of 3, as context for the
    # published peak
    src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    copy_ms = []
    for _ in range(3):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        copy_ms.append(a.elapsed_time(b))
    copy_gbs = 2 * src.numel() * 4 / (min(copy_ms) * 1e-3) / 1e9
    del src, dst

    # roofline of the traversal kernel (both launches of a step)
    b_coh, b_inc = bytes_per_ray(stats["coherent"]), bytes_per_ray(stats["incoherent"])
    achieved = (b_coh * W * H + b_inc * W * H) / ((ms_coh + ms_inc) * 1e-3) / 1e9
    traffic = measured_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if traffic is None else round((traffic["coherent"] + traffic["incoherent"]) / 2),
                "traffic_bytes_per_launch": None if traffic is None else
                {k: traffic[k] for k in ("coherent", "incoherent")},
                "traffic_source": None if traffic is None else traffic["source"],
                "kernel": "trace_kernel<closest> (persistent BVH8 traversal)",
                "bytes_per_ray": {"coherent": round(b_coh, 1), "incoherent": round(b_inc, 1)},
                "launch_ms": {"coherent": round(ms_coh, 3), "incoherent": round(ms_inc, 3)},
                "measured_copy_gbs": round(copy_gbs, 1)}

    cpu = None
    parity = {}
    cores = min(os.cpu_count() or 1, 16)
    if args.cpu_sample > 0:
        cpu = cpu_baseline_rays(args, sc, inc, W, H, cores, np, S)
    if args.parity_sample > 0:
        import oracle as O
        O.reset()
        tb = time.time()
        handles = [O.blas_create([(g.vertices, g.indices, g.transform) for g in geoms]) for geoms in sc.blas]
        ot = O.tlas_create(O.instance_records(sc.transforms, [handles[b] for b in sc.blas_of_instance],
                                              sc.custom_index, sc.mask, sc.sbt_offset, sc.flags))
        log(f"oracle BVH built in {time.time() - tb:.1f}s; full-size parity sample")
        # full-size parity: the GPU's hits from the timed configurations vs the FP64 oracle on every (W*H / n)th
        # ray of each set; epsilon-tie rays are checked against their candidate sets
        step = max(1, (W * H) // args.parity_sample)
        for name, r, rays in (("coherent", rc, S.camera_rays(sc.camera, W, H)[::step]),
                              ("incoherent", ri, inc[::step])):
            rays = rays[: args.parity_sample]
            o = O.trace(ot, rays, threads=cores)
            trace_cfg(f"closest/{name}", r)
            ctx.sync()
            idx = torch.arange(0, W * H, step, device=dev)[: args.parity_sample]
            parity[name] = hit_parity(hits[idx].cpu().numpy(), *o, tlas=ot, rays=rays)
        O.reset()
        log(f"parity vs the FP64 oracle: {json.dumps(parity)}")

    del rc, ri, hits, occl
    torch.cuda.empty_cache()
    ctx.set_config(**default_cfg)
    dense = None if args.no_dense else dense_query_bench(ctx, args, np, torch, pyrtx, S, cores)
    ctx.set_config(**default_cfg)
    render = None if args.no_render else render_bench(ctx, args, np, torch, pyrtx, S)
    ctx.set_config(**default_cfg)
    batch = None if args.no_batch else batch_bench(ctx, args, np, torch, pyrtx, S)
    nrc_roof = None if args.no_render else nrc_roofline(nrc_kernel_bench(ctx, torch, pyrtx))

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "20M-triangle ray-query scene (512 instances of 10 meshes in a 100^3 box); one step = "
                               "16,777,216 coherent + 16,777,216 incoherent closest-hit rays",
                   "unique_triangles": n_tris, "instances": int(len(sc.blas_of_instance)),
                   "rays_per_step": rays_per_step, "camera": list(sc.camera), "image": [W, H]},
        "closest_hit_mrays": {"coherent": round(mr_coh, 1), "incoherent": round(mr_inc, 1)},
        "any_hit_mrays": {k: round(v, 1) for k, v in any_res.items()},
        "traversal_counters_per_ray": stats,
        "autotuned_trace_config": tuned or None,
        "variant_96_vs_128_vgpr": variant_check or None,
        "dense_scene_queries": dense,
        "build": {"blas_tlas_ms_incl_upload": round(build_s * 1e3, 1), "blas_tlas_ms_device": round(build_dev_ms, 1),
                  "nodes": int(sum(i.node_count for i in info)), "max_depth": int(tinfo.max_depth)},
        "roofline": roofline,
        "roofline_nrc": nrc_roof,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity or None,
        "path_tracing": render,
        "batched_scenes": batch,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def hit_parity(h, o_tuv, o_ids, o_st, rel_tol=1e-5, tlas=None, rays=None):
    """Closest-hit parity of GPU hit records h (int32 [N, 6]: t, u, v bits, prim, geom, inst) against the FP64
    oracle's (t_uv, ids, status) for the same rays: IDs and hit/miss on every ray the oracle calls robust
    (status 0 miss, 1 hit), t relative and u, v absolute errors; epsilon-tie rays (status 2) against their candidate
    sets when the oracle TLAS and the rays are given (a hit outside the set, a miss where a candidate is robustly
    inside, a t off its candidate's)."""
    import numpy as np
    f = np.ascontiguousarray(h[:, :3]).view(np.float32).astype(np.float64)
    ids = np.ascontiguousarray(h[:, 3:]).view(np.uint32)
    hit, miss = o_st == 1, o_st == 0
    id_bad = hit & np.any(ids != o_ids[:, :3], axis=1)
    ok = hit & ~id_bad
    t_err = np.abs(f[ok, 0] - o_tuv[ok, 0]) / np.maximum(np.abs(o_tuv[ok, 0]), 1e-30)
    uv_err = np.abs(f[ok, 1:] - o_tuv[ok, 1:]).max(axis=1) if ok.any() else np.zeros(0)
    out = {"rays": int(len(o_st)), "ties": int((o_st == 2).sum()), "robust_hits": int(hit.sum()),
           "id_mismatch": int(id_bad.sum()), "miss_mismatch": int((miss & (ids[:, 2] != 0xFFFFFFFF)).sum()),
           "t_over_tol": int((t_err > rel_tol).sum()), "uv_over_tol": int((uv_err > rel_tol).sum()),
           "max_t_rel": float(t_err.max()) if t_err.size else 0.0,
           "max_uv_abs": float(uv_err.max()) if uv_err.size else 0.0}
    tie = o_st == 2
    if tlas is not None and tie.any():
        import oracle as O
        r = np.asarray(rays, np.float32).reshape(-1, 8)[tie]
        c = O.trace_candidates(tlas, r)
        tr = O.check_ties(r, c, ids[tie, 2], ids[tie, 1], ids[tie, 0], f[tie, 0])
        out["ties_checked"] = {k: tr[k] for k in ("hit_not_candidate", "miss_not_allowed", "t_bad", "unchecked")}
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        return ""


def cpu_baseline_rays(args, sc, inc, W, H, cores, np, S):
    """The reference's ray tracing on the host cores (oracle/cpu_baseline.cpp: FP32 raygen / closest-hit / miss over
    binned-SAH BVHs with the watertight test, std::thread x cores) on the same scene and a bounded sample of the same
    rays: median of 5 timed runs after a warm-up; value_1core: every 16th ray of the sample on one thread."""
    import oracle as O
    tb = time.time()
    cb = O.CpuBaseline(sc)
    build_s = time.time() - tb
    step = max(1, (W * H) // args.cpu_sample)
    coh_s = S.camera_rays(sc.camera, W, H)[::step][: args.cpu_sample]
    inc_s = inc[::step][: args.cpu_sample]
    cb.trace(coh_s[:65536], threads=cores)
    runs = {"coherent": [], "incoherent": []}
    for _ in range(5):
        for name, r in (("coherent", coh_s), ("incoherent", inc_s)):
            t0 = time.perf_counter()
            cb.trace(r, threads=cores)
            runs[name].append(time.perf_counter() - t0)
    med = {k: float(np.median(v)) for k, v in runs.items()}
    t1 = time.perf_counter()
    cb.trace(coh_s[::16], threads=1)
    cb.trace(inc_s[::16], threads=1)
    dt1 = time.perf_counter() - t1
    n = len(coh_s) + len(inc_s)
    out = {"value": round(n / (med["coherent"] + med["incoherent"]) / 1e6, 3), "unit": "Mrays/s", "cores": cores,
           "kind": "port",
           "closest_hit_mrays": {k: round(len(coh_s) / v / 1e6, 3) for k, v in med.items()},
           "value_1core": round((len(coh_s[::16]) + len(inc_s[::16])) / dt1 / 1e6, 4),
           "build_s": round(build_s, 2), "host_cpus": os.cpu_count(), "cpu_model": _cpu_model(),
           "sample": f"{len(coh_s)} coherent + {len(inc_s)} incoherent closest-hit rays (every {step}th ray of each "
                     f"set) on the 20M-triangle scene, traced by oracle/cpu_baseline.cpp (FP32 raygen / chit / miss of "
                     f"main.cpp:1033-1091, binned-SAH BVH2, watertight test, -O3 x86-64-v3, std::thread x {cores}); "
                     f"median of 5 runs; value_1core: every 16th of those rays on one thread"}
    log(f"cpu baseline {out['value']} Mrays/s on {cores} threads ({out['closest_hit_mrays']}), "
        f"{out['value_1core']} on one; CPU BVH build {build_s:.1f}s")
    return out


def nrc_kernel_bench(ctx, torch, pyrtx):
    """The NRC network's kernels alone on random inputs: inference of one 1920x1080 frame of cache queries and
    training on 65,536 samples (4 Adam steps of 16,384).  A warm-up round, then 10 back-to-back calls of each between
    two HIP events, so that the host's per-call enqueue overlaps the previous call's kernels; the time per call."""
    nrc = pyrtx.Nrc(ctx, (0, 0, 0), (1, 1, 1), seed=1, batch_size=16384)
    g = torch.Generator(device=ctx.device).manual_seed(1)
    q = torch.rand((2073600, 12), device=ctx.device, generator=g)
    smp = torch.rand((65536, 16), device=ctx.device, generator=g)
    out = {}
    reps = 10
    for name, fn in (("infer_2073600_ms", lambda: nrc.infer(q)), ("train_4x16384_ms", lambda: nrc.train(smp))):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        out[name] = round(a.elapsed_time(b) / reps, 4)
    nrc.close()
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0
NRC_SEEDS = (1, 2, 3, 4, 5)  # cache initialisations of the equal-relMSE comparison (median reported)
NRC_WARMUP = 2               # first frames of an NRC run: timed, training only (spp 0), nothing accumulated
NRC_FLOP_PER_QUERY = 47104    # forward: 2 x sum(fan_in x fan_out) over the 6 layers
NRC_FLOP_PER_SAMPLE = 141312  # forward + data gradient + weight gradient


def nrc_roofline(k):
    """MFMA roofline of the cache network's kernels from nrc_kernel_bench's HIP-event times: inference of 2,073,600
    queries in one launch, and 4 training steps of 16,384 samples (the training kernel plus the Adam kernel that sums
    the block partials and updates the weights, per step)."""
    inf_tf = 2073600 * NRC_FLOP_PER_QUERY / (k["infer_2073600_ms"] * 1e-3) / 1e12
    tr_tf = 65536 * NRC_FLOP_PER_SAMPLE / (k["train_4x16384_ms"] * 1e-3) / 1e12
    return {"infer": {"bound": "mfma", "achieved": round(inf_tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": round(inf_tf / MFMA_BF16_PEAK_TFLOPS, 4),
                      "kernel": "nrc_infer_lds_kernel", "units": "2,073,600 queries x 47,104 FLOP",
                      "launch_ms": k["infer_2073600_ms"]},
            "train": {"bound": "mfma", "achieved": round(tr_tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": round(tr_tf / MFMA_BF16_PEAK_TFLOPS, 4),
                      "kernel": "nrc_train_kernel + nrc_adam_kernel", "units": "4 x 16,384 samples x 141,312 FLOP",
                      "ms": k["train_4x16384_ms"]}}


def measured_traffic():
    """HBM bytes per traversal launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-calibrated) as recorded in the
    committed profile of this kernel and workload; None when the profile is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "trace_kernel_hbm_traffic_20M_query_scene.md")
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"bench-traffic-bytes-per-launch: coherent (\d+) incoherent (\d+)", txt)
    if not m:
        return None
    c = re.search(r"measured at commit ([0-9a-f]{7,40})", txt, re.I)
    return {"coherent": int(m.group(1)), "incoherent": int(m.group(2)),
            "source": "profiles/trace_kernel_hbm_traffic_20M_query_scene.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, "
                      "copy-calibrated" + (f", measured at commit {c.group(1)})" if c else ")")}


def dense_query_bench(ctx, args, np, torch, pyrtx, S, cores):
    """Closest- and any-hit ray queries where every ray traverses a BLAS: the closed Cornell box with the displaced
    level-8 icosphere (1,310,744 triangles), 4096 x 4096 camera rays from the scene's camera and 16,777,216
    incoherent rays (origins uniform in the box, directions uniform on the sphere).  Each set and query type is
    autotuned, timed over 5 launches, counted (node fetches, triangle tests per ray: HBM roofline of the launch) and
    checked against the FP64 oracle on every 64th ray (epsilon ties against their candidate sets)."""
    W = H = 4096
    sc, _ = S.cornell_scene(8, width=W, height=H)
    dev = ctx.device
    blases, tlas = ctx.upload_scene(sc)
    ctx.sync()
    sets = {"coherent": S.camera_rays(sc.camera, W, H), "incoherent": S.incoherent_rays(sc, W * H, seed=77)}
    hits = torch.empty((W * H, 6), dtype=torch.int32, device=dev)
    occl = torch.empty((W * H,), dtype=torch.int32, device=dev)
    default_cfg = ctx.config()
    out = {"workload": "closed Cornell box + displaced level-8 icosphere (1,310,744 triangles); 16,777,216 camera "
                       "rays (4096 x 4096) and 16,777,216 incoherent rays per launch",
           "closest_hit_mrays": {}, "any_hit_mrays": {}, "traversal_counters_per_ray": {}, "roofline": {},
           "autotuned_trace_config": {}, "parity_vs_oracle": {}}
    ot = None
    if args.parity_sample > 0:
        import oracle as O
        O.reset()
        handles = [O.blas_create([(g.vertices, g.indices, g.transform) for g in geoms]) for geoms in sc.blas]
        ot = O.tlas_create(O.instance_records(sc.transforms, [handles[b] for b in sc.blas_of_instance],
                                              sc.custom_index, sc.mask, sc.sbt_offset, sc.flags))
    for name, rays in sets.items():
        r = torch.from_numpy(rays).to(dev)
        for qn, mode, res in (("closest", pyrtx.RTX_QUERY_CLOSEST_HIT, hits), ("any", pyrtx.RTX_QUERY_ANY_HIT, occl)):
            ctx.set_config(**default_cfg)
            if not args.no_autotune:
                cfg, _ = ctx.autotune(tlas, r, mode, reps=2)
                out["autotuned_trace_config"][f"{qn}/{name}"] = cfg
            if qn == "closest":
                ctx.set_stats(True)
                ctx.trace_ray_buffer(tlas, r, mode, 0xFF, res)
                ctx.sync()
                st = ctx.stats()
                ctx.set_stats(False)
                n = max(1, st["rays"])
                cnt = {k: st[k] / n for k in ("node_fetches", "triangle_tests", "instance_entries")}
                cnt["hit_fraction"] = float((hits[:, 5].cpu().numpy().view(np.uint32) != 0xFFFFFFFF).mean())
                out["traversal_counters_per_ray"][name] = cnt
            ctx.trace_ray_buffer(tlas, r, mode, 0xFF, res)
            ctx.sync()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(5):
                ctx.trace_ray_buffer(tlas, r, mode, 0xFF, res)
            b.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / 5
            out[f"{qn}_hit_mrays"][name] = round(W * H / ms / 1e3, 1)
            if qn == "closest":
                c = out["traversal_counters_per_ray"][name]
                bpr = 32 + 24 + 64 * c["node_fetches"] + 48 * c["triangle_tests"] + 64 * c["instance_entries"]
                gbs = bpr * W * H / (ms * 1e-3) / 1e9
                out["roofline"][name] = {"bound": "hbm", "bytes_per_ray": round(bpr, 1), "launch_ms": round(ms, 3),
                                         "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": round(gbs / HBM_PEAK_GBS, 4)}
                if ot is not None:
                    sample = rays[::64]
                    o = O.trace(ot, sample, threads=cores)
                    out["parity_vs_oracle"][name] = hit_parity(hits[::64].cpu().numpy(), *o, tlas=ot, rays=sample)
        del r
        log(f"dense scene {name}: closest {out['closest_hit_mrays'][name]} / any {out['any_hit_mrays'][name]} "
            f"Mrays/s, counters {out['traversal_counters_per_ray'][name]}, parity {out['parity_vs_oracle'].get(name)}")
    if ot is not None:
        O.reset()
    ctx.sync()
    del hits, occl, blases, tlas
    torch.cuda.empty_cache()
    return out


def render_bench(ctx, args, np, torch, pyrtx, S):
    """Plain 64-spp path tracing vs the neural radiance cache at equal relMSE (1920x1080 Cornell box)."""
    W, H, SPP = 1920, 1080, 64
    sc, lights = S.cornell_scene(8, width=W, height=H)
    dev = ctx.device
    blases, tlas = ctx.upload_scene(sc)
    rec = torch.from_numpy(np.ascontiguousarray(sc.records).view(np.uint8).reshape(-1).copy()).to(dev)
    lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).reshape(-1, 12).copy()).to(dev)
    ctx.sync()
    log(f"path tracing scene: {sc.unique_triangle_count} triangles, {W}x{H}")

    def ev():
        return torch.cuda.Event(enable_timing=True)

    def relmse(x, ref):
        return torch.mean((x - ref) ** 2 / (ref ** 2 + 0.01))

    tuned = None
    if not args.no_autotune:
        # tuned on renders of 4 samples, which run as one whole-path launch of 4 views of the camera like the
        # timed render's groups of 16 (1-sample renders favour shapes that are ~3% slower on those launches)
        tuned, tms = ctx.autotune_render(tlas, rec, sc.camera, lt, W, H, spp=4)
        log(f"autotuned path tracing: {tuned} ({tms:.2f} ms per 4-spp render)")
    ctx.render(tlas, rec, sc.camera, lt, W, H, spp=2, sample_offset=1 << 24)  # warm-up
    ctx.sync()
    r0 = ctx.render_ray_count()
    a, b = ev(), ev()
    a.record()
    film = ctx.render(tlas, rec, sc.camera, lt, W, H, spp=SPP)
    b.record()
    torch.cuda.synchronize()
    ms_plain = a.elapsed_time(b)
    n_rays = ctx.render_ray_count() - r0
    log(f"plain {SPP} spp: {ms_plain:.1f} ms, {n_rays} rays, {n_rays / ms_plain / 1e3:.1f} Mrays/s")

    acc = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
    chunks = max(1, args.ref_spp // 64)
    tr = time.time()
    for k in range(chunks):
        ctx.render(tlas, rec, sc.camera, lt, W, H, spp=64, sample_offset=(1 << 20) + 64 * k, accum64=acc)
        if (k + 1) % 16 == 0:
            ctx.sync()
            log(f"reference {64 * (k + 1)}/{64 * chunks} spp ({time.time() - tr:.0f}s)")
    ctx.sync()
    ref = (acc / (64 * chunks)).float()
    err_plain = float(relmse(film[..., :3], ref))
    log(f"plain relMSE {err_plain:.5f} vs {64 * chunks}-spp reference")

    cpu_pt = None
    if args.cpu_sample > 0:
        # the same estimator on the host cores (oracle/cpu_baseline.cpp, FP32): 1-spp renders of the full film,
        # median of 3; the 64-spp time is 64 of them
        import oracle as O
        cores = min(os.cpu_count() or 1, 16)
        tb = time.time()
        cb = O.CpuBaseline(sc)
        cb_build = time.time() - tb
        cb.render(sc.records, sc.camera, lights, 192, 108, 1, threads=cores)
        ts, nr = [], 0
        for k in range(3):
            t0 = time.perf_counter()
            _, nr = cb.render(sc.records, sc.camera, lights, W, H, 1, sample_offset=k, threads=cores)
            ts.append(time.perf_counter() - t0)
        tcpu = float(np.median(ts))
        cpu_pt = {"ms_per_spp": round(tcpu * 1e3, 1), "mrays_per_s": round(nr / tcpu / 1e6, 2),
                  "ms_64spp": round(64 * tcpu * 1e3, 0), "cores": cores, "build_s": round(cb_build, 2),
                  "sample": f"{W}x{H} at 1 spp (median of 3 renders; 64 spp = 64 x that), oracle/cpu_baseline.cpp "
                            f"FP32 estimator, std::thread x {cores}"}
        del cb
        log(f"cpu path tracing: {cpu_pt['ms_per_spp']} ms per spp ({cpu_pt['mrays_per_s']} Mrays/s on {cores} "
            f"threads), 64 spp ~{cpu_pt['ms_64spp']:.0f} ms")

    lo, hi = S.scene_bounds(sc)
    # The NRC frame's launches (a render wavefront to the cache vertex beside the training paths' whole-path
    # launch) want their own launch shape: a few shapes timed on a scratch cache (untimed setup, like the other
    # autotuning); the plain render keeps its own.
    nrc_cfg = ctx.config()
    if not args.no_autotune:
        base = ctx.config()
        best = None
        scratch_nrc = pyrtx.Nrc(ctx, lo, hi, seed=2, batch_size=16384, cache_vertex=1)
        for bs, bpc in ((256, 4), (256, 5), (128, 8), (128, 10)):
            for leaf, refill, interval, chunk in ((1, 48, 4, 128), (8, 48, 4, 128), (1, 32, 8, 1), (8, 32, 8, 1)):
                cfg = dict(base, block_size=bs, blocks_per_cu=bpc, leaf_lanes=leaf, refill_lanes=refill,
                           refill_interval=interval, ray_chunk=chunk)
                ctx.set_config(**cfg)
                ts = []
                for k in range(3):
                    a0, b0 = ev(), ev()
                    a0.record()
                    ctx.render_nrc(tlas, rec, sc.camera, lt, scratch_nrc, W, H, spp=1, train_paths=65536,
                                   sample_offset=(1 << 22) + k)
                    b0.record()
                    torch.cuda.synchronize()
                    ts.append(a0.elapsed_time(b0))
                t = sorted(ts)[1]
                if best is None or t < best[0]:
                    best = (t, cfg)
        scratch_nrc.close()
        nrc_cfg = best[1]
        log(f"autotuned NRC frame: {nrc_cfg} ({best[0]:.2f} ms per frame)")
    ctx.set_config(**nrc_cfg)
    # Equal-relMSE time over several cache initialisations: the first NRC_WARMUP frames (untrained cache) only
    # train (spp 0: their training paths and Adam steps, exactly as in a rendered frame, no pixel paths) and are
    # timed, then rendered frames are averaged with equal weights.  Averaging frames of the untrained cache leaves a
    # bias that depends on the initialisation (2.1x - 4.1x over 8 seeds, median 2.5x, against 3.4x - 3.8x with two
    # warm-up frames, at the 324-ms plain render of that time); a warm-up frame that renders costs 2.9 ms instead of
    # 1.8 ms and its image would be discarded.
    accn = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
    film_nrc = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    per_seed = []
    for si, seed in enumerate(NRC_SEEDS):
        frames = args.nrc_frames if si == 0 else min(args.nrc_frames, 64)
        nrc = pyrtx.Nrc(ctx, lo, hi, seed=seed, batch_size=16384, cache_vertex=1)
        accn.zero_()
        errs = torch.full((frames,), float("inf"), dtype=torch.float32, device=dev)
        evs = [(ev(), ev()) for _ in range(frames)]
        torch.cuda.synchronize()
        tw = time.perf_counter()
        for fr in range(frames):
            acc_fr = accn if fr >= NRC_WARMUP else None
            evs[fr][0].record()
            ctx.render_nrc(tlas, rec, sc.camera, lt, nrc, W, H, spp=1 if acc_fr is not None else 0,
                           train_paths=65536, sample_offset=fr, film=film_nrc, accum64=acc_fr)
            evs[fr][1].record()
            if acc_fr is not None:
                errs[fr] = relmse((accn / (fr + 1 - NRC_WARMUP)).float(), ref)
        torch.cuda.synchronize()
        wall_nrc = time.perf_counter() - tw
        fms = np.array([x.elapsed_time(y) for x, y in evs])
        e = errs.cpu().numpy()
        cum = np.cumsum(fms)
        hit = np.nonzero(e <= err_plain)[0]
        f = int(hit[0]) if len(hit) else None
        ms_nrc = float(cum[f]) if f is not None else None
        steps, loss = nrc.status()
        nrc.close()
        per_seed.append({"seed": seed, "frames": frames, "frame_ms": round(float(fms[NRC_WARMUP:].mean()), 3),
                         "warmup_frame_ms": round(float(fms[:NRC_WARMUP].mean()), 3),
                         "frames_to_plain_relmse": None if f is None else f + 1,
                         "ms_to_plain_relmse": None if ms_nrc is None else round(ms_nrc, 2),
                         "speedup": None if ms_nrc is None else round(ms_plain / ms_nrc, 2),
                         "relmse_final": round(float(e[-1]), 6), "adam_steps": int(steps), "wall_s": round(wall_nrc, 2)})
        log(f"NRC seed {seed}: frame {fms[NRC_WARMUP:].mean():.2f} ms (training-only warm-up frames "
            f"{fms[:NRC_WARMUP].mean():.2f} ms), relMSE after {frames} frames {e[-1]:.5f}; reaches plain "
            f"relMSE at frame {None if f is None else f + 1} ({ms_nrc} ms)")
    sp = np.array([r["speedup"] or 0.0 for r in per_seed])
    first = per_seed[0]
    return {
        "workload": f"closed Cornell box + displaced level-8 icosphere ({sc.unique_triangle_count} triangles), "
                    f"{W}x{H}, depth 8, RR after 3, seed 1234",
        "plain": {"spp": SPP, "ms": round(ms_plain, 2), "rays": int(n_rays),
                  "mrays_per_s": round(n_rays / ms_plain / 1e3, 1), "relmse": round(err_plain, 6)},
        "reference_spp": 64 * chunks,
        "nrc": {"frames": args.nrc_frames, "train_paths_per_frame": 65536, "batch_size": 16384,
                "cache_vertex": 1, "frame_ms": first["frame_ms"], "autotuned_trace_config": nrc_cfg,
                "warmup_frames_not_accumulated": NRC_WARMUP, "warmup_frames_train_only": True,
                "warmup_frame_ms": first["warmup_frame_ms"],
                "relmse_final": first["relmse_final"], "adam_steps": first["adam_steps"],
                "frames_to_plain_relmse": first["frames_to_plain_relmse"],
                "ms_to_plain_relmse": first["ms_to_plain_relmse"], "per_seed": per_seed},
        "equal_relmse_speedup": round(float(np.median(sp)), 2),
        "equal_relmse_speedup_range": [round(float(sp.min()), 2), round(float(sp.max()), 2)],
        "autotuned_trace_config": tuned,
        "cpu_baseline": cpu_pt,
    }


def batch_bench(ctx, args, np, torch, pyrtx, S):
    """The batched-scenes workload: uneven scenes built by one batched BLAS build and rendered by one wavefront replayed from a
    HIP graph, against the same scenes built and rendered one at a time."""
    n, spp, depth = args.batch_scenes, 16, 6
    t0 = time.time()
    sc, views, lights = S.batch_scenes(count=n, seed=7)
    tris = sum(v["triangles"] for v in views)
    npix = sum(v["width"] * v["height"] for v in views)
    log(f"batch: {n} scenes, {tris} triangles, {npix} pixels (generated in {time.time() - t0:.1f}s)")
    dev = ctx.device

    def ev():
        return torch.cuda.Event(enable_timing=True)

    lists = []
    for geoms in sc.blas:
        lists.append([{"vertices": torch.from_numpy(g.vertices.astype(np.float32)).to(dev),
                       "indices": torch.from_numpy(g.indices.astype(np.uint32).view(np.int32)).to(dev),
                       "transform": None} for g in geoms])
    torch.cuda.synchronize()
    ctx.build_blas_batch(lists)  # warm-up (allocations)
    ctx.sync()
    a, b = ev(), ev()
    a.record()
    blases = ctx.build_blas_batch(lists)
    b.record()
    torch.cuda.synchronize()
    ms_build_batch = a.elapsed_time(b)
    a.record()
    singles = [ctx.build_blas(g) for g in lists]
    b.record()
    torch.cuda.synchronize()
    ms_build_seq = a.elapsed_time(b)
    del singles
    inst = pyrtx.instance_records(sc, [x.address for x in blases])
    tlas = ctx.build_tlas(inst)
    rec = torch.from_numpy(np.ascontiguousarray(sc.records).view(np.uint8).reshape(-1).copy()).to(dev)
    lt = torch.from_numpy(np.ascontiguousarray(lights, np.float32).copy()).to(dev)
    ctx.sync()
    log(f"batch BLAS builds: batched {ms_build_batch:.1f} ms, one by one {ms_build_seq:.1f} ms")
    tuned, width, width_ms = None, 8, {}
    if not args.no_autotune:
        # autotuner: BVH width (a build-time choice: the batch rebuilt at each width) x the traversal launch
        # configuration (rtx_autotune_render_batch on each build); the fastest 1-spp batch render wins
        best = None
        base_cfg = ctx.config()
        keep = [blases, tlas]  # every width's structures stay alive (inst below refers to the width-8 BLASes)
        for w in (4, 6, 8):
            flags = pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT | (pyrtx.RTX_BUILD_MAX_WIDTH(w) if w < 8 else 0)
            bw = ctx.build_blas_batch(lists, flags) if w < 8 else blases
            tw = ctx.build_tlas(pyrtx.instance_records(sc, [x.address for x in bw]), flags) if w < 8 else tlas
            ctx.set_config(**base_cfg)
            cfg, tms = ctx.autotune_render_batch(tw, rec, views, lt, spp=1, max_depth=depth)
            keep += [bw, tw]
            width_ms[w] = round(tms, 3)
            log(f"autotuned batched render at width {w}: {cfg} ({tms:.2f} ms per 1-spp batch)")
            if best is None or tms < best[0]:
                best = (tms, w, cfg, bw, tw)
        _, width, tuned, blases, tlas = best
        ctx.set_config(**tuned)
        log(f"batched render: width {width}")
    films = ctx.render_batch(tlas, rec, views, lt, spp=spp, max_depth=depth, sample_offset=1 << 20, graph=True)
    ctx.sync()
    reps = 3
    r0 = ctx.render_ray_count()
    a.record()
    for k in range(reps):
        ctx.render_batch(tlas, rec, views, lt, spp=spp, max_depth=depth, sample_offset=spp * k, graph=True,
                         films=films)
    b.record()
    torch.cuda.synchronize()
    ms_graph = a.elapsed_time(b) / reps
    rays = (ctx.render_ray_count() - r0) // reps
    a.record()
    ctx.render_batch(tlas, rec, views, lt, spp=spp, max_depth=depth, films=films)
    b.record()
    torch.cuda.synchronize()
    ms_direct = a.elapsed_time(b)
    # the same views one at a time, each scene in a TLAS of its own
    tl1 = [ctx.build_tlas(inst[k:k + 1]) for k in range(n)]
    outs = [torch.empty((v["height"], v["width"], 4), dtype=torch.float32, device=dev) for v in views]
    ctx.sync()
    a.record()
    for k, v in enumerate(views):
        ctx.render(tl1[k], rec, v["camera"], lt[v["light_first"]:v["light_first"] + v["light_count"]], v["width"],
                   v["height"], spp=spp, max_depth=depth, film=outs[k])
    b.record()
    torch.cuda.synchronize()
    ms_seq = a.elapsed_time(b)
    log(f"batch render {spp} spp: graph {ms_graph:.1f} ms ({rays / ms_graph / 1e3:.0f} Mrays/s), direct "
        f"{ms_direct:.1f} ms, one view at a time {ms_seq:.1f} ms")
    recorded = recorded_frame_bench(ctx, lists, sc, rec, views, lt, spp, depth, width, tuned, np, torch, pyrtx)
    return {"workload": f"{n} scenes (seed 7): displaced icosphere + ground quad + area light each, triangles "
                        f"log-uniform in [1e3, 2e6], films from 256^2 / 512^2 / 1024x768; {spp} spp, depth {depth}",
            "scenes": n, "triangles": int(tris), "pixels": int(npix),
            "build_ms": {"batched": round(ms_build_batch, 2), "one_by_one": round(ms_build_seq, 2)},
            "render_ms": {"batched_graph": round(ms_graph, 1), "batched_direct": round(ms_direct, 1),
                          "one_view_at_a_time": round(ms_seq, 1)},
            "rays_per_frame": int(rays), "mrays_per_s": round(rays / ms_graph / 1e3, 1),
            "autotuned_trace_config": tuned, "autotuned_bvh_width": width, "ms_per_1spp_by_width": width_ms or None,
            "recorded_build_and_render": recorded}


def recorded_frame_bench(ctx, lists, sc, rec, views, lt, spp, depth, width, tuned, np, torch, pyrtx):
    """The batch's whole frame - batched BLAS build, TLAS build, batched render - recorded by a capture of a
    context's stream into one HIP graph (the command-buffer form of the reference's recorded build and trace) and
    replayed; each replay rebuilds both levels from the device data.  A TLAS built inside the recording cannot be
    read back on the host, so the render is recorded in both integrator forms, each guarded by the content bits the
    device reads from the TLAS at replay.  The replays are checked bit-identical to the same calls made directly."""
    s = torch.cuda.Stream()
    c = pyrtx.Context(0, stream=s)
    try:
        if tuned:
            c.set_config(**tuned)
        flags = pyrtx.RTX_BUILD_PREFER_FAST_TRACE_BIT | (pyrtx.RTX_BUILD_MAX_WIDTH(width) if width < 8 else 0)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            blases = c.build_blas_batch(lists, flags)
            inst = torch.from_numpy(pyrtx.instance_records(sc, [x.address for x in blases]).view(np.uint8)).to(
                c.device)
            tlas = c.build_tlas(inst, flags)
            films = torch.empty((sum(v["width"] * v["height"] for v in views), 4), dtype=torch.float32,
                                device=c.device)
            # the direct calls the recording must reproduce (in place, as the recording will rebuild)
            c.build_blas_batch(lists, flags, into=blases, scratch=blases[0].keepalive[0])
            c.build_tlas(inst, flags, into=tlas, scratch=tlas.keepalive[0])
            c.render_batch(tlas, rec, views, lt, spp=spp, max_depth=depth, films=films)
            c.sync()
            direct = films.clone()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
            c.build_blas_batch(lists, flags, into=blases, scratch=blases[0].keepalive[0])
            c.build_tlas(inst, flags, into=tlas, scratch=tlas.keepalive[0])
            c.render_batch(tlas, rec, views, lt, spp=spp, max_depth=depth, films=films)
        form = c.render_form()
        g.replay()
        torch.cuda.synchronize()
        same = bool(torch.equal(films, direct))
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 3
        with torch.cuda.stream(s):
            a.record()
            for _ in range(reps):
                g.replay()
            b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        c.sync()
        log(f"recorded frame (batched build + TLAS build + {spp}-spp batched render, one graph): {ms:.1f} ms, "
            f"replay equals direct calls: {same}")
        del g
        return {"ms_per_frame": round(ms, 1), "replay_equals_direct": same,
                "render_form": {pyrtx.RTX_RENDER_FORM_WHOLE_PATH: "whole-path", pyrtx.RTX_RENDER_FORM_WAVEFRONT:
                                "wavefront", pyrtx.RTX_RENDER_FORM_DEVICE_CHOSEN: "device-chosen"}.get(form, str(form))}
    finally:
        c.close()


if __name__ == "__main__":
    main()
