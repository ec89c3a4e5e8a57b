"""Drivers of the batched and served workloads over the C-ABI (used by bench.py).

  batched_form   256 tile jobs drawn from 16 seeded scenes (triangle counts log-uniform in
                 [10k, 2M], plain_gpu generator), tile sizes uniform over {32,64,128,256}^2,
                 samples log-uniform in [1, 64]; every pass of every job goes to one
                 mcr_render_batch call (one wavefront run, bounce iterations replayed from a
                 captured graph), with the path-slot count picked by `autotune` on a sample.
  served_mix     Poisson arrivals over 32 resident scenes (50k-4M triangles) of requests with a
                 resolution uniform over {320x240, 640x480, 1280x720, 1920x1080} and samples
                 log-uniform in [4, 256]; 30% of arrivals are camera moves that cancel an
                 in-flight render of the same scene and submit the moved view. Reports
                 time-to-first-preview and time-to-target-samples percentiles and Mrays/s.
  textured_form  the 1.38M-triangle workload with maps and an HDR environment: Mrays/s, and the
                 upload costs of a scene version (geometry, shading with texels, shading only).
Log-uniform draws are exp(U[ln a, ln b]) rounded to an integer.
"""
import json
import math
import os
import time

import numpy as np

from . import abi, native
from .compile import compile_camera, compile_scene
from .raytrace import trace_settings
from .scenes import orbit_camera, sized_gpu_scene
from .types import Tile

RESOLUTIONS = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]
TILE_SIZES = [32, 64, 128, 256]


def log_uniform(rng, a, b, size=None):
    return np.rint(np.exp(rng.uniform(math.log(a), math.log(b), size))).astype(np.int64)


def resident_scenes(ctx, seed, count, lo, hi, log=print):
    """`count` plain_gpu scenes with log-uniform triangle counts, BVHs built on the GPU."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = log_uniform(rng, lo, hi, count)
    out = []
    t0 = time.time()
    for i, n in enumerate(sizes):
        scene, options = sized_gpu_scene(seed * 1000 + i, int(n))
        sc = compile_scene(scene, ctx.build_bvh)
        gpu = native.Scene(ctx)
        gpu.set_geometry(sc.geometry, 1)
        gpu.set_shading(sc.shading, 1)
        out.append((gpu, int(sc.geometry.triangleCount), options))
    log(f'{count} scenes, {sum(n for _, n, _ in out):,} triangles, built in {time.time() - t0:.1f} s')
    return out


# ---- precision form: the probe-ray set ----------------------------------------------------------
def camera_rays(camera, xs, ys, width, height):
    """cameraRay (tracer.ts:705-727) for sub-pixel positions, in FP64 numpy: origin on the near
    plane, unit direction towards the far plane. Returns n x 6."""
    p = np.asarray(camera.projectionInverse, np.float64).reshape(4, 4).T  # column-major
    m = np.asarray(camera.world, np.float64).reshape(4, 4).T
    nx, ny = xs / width * 2 - 1, 1 - ys / height * 2

    def unproject(z):
        v = np.stack([nx, ny, np.full_like(nx, z), np.ones_like(nx)])
        h = p @ v
        w = (m @ np.vstack([h[:3] / h[3], np.ones_like(nx)]))[:3]
        return w.T
    a, b = unproject(-1.0), unproject(1.0)
    d = b - a
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.hstack([a, d])


def probe_rays(geometry, camera, width, height, n, seed, closest):
    """The precision probe set: n/2 primary rays through uniform random sub-pixel positions and n/2
    secondary rays leaving their hits (offset by the geometry epsilon along the geometric normal,
    towards the viewer) in cosine-distributed directions. Origins and directions are rounded to
    float32, so every consumer traces identical rays. closest(rays) -> (tri, t, ..., front) gives the
    primary hits. Returns n x 6 float32 (primaries first)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    half = n // 2
    prim = camera_rays(camera, rng.uniform(0, width, half), rng.uniform(0, height, half), width, height)
    prim = prim.astype(np.float32)
    sec = []
    got = 0
    src = prim
    while got < n - half:
        tri, t, _, _, front = closest(src.astype(np.float64))[:5]
        ok = np.flatnonzero(tri >= 0)
        r = src[ok].astype(np.float64)
        tr = np.asarray(geometry.triangles, np.float64).reshape(-1, 9)[tri[ok]]
        g = np.cross(tr[:, 3:6], tr[:, 6:9])
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        g *= np.where(front[ok] != 0, 1.0, -1.0)[:, None]
        # the normal facing the incoming ray
        g *= np.where((g * r[:, 3:6]).sum(1) > 0, -1.0, 1.0)[:, None]
        hit = r[:, :3] + r[:, 3:6] * t[ok][:, None]
        u1, u2 = rng.random(len(ok)), rng.random(len(ok))
        phi, rad = 2 * np.pi * u1, np.sqrt(u2)
        lx, ly, lz = rad * np.cos(phi), rad * np.sin(phi), np.sqrt(np.maximum(0, 1 - u2))
        sgn = np.where(g[:, 2] >= 0, 1.0, -1.0)  # Duff et al. 2017 basis
        a = -1.0 / (sgn + g[:, 2])
        bb = g[:, 0] * g[:, 1] * a
        t1 = np.stack([1 + sgn * g[:, 0] ** 2 * a, sgn * bb, -sgn * g[:, 0]], 1)
        t2 = np.stack([bb, sgn + g[:, 1] ** 2 * a, -g[:, 1]], 1)
        d = t1 * lx[:, None] + t2 * ly[:, None] + g * lz[:, None]
        o = hit + g * float(geometry.epsilon)
        sec.append(np.hstack([o, d]).astype(np.float32))
        got += len(ok)
        src = camera_rays(camera, rng.uniform(0, width, half), rng.uniform(0, height, half), width,
                          height).astype(np.float32)
    return np.vstack([prim, np.vstack(sec)[:n - half]])


# ---- precision form --------------------------------------------------------------------------------
# Record sizes the two query kernels read per visit (for algorithmic bytes): the FP64 path reads a
# 64-byte binary inner record per node and a 48-byte triangle record; the wide path an 80-byte
# node and a 48-byte triangle.
FP64_NODE_BYTES, WIDE_NODE_BYTES, TRI_BYTES = 64, 80, 48


def precision_form(ctx, gpu, geometry, camera, width=1920, height=1080, n=1_000_000, seed=1234, repeats=5,
                   settings=None, render_passes=16, log=None):
    """Closest-hit throughput of the FP64 binary query and the FP32 wide query on the same resident
    probe rays (inputs and outputs in HBM), with node visits, triangle tests and the wide BVH's
    build time; with `settings`, also `render_passes` full-frame passes rendered with each
    intersection mode (same seeds): throughput, the relative difference of the frame means and the
    fraction of pixel channels whose means differ by more than 3 sigma. Returns a dict."""
    import ctypes as C
    from . import abi, native
    rays = probe_rays(geometry, camera, width, height, n, seed, lambda r: gpu.intersect_closest(r, cull=True))
    info = gpu.build_wide()
    if log:
        log(f'wide BVH: {info}')
    r32 = native.DeviceBuffer(ctx, n * 24)
    r64 = native.DeviceBuffer(ctx, n * 48)
    tri = native.DeviceBuffer(ctx, n * 4)
    tbuf = native.DeviceBuffer(ctx, n * 8)
    r32.upload(rays)
    r64.upload(rays.astype(np.float64))
    L = native.lib()
    ctx.set_timing(True)
    out = {'rays': n, 'wide_bvh': info}

    def timed(call, node_bytes):
        call(abi.Counters())  # warm
        cnt = abi.Counters()
        ms = 0.0
        for _ in range(repeats):
            call(cnt)
            ms += ctx.last_timing()['closest_ms']
        c = cnt.as_dict()
        q = max(1, c['closest_queries'])
        algo = c['node_visits'] * node_bytes + c['triangle_tests'] * TRI_BYTES
        return {'mrays_per_s': round(q / (ms / 1e3) / 1e6, 1), 'ms_per_launch': round(ms / repeats, 3),
                'node_visits_per_ray': round(c['node_visits'] / q, 2),
                'triangle_tests_per_ray': round(c['triangle_tests'] / q, 2),
                'algorithmic_gb_per_s': round(algo / (ms / 1e3) / 1e9, 1)}

    out['fp64_binary'] = timed(lambda cnt: native.check(L.mcr_intersect_closest(
        ctx.handle, gpu.handle, r64.ptr, n, 0.0, float('inf'), 1, tri.ptr, tbuf.ptr, None, None, None, 1,
        C.byref(cnt))), FP64_NODE_BYTES)
    out['fp32_wide'] = timed(lambda cnt: native.check(L.mcr_intersect_closest_wide(
        ctx.handle, gpu.handle, r32.ptr, n, 0.0, float('inf'), 1, tri.ptr, tbuf.ptr, None, None, None, 1,
        C.byref(cnt))), WIDE_NODE_BYTES)
    out['wide_speedup'] = round(out['fp32_wide']['mrays_per_s'] / out['fp64_binary']['mrays_per_s'], 2)
    ctx.set_timing(False)
    for b in (r32, r64, tri, tbuf):
        b.free()
    if settings is not None and render_passes:
        from .raytrace import render_passes as run_passes
        npx = width * height
        data = native.DeviceBuffer(ctx, render_passes * npx * 16)
        stats = {}
        try:
            for name, mode in (('fp64', abi.MCR_ISECT_FP64), ('wide_fp32', abi.MCR_ISECT_WIDE_FP32)):
                ctx.set_intersection(mode)
                try:
                    run_passes(gpu, camera, settings, width, height, range(1), data.ptr, abi.Counters())
                    cnt = abi.Counters()
                    t0 = time.perf_counter()
                    run_passes(gpu, camera, settings, width, height, range(render_passes), data.ptr, cnt)
                    dt = time.perf_counter() - t0
                finally:
                    ctx.set_intersection(abi.MCR_ISECT_FP64)
                img = data.download(np.float32, render_passes * npx * 4).reshape(render_passes, npx, 4)[..., :3]
                stats[name] = ((cnt.closest_queries + cnt.any_queries) / dt / 1e6, img.mean(0, dtype=np.float64),
                               img.var(0, dtype=np.float64, ddof=1))
                del img
                if log:
                    log(f'{name} render: {stats[name][0]:.1f} Mrays/s')
        finally:
            data.free()
        (ra, ma, va), (rb, mb, vb) = stats['fp64'], stats['wide_fp32']
        se = np.sqrt(va / render_passes + vb / render_passes)
        diff = np.abs(ma - mb)
        out['render'] = {
            'passes': render_passes, 'fp64_mrays_per_s': round(ra, 1), 'wide_fp32_mrays_per_s': round(rb, 1),
            'frame_mean_rel_diff': float(abs(mb.mean() - ma.mean()) / ma.mean()),
            'pixel_channels_outside_3_sigma': float(((diff > 3 * se) & (diff > 1e-12)).mean())}
    return out


# ---- batched form --------------------------------------------------------------------------------
def batched_jobs(rng, scenes, n_jobs=256, width=1920, height=1080):
    """(views, jobs, job_views, floats, samples): every pass of every tile job."""
    views = []
    for gpu, _, options in scenes:
        cam = compile_camera(orbit_camera(width, height, float(rng.uniform(-0.5, 0.5))))
        views.append((gpu, cam, trace_settings(options)))
    jobs, jv, off, samples = [], [], 0, 0
    for _ in range(n_jobs):
        v = int(rng.integers(len(scenes)))
        tw, th = int(rng.choice(TILE_SIZES)), int(rng.choice(TILE_SIZES))
        x, y = int(rng.integers(0, width - tw + 1)), int(rng.integers(0, height - th + 1))
        spp = int(log_uniform(rng, 1, 64))
        for p in range(spp):
            jobs.append(abi.TileJob(Tile(x, y, tw, th).to_c(), width, height, p, 1, off))
            jv.append(v)
            off += tw * th * 4
            samples += tw * th
    return views, jobs, jv, off, samples


def run_batch(ctx, views, jobs, jv, out_ptr, capacity, graph):
    ctx.set_capacity(capacity)
    ctx.set_graph(graph)
    cnt = abi.Counters()
    t = time.perf_counter()
    native.render_batch(ctx, views, jobs, jv, out=out_ptr, out_device=True, counters=cnt)
    return time.perf_counter() - t, cnt


def autotune(ctx, views, jobs, jv, out_ptr, log=print,
             capacities=(1 << 20, 1 << 21, 1 << 22, 1 << 23), graphs=(True, False)):
    """Times each (path slots, graph replay) choice on the first quarter of the jobs and returns
    the fastest, with the trial table."""
    sample = jobs[:max(1, len(jobs) // 4)]
    trials = []
    for cap in capacities:
        for g in graphs:
            run_batch(ctx, views, sample, jv[:len(sample)], out_ptr, cap, g)  # allocate, capture
            dt, cnt = run_batch(ctx, views, sample, jv[:len(sample)], out_ptr, cap, g)
            rate = (cnt.closest_queries + cnt.any_queries) / dt / 1e6
            trials.append({'slots': cap, 'graph': g, 'mrays_s': round(rate, 1)})
    best = max(trials, key=lambda t: t['mrays_s'])
    log(f'autotune: {best} of {len(trials)} trials')
    return best, trials


def batched_form(ctx, seed=1, n_jobs=256, n_scenes=16, log=print, tune=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = resident_scenes(ctx, seed, n_scenes, 10_000, 2_000_000, log)
    views, jobs, jv, floats, samples = batched_jobs(rng, scenes, n_jobs)
    out = native.DeviceBuffer(ctx, floats * 4)
    try:
        best, trials = (autotune(ctx, views, jobs, jv, out.ptr, log) if tune
                        else ({'slots': native.DEFAULT_SLOTS, 'graph': True}, []))
        run_batch(ctx, views, jobs, jv, out.ptr, best['slots'], best['graph'])  # warm
        dt, cnt = run_batch(ctx, views, jobs, jv, out.ptr, best['slots'], best['graph'])
    finally:
        out.free()
        ctx.set_capacity()
        ctx.set_graph(True)
        for gpu, _, _ in scenes:
            gpu.release()
    rays = cnt.closest_queries + cnt.any_queries
    return {'jobs': n_jobs, 'tile_passes': len(jobs), 'scenes': n_scenes,
            'triangles': int(sum(n for _, n, _ in scenes)),
            'samples': samples, 'rays': int(rays), 'seconds': round(dt, 4),
            'mrays_s': round(rays / dt / 1e6, 2), 'paths_per_s': round(samples / dt, 1),
            'slots': best['slots'], 'graph': best['graph'], 'autotune': trials}


def batched_equivalent(ctx, scenes, specs, slots=1 << 21, log=print):
    """The plain batched form of a served run's content: every request in `specs` (scene index,
    width, height, samples, camera angle), each as 64-px tiles x passes, all in one
    mcr_render_batch call. Results are overwritten pass after pass (only the time matters)."""
    views, jobs, jv = [], [], []
    for i, (k, w, h, spp, angle) in enumerate(specs):
        gpu, _, options = scenes[k]
        views.append((gpu, compile_camera(orbit_camera(w, h, angle)), trace_settings(options)))
        tiles = [(x, y, min(64, w - x), min(64, h - y)) for y in range(0, h, 64) for x in range(0, w, 64)]
        for p in range(spp):
            for (x, y, tw, th) in tiles:
                jobs.append(abi.TileJob(Tile(x, y, tw, th).to_c(), w, h, p, 1, (y * w + x * th) * 4))
                jv.append(i)
    out = native.DeviceBuffer(ctx, 1920 * 1080 * 16)
    try:
        n = max(1, len(jobs) // 16)
        run_batch(ctx, views, jobs[:n], jv[:n], out.ptr, slots, False)  # allocate
        dt, cnt = run_batch(ctx, views, jobs, jv, out.ptr, slots, False)
    finally:
        out.free()
        ctx.set_capacity()
        ctx.set_graph(True)
    rays = cnt.closest_queries + cnt.any_queries
    log(f'batched equivalent: {len(jobs)} tile passes of {len(specs)} requests in {dt:.2f} s')
    return {'requests': len(specs), 'tile_passes': len(jobs), 'slots': slots, 'seconds': round(dt, 3),
            'mrays_s': round(rays / dt / 1e6, 2),
            'mixed_view_wave_frac': round(cnt.mixed_view_waves / max(1, cnt.shade_waves), 4)}


def autotune_served(ctx, scenes, target_ms=50.0, rate=16.0, seconds=8.0, margin=0.9, slots=1 << 23,
                    preview_cus=(8, 16, 24), preview_passes=(False,), log=print):
    """Preview-lane CUs, and whether the preview lane also runs passes, for the served form: a
    short run of the mix for each pair at an arrival rate that keeps the pass lane saturated (its
    `slots`: more pass slots only buy throughput, since previews run on their own lane); the
    choice with the fewest preview CUs (the most left to the passes) whose p90 time-to-first-preview
    stays under margin * target_ms / 2 wins; if none does, the one with the shortest median
    preview. A trial's p90 is the test because its p99 rests on the two slowest of ~130 previews,
    which in 8 s trials sit far above what the long runs show for the same choice, and the half
    target keeps the long runs' p99 under the target (8 preview CUs: trial p90 31 ms, p99 56-82 ms
    in the long runs at 2 and 8 requests/s; 16 CUs: trial p90 19 ms). The trials' throughputs
    differ by less than their run-to-run spread, so they do not decide. More
    preview CUs shorten previews and take CUs from the passes; passes on the preview lane use those
    CUs between previews and lengthen previews. The margin covers the spread between a short trial
    and a long run."""
    trials = []
    for cus in preview_cus:
        for pp in preview_passes:
            r = served_mix(ctx, rate, seconds, seed=7, scenes=scenes, slots=slots, preview_cus=cus,
                           preview_passes=pp, drain_seconds=2.0, log=lambda *_: None)
            trials.append({'slots': slots, 'preview_cus': cus, 'preview_passes': pp,
                           'ttfp_ms_p50': r['ttfp_ms_p50'], 'ttfp_ms_p90': r['ttfp_ms_p90'],
                           'ttfp_ms_p99': r['ttfp_ms_p99'], 'mrays_s': r['mrays_s']})
            log(f'served autotune: {trials[-1]}')
    ok = [t for t in trials if t['ttfp_ms_p90'] is not None and t['ttfp_ms_p90'] <= margin * target_ms / 2]
    if ok:
        return min(ok, key=lambda t: (t['preview_cus'], t['preview_passes'], -t['mrays_s'])), trials
    timed = [t for t in trials if t['ttfp_ms_p50'] is not None] or trials
    return min(timed, key=lambda t: t['ttfp_ms_p50'] if t['ttfp_ms_p50'] is not None else 0), trials


# ---- served form ---------------------------------------------------------------------------------
def served_mix(ctx, rate, seconds, seed=2, n_scenes=32, scenes=None, restart_frac=0.3,
               budget=200 << 30, slots=0, window=0, max_requests=1024, drain_seconds=60.0, lanes=0,
               preview_slots=0, preview_cus=0, preview_passes=False, warmup=True, log=print):
    """Runs the arrival process for `seconds`, waits up to `drain_seconds` for the admitted
    requests, cancels what is left, and returns the latency and throughput figures."""
    rng = np.random.Generator(np.random.PCG64(seed))
    own = scenes is None
    if own:
        scenes = resident_scenes(ctx, seed, n_scenes, 50_000, 4_000_000, log)
    srv = native.Server(ctx, hbm_budget=budget, slots=slots, window=window, max_requests=max_requests,
                        lanes=lanes, preview_slots=preview_slots, preview_cus=preview_cus,
                        preview_passes=preview_passes)
    try:
        for gpu, _, _ in scenes:
            srv.add_scene(gpu)
        # A first request warms the new server (graph captures, first use of each lane and of the
        # allocation pool), and two more at once, on other scenes, capture the lanes' mixed-view
        # kernels; the first one's time to first preview is reported on its own, not in the mix.
        warm_ttfp = None
        if warmup:
            warm = []
            for k in ([0], [1 % len(scenes), 2 % len(scenes)]):
                ids = []
                for i in k:
                    gpu, _, options = scenes[i]
                    ids.append(srv.submit(gpu, compile_camera(orbit_camera(1920, 1080, 0.0)),
                                          trace_settings(options), 1920, 1080, 2, 64, 4))
                for rid in ids:
                    srv.wait(rid, 60000)
                warm += ids
            st = srv.status(warm[0])
            warm_ttfp = round(st['first_preview_ms'] - st['submit_ms'], 2)
            for rid in warm:
                srv.release(rid)
        live = []          # (rid, scene index, w, h, spp, angle)
        record = {}        # rid -> kind
        t_next = 0.0
        stats0 = srv.stats()
        t0 = time.perf_counter()
        submitted = restarts = 0
        while True:
            t_next += float(rng.exponential(1.0 / rate))
            if t_next > seconds:
                break
            delay = t0 + t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            live = [x for x in live if srv.status(x[0])['state'] in (abi.REQ_QUEUED, abi.REQ_RUNNING)]
            if live and rng.random() < restart_frac:
                old = live.pop(int(rng.integers(len(live))))
                srv.cancel(old[0])
                k, w, h, spp, angle = old[1:]
                angle += float(rng.uniform(-0.05, 0.05))
                kind = 'restart'
                restarts += 1
            else:
                k = int(rng.integers(len(scenes)))
                w, h = RESOLUTIONS[int(rng.integers(len(RESOLUTIONS)))]
                spp = int(log_uniform(rng, 4, 256))
                angle = float(rng.uniform(-0.6, 0.6))
                kind = 'new'
            gpu, _, options = scenes[k]
            cam = compile_camera(orbit_camera(w, h, angle))
            rid = srv.submit(gpu, cam, trace_settings(options), w, h, spp, 64, 4)
            live.append((rid, k, w, h, spp, angle))
            record[rid] = (kind, (k, w, h, spp, angle))
            submitted += 1
        t_arrivals = time.perf_counter() - t0
        stats_mid = srv.stats()  # the device has been kept busy by arrivals until here
        deadline = time.perf_counter() + drain_seconds
        for rid in record:
            left = (deadline - time.perf_counter()) * 1e3
            if left <= 0:
                break
            try:
                srv.wait(rid, left)
            except native.NativeError as e:
                if e.code != abi.MCR_E_TIMEOUT:
                    raise
                break
        t_end = time.perf_counter() - t0
        stats1 = srv.stats()
        for rid in record:
            if srv.status(rid)['state'] in (abi.REQ_QUEUED, abi.REQ_RUNNING):
                srv.cancel(rid)
        ttfp, ttt, cancelled_before_preview, unfinished, wait, iters = [], [], 0, 0, [], []
        for rid in record:
            st = srv.status(rid)
            if os.environ.get('MCR_SERVED_TRACE'):
                log(json.dumps({'rid': rid, 'kind': record[rid][0], 'spec': record[rid][1], **st}))
            if st['first_preview_ms'] >= 0:
                ttfp.append(st['first_preview_ms'] - st['submit_ms'])
                wait.append(st['admit_ms'] - st['submit_ms'])
                iters.append(st['preview_iterations'])
            elif st['state'] == abi.REQ_CANCELLED:
                cancelled_before_preview += 1
            if st['state'] == abi.REQ_DONE:
                ttt.append(st['done_ms'] - st['submit_ms'])
            elif st['state'] != abi.REQ_CANCELLED:
                unfinished += 1
        rays = (stats1['closest_queries'] + stats1['any_queries']) - (stats0['closest_queries'] + stats0['any_queries'])
        rays_mid = ((stats_mid['closest_queries'] + stats_mid['any_queries']) -
                    (stats0['closest_queries'] + stats0['any_queries']))
        wall_mid = max(stats_mid['uptime_ms'] - stats0['uptime_ms'], 1e-9)
        wall = max(stats1['uptime_ms'] - stats0['uptime_ms'], 1e-9)
        device = max(stats1['device_ms'] - stats0['device_ms'], 1e-9)

        def pct(a, q):
            return round(float(np.percentile(a, q)), 2) if a else None
        completed_specs = [record[rid][1] for rid in record if srv.status(rid)['state'] == abi.REQ_DONE]
        return {'completed_specs': completed_specs, 'first_request_ttfp_ms': warm_ttfp, 'rate_per_s': rate,
                'arrival_seconds': round(t_arrivals, 2), 'submitted': submitted,
                'restarts': restarts, 'previews': len(ttfp), 'completed': len(ttt),
                'cancelled_before_preview': cancelled_before_preview, 'unfinished_at_drain': unfinished,
                'ttfp_ms_p50': pct(ttfp, 50), 'ttfp_ms_p90': pct(ttfp, 90), 'ttfp_ms_p99': pct(ttfp, 99),
                'time_to_target_ms_p50': pct(ttt, 50), 'time_to_target_ms_p99': pct(ttt, 99),
                'admit_wait_ms_p50': pct(wait, 50), 'preview_iterations_p50': pct(iters, 50),
                'preview_iterations_p99': pct(iters, 99),
                'iterations': int(stats1['iterations'] - stats0['iterations']),
                'mixed_view_wave_frac': round((stats1['mixed_view_waves'] - stats0['mixed_view_waves']) /
                                              max(1, stats1['shade_waves'] - stats0['shade_waves']), 4),
                'ms_per_iteration': round(device / max(1, stats1['iterations'] - stats0['iterations']), 3),
                'mrays_s': round(rays / (wall / 1e3) / 1e6, 2),
                'mrays_s_arrivals': round(rays_mid / (wall_mid / 1e3) / 1e6, 2),
                'mrays_s_device_busy': round(rays / (device / 1e3) / 1e6, 2), 'wall_seconds': round(t_end, 2),
                'paths': int(stats1['paths'] - stats0['paths']),
                # the server thread's own work and the pass lane's time with a step in flight, as
                # fractions of the wall time
                'server_thread': {k: round((stats1[k] - stats0[k]) / wall, 4)
                                  for k in ('host_ms', 'plan_ms', 'deliver_ms', 'pass_busy_ms')}}
    finally:
        srv.close()
        if own:
            for gpu, _, _ in scenes:
                gpu.release()


def textured_form(ctx, width=1920, height=1080, steps=16, warmup=1, seed=0, plain_mrays=None, log=print):
    """The plain_gpu workload with textures and an environment (scenes.textured_gpu_scene): Mrays/s
    over `steps` one-sample passes of the resident scene, and the residency costs of a scene
    version (textures.ts TextureTable upload with the shading; progressive.ts:80-157): the geometry
    upload, the first shading upload with its texels, and a shading-only update (one material
    colour changed, texels unchanged)."""
    from .raytrace import render_passes
    from .scenes import textured_gpu_scene
    scene, camera, opts = textured_gpu_scene(seed, width, height)
    sc = compile_scene(scene, ctx.build_bvh)
    cam = compile_camera(camera)
    st = trace_settings(opts)
    texels = sum(int(t.data.nbytes) for t in sc.shading.textures)
    gpu = native.Scene(ctx)
    t0 = time.perf_counter()
    gpu.set_geometry(sc.geometry, 1)
    t1 = time.perf_counter()
    gpu.set_shading(sc.shading, 1)
    t2 = time.perf_counter()
    sc.shading.materials[0].color = (0.5, 0.5, 0.5)
    t3 = time.perf_counter()
    gpu.set_shading(sc.shading, 2)
    t4 = time.perf_counter()
    npx = width * height
    out = native.DeviceBuffer(ctx, max(steps, warmup, 1) * npx * 16)
    if warmup:
        render_passes(gpu, cam, st, width, height, range(warmup), out.ptr, abi.Counters())
    cnt = abi.Counters()
    ts = time.perf_counter()
    render_passes(gpu, cam, st, width, height, range(warmup, warmup + steps), out.ptr, cnt)
    elapsed = time.perf_counter() - ts
    c = cnt.as_dict()
    rays = c['closest_queries'] + c['any_queries']
    out.free()
    gpu.release()
    mrays = rays / elapsed / 1e6
    log(f'textured: {steps} passes in {elapsed:.3f} s, {mrays:.1f} Mrays/s')
    res = {'workload': 'textured_gpu: plain_gpu with a 512x512 sRGB colour map, 256x256 normal and roughness '
                       'maps on the terrain, 64x32 sRGB maps on the non-glass spheres, a 256x128 HDR sky as '
                       'environment and background; 1920x1080, GI on, 8 bounces, ACES',
           'textures': len(sc.shading.textures), 'texel_bytes': texels, 'steps': steps,
           'ms_per_step': round(elapsed / steps * 1e3, 2), 'mrays_s': round(mrays, 2),
           'rays_per_step': int(rays / steps),
           'residency_ms': {'geometry_upload': round((t1 - t0) * 1e3, 2),
                            'shading_upload_with_texels': round((t2 - t1) * 1e3, 2),
                            'shading_only_update': round((t4 - t3) * 1e3, 2)}}
    if plain_mrays:
        res['vs_plain'] = round(mrays / plain_mrays, 3)
    return res
