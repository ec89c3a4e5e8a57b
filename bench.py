"""Benchmark: trace throughput (Mrays/s) of the plain GPU workload on one MI355X.

Workload: a 724x724-quad fBm heightfield plus 64 subdivision-4 icospheres (1,376,032 triangles,
20% glass), a sun and 4 point lights, 1920x1080, global illumination on, 8 bounces, ACES. The BVH
is built on the GPU. One step is one full-frame pass (1 sample per pixel); the default 64 steps
are the 64-sample frame, traced in one device run with the scene resident in HBM, followed by the
on-device accumulate + encode (writeTile) of every pass in order.
Mrays/s counts every BVH query: closest hits (camera, bounce, pass-through) and shadow any-hits.

Prints one JSON line. `roofline` is computed for the kernel that takes the most device time,
from algorithmic bytes (inner-node records x 64 B + triangle records x 48 B, counted by the
kernels) over its HIP-event-timed duration. `cpu_baseline` times the oracle (the C++ restatement
of the reference's CPU path) on a bounded sample of the same frame on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'Mrays/s @1M tris 1080p 64spp vs CPU; p50/p99 time-to-first-preview in mix'
BATCHED_BEST_SLOTS = 1 << 23  # the batched form's autotuner pick on its mix (2^20 .. 2^23 tried)
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES = 64, 48


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """The host CPU's model, the cores this process may run on (its affinity mask) and the thread
    count a CPU baseline uses: the affinity's size, at most 16 (the host share of one GPU)."""
    model = 'unknown'
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    return {'model': model, 'affinity': _ranges(cpus), 'threads': max(1, min(16, len(cpus)))}


def _ranges(cpus):
    out, start = [], None
    for i, c in enumerate(cpus):
        if start is None:
            start = c
        if i + 1 == len(cpus) or cpus[i + 1] != c + 1:
            out.append(f'{start}-{c}' if c != start else f'{c}')
            start = None
    return ','.join(out)


def cpu_baseline(sc, cam, st, width, height, seconds=15.0, single_seconds=5.0):
    """Times the oracle on a sample of the frame's (tile, pass) work items in the reference's
    progressive order (64x64 tiles centre-first, pass-major from pass 1) with one thread per host
    core, sized to about `seconds` of work; then one thread alone on about `single_seconds`."""
    import oracle
    from meercadraytracer_amd import abi
    host = host_cpu()
    threads = host['threads']
    tiles = oracle.make_tiles(width, height, 64)

    def jobs_for(k):
        jobs, off = [], 0
        for i in range(k):
            x, y, w, h = tiles[i % len(tiles)]
            jobs.append(abi.TileJob(abi.Tile(x, y, w, h), width, height, 1 + i // len(tiles), 1, off))
            off += w * h * 4
        return jobs
    # calibrate: one tile per thread
    probe = jobs_for(threads)
    secs, cnt = oracle.render_items_mt(sc, cam, st, probe, threads)
    per_tile = secs / max(1, len(probe)) * threads
    k = int(min(64 * len(tiles), max(threads, seconds / max(per_tile, 1e-6) * threads)))
    jobs = jobs_for(k)
    secs, cnt = oracle.render_items_mt(sc, cam, st, jobs, threads)
    rays = cnt['closest_queries'] + cnt['any_queries']
    k1 = int(max(1, min(k, single_seconds / max(per_tile, 1e-6))))
    secs1, cnt1 = oracle.render_items_mt(sc, cam, st, jobs_for(k1), 1)
    rays1 = cnt1['closest_queries'] + cnt1['any_queries']
    return {'value': rays / secs / 1e6, 'unit': 'Mrays/s', 'cores': threads, 'kind': 'port',
            'cpu_model': host['model'], 'affinity': host['affinity'],
            'single_thread': {'value': round(rays1 / secs1 / 1e6, 3), 'unit': 'Mrays/s', 'items': k1,
                              'seconds': round(secs1, 2)},
            'sample': f'{k} (tile, pass) items = {k / len(tiles):.2f} passes of {len(tiles)} centre-first 64x64 tiles '
                      f'of the same 1920x1080 frame from pass 1, '
                      f'{rays} queries in {secs:.2f} s (portable-math oracle, std::thread pool)'}


KERNEL_SYMBOLS = {'extend (closest hit)': 'k_extend', 'shadow (any hit)': 'k_shadow'}


def _pmc_run(counters, bench_args, outdir, timeout=240):
    """One rocprofv3 counter pass (the counters of one hardware pass, space separated) over a
    child bench run; returns {kernel symbol: {counter: [value per dispatch, ...]}} in dispatch
    order."""
    import csv
    import subprocess
    names = counters.split()
    d = os.path.abspath(os.path.join(outdir, '_'.join(names)))
    os.makedirs(d, exist_ok=True)
    cmd = ['rocprofv3', '--pmc', *names, '--output-format', 'csv', '-d', d, '-o', 'pmc', '--',
           sys.executable, os.path.abspath(__file__), *bench_args]
    env = dict(os.environ, TMPDIR='/tmp')
    with open(os.path.join(d, 'rocprofv3.log'), 'w') as logf:
        subprocess.run(cmd, cwd='/tmp', env=env, stdout=logf, stderr=subprocess.STDOUT, timeout=timeout, check=True)
    rows = []
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith('counter_collection.csv'):
                rows += list(csv.DictReader(open(os.path.join(root, f))))
    rows.sort(key=lambda r: int(r['Dispatch_Id']))
    out = {}
    for r in rows:
        per = out.setdefault(kernel_symbol(r['Kernel_Name']), {})
        per.setdefault(r['Counter_Name'], []).append(float(r['Counter_Value']))
    return out


def kernel_symbol(name):
    """'void k_shadow<false>(ViewDev const*, ...)' -> 'k_shadow'."""
    name = name.split('(')[0].strip()
    if name.startswith('void '):
        name = name[5:]
    return name.split('<')[0]


def measure_traffic(args, kernel, launches):
    """HBM bytes per launch of `kernel` over the timed launches of the same bench command, from
    FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a 4 GiB device copy."""
    import shutil
    import tempfile
    if not shutil.which('rocprofv3') or launches == 0:
        return None
    outdir = tempfile.mkdtemp(prefix='mcr_pmc_', dir='/tmp')
    # Only the plain run: the other sections launch the same kernels and would mix into the tail.
    bench_args = ['--steps', str(args.steps), '--warmup', str(args.warmup), '--no-cpu', '--no-traffic',
                  '--no-batched', '--no-served', '--no-precision', '--no-textures', '--no-c0',
                  '--width', str(args.width), '--height', str(args.height)]
    copy_args = ['--copy-calibration', '4096']
    res = {}
    for counter in ('FETCH_SIZE', 'WRITE_SIZE'):
        copy = _pmc_run(counter, copy_args, os.path.join(outdir, 'copy'))
        copied_kib = sum(copy.get('__amd_rocclr_copyBuffer', {}).get(counter, []))
        factor = (4096 * 1024) / copied_kib if copied_kib else None
        per = _pmc_run(counter, bench_args, outdir).get(KERNEL_SYMBOLS[kernel], {})
        vals = (per.get(counter) or next(iter(per.values()), []))[-launches:]
        if not factor or len(vals) < launches:
            return None
        res[counter] = (sum(vals) * factor * 1024 / launches, factor)
    shutil.rmtree(outdir, ignore_errors=True)
    return {'bytes_per_launch': int(res['FETCH_SIZE'][0] + res['WRITE_SIZE'][0]),
            'read_bytes_per_launch': int(res['FETCH_SIZE'][0]), 'write_bytes_per_launch': int(res['WRITE_SIZE'][0]),
            'fetch_scale': round(res['FETCH_SIZE'][1], 4), 'write_scale': round(res['WRITE_SIZE'][1], 4)}


# Counter passes of the traversal evidence: one rocprofv3 run each (a hardware pass holds at most
# 8 SQ, 4 TCC and 2 GRBM counters, and derived metrics bring their own inputs, so each derived
# metric runs alone). GRBM_GUI_ACTIVE (GPU cycles per dispatch) weights the per-dispatch
# percentages into one figure per kernel.
COUNTER_PASSES = ['GRBM_GUI_ACTIVE', 'TCC_HIT_sum TCC_MISS_sum', 'VALUBusy', 'VALUUtilization',
                  'OccupancyPercent', 'MemUnitStalled']
COUNTER_KERNELS = ('k_extend', 'k_shadow', 'k_shade', 'k_resolve')


def kernel_counters(args, passes=None, outdir=None):
    """L2 (TCC) hit rate, VALU busy and utilisation, occupancy and memory-unit stalls of the
    wavefront kernels over one plain run of `args.steps` passes (every launch of the run),
    each counter pass a rocprofv3 run of its own over the same command."""
    import shutil
    import tempfile
    if not shutil.which('rocprofv3'):
        return {'error': 'rocprofv3 not found'}
    keep = outdir is not None
    if keep:
        os.makedirs(outdir, exist_ok=True)
    else:
        outdir = tempfile.mkdtemp(prefix='mcr_pmc_', dir='/tmp')
    bench_args = ['--steps', str(args.steps), '--warmup', '0', '--no-cpu', '--no-traffic',
                  '--no-batched', '--no-served', '--no-precision', '--no-textures', '--no-c0',
                  '--width', str(args.width), '--height', str(args.height)]
    raw, errors = {}, {}
    for counters in (passes or COUNTER_PASSES):
        log(f'counter pass: {counters} ...')
        try:
            res = _pmc_run(counters, bench_args, outdir, timeout=180)
        except Exception as e:  # an unknown counter on this ROCm fails its own pass only
            errors[counters] = str(e)[:200]
            continue
        for k, per in res.items():
            if k in COUNTER_KERNELS:
                raw.setdefault(k, {}).update(per)
    if not keep:
        shutil.rmtree(outdir, ignore_errors=True)
    out = {}
    for k, per in raw.items():
        cyc = per.get('GRBM_GUI_ACTIVE')
        o = {'dispatches': max(len(v) for v in per.values())}
        if cyc:
            o['gpu_cycles_total'] = sum(cyc)
        hit, miss = sum(per.get('TCC_HIT_sum', [])), sum(per.get('TCC_MISS_sum', []))
        if hit + miss > 0:
            o['l2_hit_pct'] = round(100.0 * hit / (hit + miss), 2)
            o['l2_requests_per_dispatch'] = int((hit + miss) / len(per['TCC_HIT_sum']))
        for name in ('VALUBusy', 'VALUUtilization', 'OccupancyPercent', 'MemUnitStalled'):
            v = per.get(name)
            if not v:
                continue
            if cyc and len(cyc) == len(v) and sum(cyc) > 0:
                o[name] = round(sum(a * w for a, w in zip(v, cyc)) / sum(cyc), 2)
            else:
                o[name] = round(sum(v) / len(v), 2)
        out[k] = o
    return {'workload': f'plain_gpu, {args.width}x{args.height}, {args.steps} passes (every launch of the run)',
            'kernels': out, 'failed_passes': errors,
            'note': 'percentages are per-dispatch values weighted by GRBM_GUI_ACTIVE cycles of the same dispatch'}


def c0(ctx):
    """The CPU reference configuration (plain_cpu: 10,242 triangles, 256x256, 16 spp, 4 bounces,
    GI off, ACES, sRGB) on the GPU through raytrace_compiled, and the whole frame on the host's
    cores with the oracle (16 passes of centre-first 64x64 tiles, pass-major)."""
    import oracle
    from meercadraytracer_amd import abi, scenes
    from meercadraytracer_amd.raytrace import raytrace_compiled, trace_settings
    sc, cam, o = scenes.compiled('plain_cpu', 0, build_bvh=ctx.build_bvh)
    raytrace_compiled(sc, cam, 256, 256, o, ctx=ctx)  # warm
    cnt = abi.Counters()
    t = time.perf_counter()
    raytrace_compiled(sc, cam, 256, 256, o, ctx=ctx, counters=cnt)
    dt = time.perf_counter() - t
    gpu_rays = cnt.closest_queries + cnt.any_queries
    tiles = oracle.make_tiles(256, 256, 64)
    jobs = [abi.TileJob(abi.Tile(*tl), 256, 256, p, 1, 0) for p in range(16) for tl in tiles]
    threads = host_cpu()['threads']
    secs, c = oracle.render_items_mt(sc, cam, trace_settings(o), jobs, threads)
    cpu_rays = c['closest_queries'] + c['any_queries']
    return {'workload': 'plain_cpu: 10,242 triangles, 256x256, 16 spp, 4 bounces, GI off, ACES, sRGB',
            'gpu': {'seconds': round(dt, 4), 'rays': int(gpu_rays), 'mrays_s': round(gpu_rays / dt / 1e6, 2),
                    'note': 'raytrace_compiled: scene upload, 16 passes, encode and readback'},
            'cpu': {'seconds': round(secs, 3), 'rays': int(cpu_rays), 'mrays_s': round(cpu_rays / secs / 1e6, 2),
                    'cores': threads, 'kind': 'port'}}


def builder(ctx, repeats=10):
    """The binned-SAH builder on the plain GPU scene's 1,376,032 triangles already in HBM
    (mcr_build_bvh_device), `repeats` builds after one warm-up: host wall time per build."""
    from meercadraytracer_amd import native, scenes
    from meercadraytracer_amd.compile import compile_scene
    captured = {}

    def keep(tris, count):
        captured['tris'], captured['count'] = tris.copy(), count
        return ctx.build_bvh(tris, count)
    scene, _, _ = scenes.plain_gpu_scene(0)
    compile_scene(scene, keep)
    tris, n = captured['tris'], captured['count']
    dt = native.DeviceBuffer(ctx, tris.nbytes)
    dt.upload(tris)
    db = native.DeviceBuffer(ctx, (2 * n - 1) * 24)
    dn = native.DeviceBuffer(ctx, (2 * n - 1) * 12)
    do = native.DeviceBuffer(ctx, n * 4)
    import ctypes as C
    nc = C.c_uint32()
    L = native.lib()
    walls = []
    for i in range(repeats + 1):
        t = time.perf_counter()
        native.check(L.mcr_build_bvh_device(ctx.handle, dt.ptr, n, db.ptr, dn.ptr, do.ptr, C.byref(nc)))
        if i:
            walls.append((time.perf_counter() - t) * 1e3)
    for b in (dt, db, dn, do):
        b.free()
    walls.sort()
    return {'workload': 'plain_gpu triangles (1,376,032) in HBM, mcr_build_bvh_device', 'triangles': int(n),
            'nodes': int(nc.value), 'builds': repeats, 'ms_median': round(walls[len(walls) // 2], 2),
            'ms_min': round(walls[0], 2), 'ms_max': round(walls[-1], 2)}


def served(args, ctx, workloads):
    """The served mix at each arrival rate over the same 32 resident scenes. The top-level
    figures are those of the highest rate (the one meant to keep the device saturated)."""
    scenes = workloads.resident_scenes(ctx, 2, 32, 50_000, 4_000_000, log)
    try:
        tuned, trials = None, []
        slots, cus, pp = args.served_slots, 0, False
        if not slots:
            tuned, trials = workloads.autotune_served(ctx, scenes, log=log)
            slots, cus, pp = tuned['slots'], tuned['preview_cus'], tuned['preview_passes']
        sweep = []
        for rate in [float(r) for r in args.served_rates.split(',') if r]:
            log(f'served mix at {rate}/s for {args.served_seconds} s, {slots} pass slots ...')
            r = workloads.served_mix(ctx, rate, args.served_seconds, scenes=scenes, slots=slots, preview_cus=cus,
                                     preview_passes=pp, drain_seconds=10.0, log=log)
            specs = r.pop('completed_specs')
            log(json.dumps(r))
            sweep.append(r)
        # The requests the last rate completed, as one batch: on the pass lane's slots, and on the
        # slot count the batched form's own autotuner picks (its best).
        same = workloads.batched_equivalent(ctx, scenes, specs, slots=slots, log=log)
        best = same if slots == BATCHED_BEST_SLOTS else workloads.batched_equivalent(
            ctx, scenes, specs, slots=BATCHED_BEST_SLOTS, log=log)
    finally:
        for gpu, _, _ in scenes:
            gpu.release()
    top = sweep[-1]
    return {'workload': '32 resident plain_gpu scenes (50k-4M triangles, log-uniform), Poisson arrivals, '
                        'resolutions {320x240, 640x480, 1280x720, 1920x1080}, samples log-uniform [4, 256], '
                        '30% camera-move restarts, 64-px tiles, 4x4 preview, HBM budget 200 GB; previews on a '
                        'preview lane (2^18 slots, preview_cus CUs), passes on a pass lane (slots) on the rest',
            'triangles': int(sum(n for _, n, _ in scenes)),
            'ttfp_ms_p50': top['ttfp_ms_p50'], 'ttfp_ms_p99': top['ttfp_ms_p99'],
            'saturated_mrays_s': top['mrays_s'], 'mrays_s_during_arrivals': top['mrays_s_arrivals'],
            'slots': slots, 'preview_cus': cus, 'preview_passes': pp, 'autotune': trials,
            'batched_same_scenes': same, 'vs_batched_same_scenes': round(top['mrays_s'] / same['mrays_s'], 3),
            'batched_best_slots': best, 'vs_batched_best_slots': round(top['mrays_s'] / best['mrays_s'], 3),
            'sweep': sweep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=64)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-traffic', action='store_true', help='skip the rocprofv3 HBM-traffic passes')
    ap.add_argument('--width', type=int, default=1920)
    ap.add_argument('--height', type=int, default=1080)
    ap.add_argument('--no-batched', action='store_true', help='skip the batched-form measurement')
    ap.add_argument('--no-served', action='store_true', help='skip the served-mix measurement')
    ap.add_argument('--served-rates', default='2,8,16',
                    help='request arrival rates (per second) to sweep; the last one should saturate the device')
    ap.add_argument('--served-seconds', type=float, default=12.0, help='arrival window per rate')
    ap.add_argument('--served-slots', type=int, default=0, help='server path slots (0 = default)')
    ap.add_argument('--no-precision', action='store_true', help='skip the precision-form probe measurement')
    ap.add_argument('--no-textures', action='store_true', help='skip the textured-workload measurement')
    ap.add_argument('--no-c0', action='store_true', help='skip the CPU configuration (C0) on GPU and host')
    ap.add_argument('--only', choices=['plain', 'batched', 'served', 'c0', 'precision', 'textures', 'counters', 'build'],
                    default=None,
                    help='run one workload and print its JSON object only')
    ap.add_argument('--pmc-dir', default=None, help='--only counters: keep the rocprofv3 output here')
    ap.add_argument('--shade-sort', choices=['on', 'off'], default=None,
                    help='material-sorted shading on or off (default: the library default)')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='only run a device-to-device copy of MIB MiB (for scaling HBM counters)')
    args = ap.parse_args()
    if args.copy_calibration:
        from meercadraytracer_amd import native
        ctx = native.default_context()
        n = args.copy_calibration << 20
        a, b = native.DeviceBuffer(ctx, n), native.DeviceBuffer(ctx, n)
        a.zero()
        native.check(native.lib().mcr_copy(ctx.handle, b.ptr, a.ptr, n, 2))
        print(json.dumps({'copy_bytes': n}), flush=True)
        return
    if args.gpus != 1:
        print(f'bench.py measures one GPU: --gpus {args.gpus} is not supported '
              '(tiles shard across GPUs by the scheduler; only the one-GPU path is measured)', file=sys.stderr)
        sys.exit(2)

    if args.only == 'counters':  # child runs only: this process does not touch the GPU
        print(json.dumps(kernel_counters(args, outdir=args.pmc_dir)), flush=True)
        return
    from meercadraytracer_amd import abi, native, scenes, workloads
    from meercadraytracer_amd.compile import compile_camera, compile_scene
    from meercadraytracer_amd.raytrace import render_passes, trace_settings
    from meercadraytracer_amd.types import Tile

    W, H = args.width, args.height
    ctx = native.default_context()
    if args.shade_sort:
        ctx.set_shade_sort(args.shade_sort == 'on')
    if args.only == 'batched':
        print(json.dumps(workloads.batched_form(ctx, log=log)), flush=True)
        return
    if args.only == 'served':
        print(json.dumps(served(args, ctx, workloads)), flush=True)
        return
    if args.only == 'c0':
        print(json.dumps(c0(ctx)), flush=True)
        return
    if args.only == 'build':
        print(json.dumps(builder(ctx)), flush=True)
        return
    if args.only == 'textures':
        print(json.dumps(workloads.textured_form(ctx, W, H, log=log)), flush=True)
        return
    t0 = time.time()
    scene, camera, opts = scenes.plain_gpu_scene(0, W, H)
    build_times = []

    def timed_build(tris, count):
        t = time.time()
        r = ctx.build_bvh(tris, count)
        build_times.append(time.time() - t)
        return r
    tc = time.time()
    sc = compile_scene(scene, timed_build)
    compile_host_s = time.time() - tc
    # compileGeometry on the device (flatten, build, BVH-order gather, tangents) on the same scene.
    from meercadraytracer_amd.compile import compile_scene_device, compile_scene_resident
    compile_scene_device(scene, ctx)  # allocations and first-touch
    t1 = time.time()
    dev_sc = compile_scene_device(scene, ctx)
    compile_device_s = time.time() - t1
    # ... then made resident from the downloaded arrays, against compiling straight into a scene.
    tmp = native.Scene(ctx)
    t1 = time.time()
    tmp.set_geometry(dev_sc.geometry, 1)
    tmp.set_shading(dev_sc.shading, 1)
    upload_s = time.time() - t1
    compile_scene_resident(scene, tmp, 2)  # first-touch
    t1 = time.time()
    compile_scene_resident(scene, tmp, 3)
    resident_s = time.time() - t1
    tmp.release()
    del dev_sc
    cam = compile_camera(camera)
    st = trace_settings(opts)
    log(f'compiled {sc.geometry.triangleCount} triangles, {sc.geometry.bvh.nodeCount} nodes in '
        f'{time.time() - t0:.1f} s (GPU build {build_times[0] * 1e3:.1f} ms incl. copies)')
    gpu = native.Scene(ctx)
    gpu.set_geometry(sc.geometry, 1)
    gpu.set_shading(sc.shading, 1)
    npx = W * H
    data = native.DeviceBuffer(ctx, max(args.steps, args.warmup, 1) * npx * 16)
    acc = native.DeviceBuffer(ctx, npx * 16)
    image = native.DeviceBuffer(ctx, npx * 4)
    acc.zero()
    tile = Tile(0, 0, W, H)

    def run(passes, cnt):
        # All the passes stream through the wavefront in one device run; each pass's frame is
        # then summed and encoded in pass order (writeTile).
        render_passes(gpu, cam, st, W, H, passes, data.ptr, cnt)
        for i, p in enumerate(passes):
            native.encode_tile_device(ctx, image.ptr, W, tile, data.ptr + i * npx * 16, acc.ptr, p + 1, True)

    ctx.set_timing(True)
    if args.warmup:
        run(range(args.warmup), abi.Counters())
        log(f'warmup: {args.warmup} passes done')
    cnt = abi.Counters()
    t_start = time.perf_counter()
    run(range(args.warmup, args.warmup + args.steps), cnt)
    elapsed = time.perf_counter() - t_start
    timing = ctx.last_timing()
    log(f'timed: {args.steps} passes in {elapsed:.3f} s')
    c = cnt.as_dict()
    rays = c['closest_queries'] + c['any_queries']
    value = rays / elapsed / 1e6
    kern = {
        'extend (closest hit)': (timing['closest_ms'], timing['closest_launches'],
                                 c['closest_node_visits'] * NODE_BYTES + c['closest_triangle_tests'] * TRI_BYTES),
        'shadow (any hit)': (timing['any_ms'], timing['any_launches'],
                             c['any_node_visits'] * NODE_BYTES + c['any_triangle_tests'] * TRI_BYTES),
    }
    name = max(kern, key=lambda k: kern[k][0])
    ms, launches, algo = kern[name]
    achieved = algo / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': None, 'kernel': name,
                'avg_launch_ms': round(ms / max(1, launches), 4),
                'algorithmic_bytes_per_launch': int(algo / max(1, launches))}
    result = {
        'metric': METRIC, 'value': round(value, 2), 'unit': 'Mrays/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(elapsed / args.steps * 1e3, 2), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic (seeded scene generator)',
        'config': {'workload': 'plain_gpu: 724x724-quad fBm heightfield + 64 icospheres (1,376,032 triangles, '
                               '20% glass), sun + 4 point lights, 1920x1080, GI on, 8 bounces, ACES; one step = '
                               'one 1-spp pass of the 64-spp frame, scene resident in HBM',
                   'triangles': int(sc.geometry.triangleCount), 'width': W, 'height': H, 'spp_per_step': 1,
                   'frame_spp': 64, 'paths_per_step': npx,
                   'rays_per_step': int(rays / args.steps),
                   'kernel_ms': {'extend': round(timing['closest_ms'] / args.steps, 2),
                                 'shadow': round(timing['any_ms'] / args.steps, 2),
                                 'shade': round(timing['shade_ms'] / args.steps, 2),
                                 'device_total': round(timing['total_ms'] / args.steps, 2)},
                   'bvh_build_ms_incl_copies': round(build_times[0] * 1e3, 1),
                   'compile_geometry_ms': {'host_mirror_with_gpu_build': round(compile_host_s * 1e3, 1),
                                           'device': round(compile_device_s * 1e3, 1),
                                           'device_then_upload': round((compile_device_s + upload_s) * 1e3, 1),
                                           'resident': round(resident_s * 1e3, 1)}},
        'roofline': roofline,
    }
    if args.shade_sort == 'on':
        lanes = c['shade_class_lanes']
        result['config']['shade_sort'] = {
            'class_lanes': lanes,
            'class_mixed_wave_frac': round(c['class_mixed_waves'] / max(1, sum(lanes) / 64), 4)}
    if not args.no_traffic:
        log('HBM traffic passes (rocprofv3 FETCH_SIZE / WRITE_SIZE) ...')
        try:
            t = measure_traffic(args, name, launches)
        except Exception as e:  # the counters are a report, never a reason to fail the bench
            log(f'traffic measurement failed: {e}')
            t = None
        if t:
            roofline['traffic'] = t['bytes_per_launch']
            roofline['traffic_detail'] = t
    for b in (data, acc, image):
        b.free()
    if args.only == 'plain':  # the plain workload alone (and the CPU baseline unless --no-cpu)
        args.no_precision = args.no_textures = args.no_batched = args.no_served = args.no_c0 = True
    if not args.no_precision or args.only == 'precision':
        log('precision form (1M probe rays, FP64 binary vs FP32 wide) ...')
        result['precision'] = workloads.precision_form(ctx, gpu, sc.geometry, cam, W, H, settings=st, log=log)
        if args.only == 'precision':
            print(json.dumps(result['precision']), flush=True)
            return
    gpu.release()
    if not args.no_textures:
        log('textured workload ...')
        result['textures'] = workloads.textured_form(ctx, W, H, plain_mrays=value, log=log)
    if not args.no_batched:
        log('batched form ...')
        result['batched'] = workloads.batched_form(ctx, log=log)
    if not args.no_served:
        log('served mix ...')
        result['served'] = served(args, ctx, workloads)
    if not args.no_c0:
        log('C0 (the CPU configuration) on the GPU and on the host ...')
        result['c0'] = c0(ctx)
    if not args.no_cpu:
        log('cpu baseline ...')
        result['cpu_baseline'] = cpu_baseline(sc, cam, st, W, H, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
