#!/usr/bin/env python3
"""Benchmark: TOF single scatter simulation on one MI355X.

One step = one full scatter estimate of the workload with the mu-map, scatter
points and activity image resident in HBM: trace every scatter-point ->
detector half path (table build), refresh the emission terms, fused scatter
sum over (scatter point x LOR x TOF bin) in FP32.  Prints one JSON
line with throughput in scatter-point x LOR x TOF-bin evaluations per second,
p50/p99 step latency, the roofline of the dominant kernel (the fused scatter
sum, FP32 VALU), the error against an FP64 run of the same kernels, and a CPU
baseline (the float64 oracle timed on a bounded scatter-point sample).
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("scatter-pt×LOR×TOF-bin evals/s; p50/p99 latency at ≤1e-4 rel err vs FP64")
FP32_VALU_PEAK_TFLOPS = 157.3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=10)
    p.add_argument('--warmup', type=int, default=2)
    p.add_argument('--workload', default='config2',
                   help="config1, config2, small, or randomN (the uneven random instance of seed N)")
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--no-fp64-check', action='store_true')
    p.add_argument('--cpu-threads', type=int, default=0)
    p.add_argument('--no-stages', action='store_true', help='skip the full-size sinogram stages (config 2)')
    p.add_argument('--default-batch', type=int, default=64, metavar='N',
                   help='the default line\'s batched reading: N uneven random instances (0: none)')
    p.add_argument('--default-serve', type=float, default=120.0, metavar='SECONDS',
                   help='the default line\'s served reading: a Poisson stream of this length (0: none)')
    p.add_argument('--calibrate-copy', action='store_true',
                   help='run only a 1 GiB device copy (3 launches) to scale FETCH_SIZE/WRITE_SIZE')
    p.add_argument('--batch', type=int, default=0, metavar='N',
                   help='instead of the default line: N uneven random instances (seeds 0..N-1), '
                        'the one-instance path against the batched path')
    p.add_argument('--streams', type=int, default=4, help='contexts of the served path (and of --batch)')
    p.add_argument('--batch-streams', type=int, default=2, help='contexts of the default line\'s batched reading')
    p.add_argument('--batch-kind', default='random', choices=['random', 'config1'],
                   help='uneven random instances (seeds 0..N-1) or config 1 instances of N seeds')
    p.add_argument('--serve', type=float, default=0.0, metavar='SECONDS',
                   help='instead of the default line: an open-loop Poisson request stream of this '
                        'length over resident subjects (throughput, p50/p99 latency)')
    p.add_argument('--rate', type=float, default=50.0, help='requests per second of --serve')
    p.add_argument('--subjects', type=int, default=16, help='subjects of --serve (uneven random instances)')
    p.add_argument('--subject-kind', default='random', choices=['random', 'config1'],
                   help='subjects of --serve: uneven random instances or config 1 instances')
    p.add_argument('--autotune', action='store_true',
                   help='--serve: pick each subject\'s tile shape by timing when it is loaded')
    p.add_argument('--noise-fields', type=int, default=3,
                   help='precomputed 5%% noise fields per subject that the requests cycle through')
    return p.parse_args()


def instance_evals(w):
    nru, ndu = w.n_rings_used, w.n_det_used
    lors = nru * (ndu * (ndu - 1) // 2) + nru * (nru - 1) * (ndu * (ndu + 1) // 2)
    return len(w.scatter_points) * lors * w.n_bins


def batch_reading(kind, n, streams, fp64_sample=4):
    """Batched reading (config 4 with kind 'random': n uneven instances of
    seeds 0..n-1): the one-instance path (create, trace, scatter, download,
    destroy, one at a time) against opensss_amd.batch.run_batch (several in
    flight on `streams` contexts, device blocks reused); outputs compared bit
    for bit, and the FP32 outputs of `fp64_sample` instances against FP64
    runs."""
    from opensss_amd import _native, phantoms
    from opensss_amd.batch import run_batch, run_plain
    water = np.load(os.path.join(ROOT, 'opensss_amd', 'data', 'attenuation_water.npy'))
    if kind == 'config1':
        ws = [phantoms.config1(seed=seed) for seed in range(n)]
    else:
        ws = [phantoms.random_instance(seed) for seed in range(n)]
    evals = sum(instance_evals(w) for w in ws)
    print(f'[bench] batch of {len(ws)} {kind} instances, {evals:.3e} evals', file=sys.stderr, flush=True)
    run_plain(ws[:1], water)   # library and device warm-up
    t0 = time.perf_counter()
    plain = run_plain(ws, water)
    t_plain = time.perf_counter() - t0
    print(f'[bench] one-instance path {t_plain:.3f} s', file=sys.stderr, flush=True)
    ctxs = [_native.Context(0) for _ in range(max(1, streams))]
    run_batch(ws[:2 * len(ctxs)], water, contexts=ctxs)   # contexts and their caches warm, as the default one
    t0 = time.perf_counter()
    batched = run_batch(ws, water, contexts=ctxs)
    t_batch = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    print(f'[bench] batched path {t_batch:.3f} s', file=sys.stderr, flush=True)
    same = all(np.array_equal(a, b) for a, b in zip(plain, batched))
    graph = graph_reading(ws, water, plain, evals, t_plain)
    errs = []
    for i in range(0, len(ws), max(1, len(ws) // max(1, fp64_sample)))[:fp64_sample]:
        ref = run_plain([ws[i]], water, precision='fp64')[0]
        if np.abs(ref).max() > 0:
            errs.append(rel_err_stats(batched[i], ref))
    return {
        'workload': f'{kind}-instances-0..{len(ws) - 1}', 'instances': len(ws), 'evals': int(evals),
        'streams': streams,
        'value': evals / t_batch, 'unit': 'evals/s',
        'one_instance_path': {'value': evals / t_plain, 'seconds': t_plain},
        'batched_path': {'value': evals / t_batch, 'seconds': t_batch},
        'speedup': t_plain / t_batch, 'outputs_identical': bool(same),
        'graph_batch': graph,
        'max_rel_err_vs_fp64': max(e['max_rel_err_cells_above_1pct_peak'] for e in errs) if errs else None,
        'fp64_checked_instances': len(errs),
    }


def graph_reading(ws, water, plain, evals, t_plain, iterations=3):
    """The graph batch (opensss_amd.batch.run_graph_batch) on the same
    instances: resident groups under 60% of the free device memory, one
    persistent scatter launch per kernel variant per group, replayed as a
    HIP graph.  Launch 1 takes the instances' own images (outputs compared
    bit for bit with the one-instance path); launches 2..`iterations` take
    new images (activity x (1 + 0.05 N(0,1)), the repeated estimates of a
    reconstruction), timed against the one-instance path run once per
    image."""
    from opensss_amd import _native
    from opensss_amd.batch import run_graph_batch
    rng = np.random.default_rng(11)
    fields = {}

    def image(k, j):
        if j == 0:
            return ws[k].activity
        if (k, j) not in fields:
            fields[(k, j)] = ws[k].activity * np.clip(1.0 + 0.05 * rng.standard_normal(ws[k].activity.shape), 0.0, None)
        return fields[(k, j)]

    ctx = _native.Context(0)
    try:
        run_graph_batch(ws[:2], water, context=ctx)   # warm: kernels, graph code paths
        timings = []
        t0 = time.perf_counter()
        got = run_graph_batch(ws, water, context=ctx, iterations=iterations, activity=image, timings=timings)
        wall = time.perf_counter() - t0
    finally:
        ctx.close()
    # the outputs are the last launch's: compare them with the one-instance
    # path on those images for a few instances (launch 1 equals `plain` by
    # the same kernels; tests/test_gpu_graph_batch.py checks every launch)
    from opensss_amd.batch import run_plain
    import dataclasses
    check = list(range(0, len(ws), max(1, len(ws) // 4)))[:4]
    same = all(np.array_equal(got[k], run_plain([dataclasses.replace(ws[k], activity=image(k, iterations - 1))],
                                                water)[0]) for k in check)
    first = sum(t['launch_ms'][0] for t in timings) / 1e3
    later = sum(sum(t['launch_ms'][1:]) for t in timings) / 1e3
    setup = sum(t['setup_s'] for t in timings)
    per_image = later / max(1, iterations - 1)
    print(f'[bench] graph batch: {len(timings)} resident groups, set-up {setup:.3f} s, first launches '
          f'{first:.3f} s, later launches {per_image:.3f} s per image set', file=sys.stderr, flush=True)
    return {
        'groups': len(timings), 'iterations': iterations, 'wall_seconds': wall, 'setup_seconds': setup,
        'first_launch_seconds': first, 'first_launch_value': evals / first,
        'reestimate_seconds_per_image_set': per_image, 'reestimate_value': evals / per_image,
        'reestimate_ratio_to_one_instance_path': t_plain / per_image,
        'one_shot_ratio_to_one_instance_path': t_plain / (setup + first),
        'outputs_identical_to_one_instance_path': bool(same), 'checked_instances': len(check),
    }


def config1_graph_reading(n=64, iterations=5):
    """n config-1 subjects (one-bin, non-TOF; seeds 0..n-1) in one graph
    batch: per image set, the graph launch against the one-instance path."""
    from opensss_amd import _native, phantoms
    from opensss_amd.batch import GraphBatch, _subject, run_plain
    water = np.load(os.path.join(ROOT, 'opensss_amd', 'data', 'attenuation_water.npy'))
    ws = [phantoms.config1(seed=seed) for seed in range(n)]
    evals = sum(instance_evals(w) for w in ws)
    run_plain(ws[:1], water)
    t0 = time.perf_counter()
    plain = run_plain(ws, water)
    t_plain = time.perf_counter() - t0
    ctx = _native.Context(0)
    subs, batch = [], None
    try:
        for w in ws:
            s = _subject(w, water, ctx, 'fp32')
            subs.append(s)
            s.set_activity(w.activity)
            s.trace()
        batch = GraphBatch(subs, ctx)
        ms = []
        for _ in range(max(2, iterations)):
            batch.launch(graph=True)
            ms.append(batch.elapsed_ms())
        same = all(np.array_equal(batch.output(i), plain[i]) for i in range(n))
        info = batch.info
    finally:
        if batch is not None:
            batch.close()
        for s in subs:
            s.close()
        ctx.close()
    launch = float(np.median(ms[1:])) / 1e3
    print(f'[bench] {n} config-1 subjects: one-instance path {t_plain:.4f} s, graph launch {launch * 1e3:.2f} ms',
          file=sys.stderr, flush=True)
    return {
        'workload': f'config1-subjects-0..{n - 1}', 'subjects': n, 'evals': int(evals),
        'one_instance_path': {'value': evals / t_plain, 'seconds': t_plain},
        'graph_launch': {'value': evals / launch, 'seconds': launch, 'scatter_launches': info['scatter_launches']},
        'ratio_to_one_instance_path': t_plain / launch, 'outputs_identical': bool(same),
    }


def batch_bench(args):
    r = batch_reading(args.batch_kind, args.batch, args.streams)
    line = {'metric': 'batched scatter-pt x LOR x TOF-bin evals/s over uneven instances', 'n_gpus': 1,
            'higher_is_better': True, 'dtype': 'f32', 'data': 'synthetic', **r}
    print(json.dumps(line))


def serve_reading(kind, n_subjects, rate, seconds, streams, autotune=False, noise_fields=3, fp64_sample=4):
    """Served reading (config 5 with kind 'random'): Poisson arrivals at
    `rate` for `seconds`, each request a subject (uniform over `n_subjects`)
    and its activity times (1 + 0.05 N(0,1)) per voxel.  Subjects are loaded
    (created and traced) before the stream starts and stay resident.
    Reports completed evaluations and requests per second over the stream,
    request latency (arrival to result in host memory) p50/p99, the
    one-instance path's rate on the first requests, and the FP32 error
    against FP64 runs of sampled requests."""
    from opensss_amd import phantoms
    from opensss_amd.batch import run_plain
    from opensss_amd.serve import Request, ScatterServer, serve
    water = np.load(os.path.join(ROOT, 'opensss_amd', 'data', 'attenuation_water.npy'))
    rng = np.random.default_rng(2024)
    if kind == 'config1':
        subs = {k: phantoms.config1(seed=k) for k in range(n_subjects)}
    else:
        subs = {k: phantoms.random_instance(k) for k in range(n_subjects)}
    fields = {k: [w.activity * (1.0 + 0.05 * rng.standard_normal(w.activity.shape))
                  for _ in range(max(1, noise_fields))] for k, w in subs.items()}
    n_req = int(rng.poisson(rate * seconds))
    arrivals = np.sort(rng.uniform(0.0, seconds, size=n_req))
    keys = rng.integers(0, n_subjects, size=n_req)
    reqs = [Request(key=int(k), activity=fields[int(k)][i % len(fields[int(k)])], arrival=float(t), tag=i)
            for i, (t, k) in enumerate(zip(arrivals, keys))]
    evals = {k: instance_evals(w) for k, w in subs.items()}
    print(f'[bench] serve: {n_req} requests over {seconds:.0f} s, {n_subjects} subjects', file=sys.stderr,
          flush=True)

    server = ScatterServer(subs, water, n_contexts=streams, autotune=autotune)
    t0 = time.perf_counter()
    for k in subs:
        server.load(k)
    for c in server.ctxs:
        c.synchronize()
    t_load = time.perf_counter() - t0
    # warm-up: one request per subject (not timed)
    serve(server, [Request(key=k, activity=subs[k].activity) for k in subs], on_result=lambda r, v: None)
    sample = set(range(0, n_req, max(1, n_req // max(1, fp64_sample))))
    kept = {}

    def on_result(req, view):
        if req.tag in sample:
            kept[req.tag] = view.copy()

    serve(server, reqs, on_result=on_result)
    t_end = max(r.done for r in reqs)
    server_stats = dict(server.stats)
    resident = server.resident_bytes()
    server.close()
    lat = np.array([r.done - r.arrival for r in reqs])
    total = sum(evals[r.key] for r in reqs)
    served_rate = total / t_end
    print(f'[bench] served {n_req / t_end:.1f} req/s, p99 {np.percentile(lat, 99) * 1e3:.0f} ms', file=sys.stderr,
          flush=True)
    # the one-instance path (create, trace, scatter, download, destroy) on the first requests
    head = reqs[:min(16, n_req)]
    plain_inst = [dataclasses.replace(subs[r.key], activity=r.activity) for r in head]
    run_plain(plain_inst[:1], water)
    t0 = time.perf_counter()
    run_plain(plain_inst, water)
    t_plain = time.perf_counter() - t0
    plain_rate = sum(evals[r.key] for r in head) / t_plain
    # FP32 against FP64 on the sampled requests
    errs = []
    for i in sorted(kept):
        r = reqs[i]
        w = dataclasses.replace(subs[r.key], activity=r.activity)
        ref = run_plain([w], water, precision='fp64')[0]
        if np.abs(ref).max() > 0:
            errs.append(rel_err_stats(kept[i], ref))
    saturated = bool(n_req / t_end < 0.95 * rate)
    return {
        'workload': f'{kind}-subjects-0..{n_subjects - 1}', 'requests': n_req, 'offered_rate_per_s': rate,
        'seconds': seconds, 'contexts': streams, 'autotune': bool(autotune),
        'noise': '5% multiplicative per voxel',
        'value': served_rate, 'unit': 'evals/s',
        'completed_per_s': n_req / t_end,
        'latency_ms': {'p50': float(np.percentile(lat, 50) * 1e3), 'p99': float(np.percentile(lat, 99) * 1e3),
                       'max': float(lat.max() * 1e3)},
        'subject_load_s': t_load, 'resident_bytes': int(resident), 'server': server_stats,
        'one_instance_path': {'value': plain_rate, 'requests': len(head), 'seconds': t_plain},
        # below saturation the served rate is the offered load, not a capacity
        'saturated': saturated,
        'ratio_to_one_instance_path': served_rate / plain_rate,
        'max_rel_err_vs_fp64': max(e['max_rel_err_cells_above_1pct_peak'] for e in errs) if errs else None,
        'mean_rel_err_vs_fp64': float(np.mean([e['mean_rel_err_cells_above_1pct_peak'] for e in errs]))
        if errs else None,
    }


def serve_bench(args):
    r = serve_reading(args.subject_kind, args.subjects, args.rate, args.serve, args.streams, args.autotune,
                      args.noise_fields)
    line = {'metric': 'served scatter-pt x LOR x TOF-bin evals/s; p50/p99 request latency', 'n_gpus': 1,
            'higher_is_better': True, 'dtype': 'f32', 'data': 'synthetic', **r}
    print(json.dumps(line))


def calibrate_copy():
    """Reference kernel for the HBM counters: 1 GiB read + 1 GiB written per launch."""
    import torch
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({'calibration': 'torch.mul copy', 'bytes_read': n * 4, 'bytes_written': n * 4,
                      'launches': 3}))


def make_workload(name):
    from opensss_amd import phantoms
    if name == 'config1':
        return phantoms.config1()
    if name == 'small':
        return phantoms.small_tof()
    if name.startswith('random'):
        return phantoms.random_instance(int(name[len('random'):]))
    if name != 'config2':
        raise SystemExit(f'unknown workload {name!r}')
    return phantoms.config2()


def subject(w, precision, points=None):
    from opensss_amd.sss import ScatterSubject
    water = np.load(os.path.join(ROOT, 'opensss_amd', 'data', 'attenuation_water.npy'))
    return ScatterSubject(w.attenuation, w.image_size, w.geometry, w.normals, w.det_size, water,
                          w.energy_resolution, w.energy_threshold, w.n_rings_used, w.n_det_used,
                          w.tof_resolution, w.tof_range, w.n_bins,
                          scatter_points=w.scatter_points if points is None else points,
                          precision=precision)


def profile_traffic(workload):
    """HBM bytes per k_scatter launch from the committed rocprofv3 summary of
    this workload (calibrated FETCH_SIZE + WRITE_SIZE), or (None, None)."""
    import re
    path = os.path.join(ROOT, 'profiles', f'k_scatter_{workload}.md')
    if not os.path.exists(path):
        return None, None
    m = re.search(r'HBM traffic per launch \(bytes\):\s*([0-9.eE+]+)', open(path).read())
    return (float(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


def rel_err_stats(got, want, frac=0.01):
    """Max and mean relative error on cells above frac of the peak, max
    absolute error over the peak."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    peak = np.abs(want).max()
    big = np.abs(want) >= frac * peak
    rel = np.abs(got[big] - want[big]) / np.abs(want[big])
    return {'max_rel_err_cells_above_1pct_peak': float(rel.max()),
            'mean_rel_err_cells_above_1pct_peak': float(rel.mean()),
            'max_abs_err_over_peak': float(np.abs(got - want).max() / peak)}


def cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_threads():
    """Host threads of the CPU baseline: the process's CPU share as the
    environment states it (OMP_NUM_THREADS), else every CPU."""
    env = os.environ.get('OMP_NUM_THREADS', '')
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_baseline(w, threads, target_s=12.0):
    """Oracle (float64 C restatement) on the first k scatter points of each of
    the first `threads` z slices (OpenMP runs one slice per thread); k is
    sized from a one-point-per-slice run so the timed run takes about
    target_s seconds.  evals/s scaled by the sample's evaluation count.
    Returns the baseline and the oracle's compact result for the sample, with
    the sample's points, to check the GPU path against."""
    from oracle import oracle as ora
    from opensss_amd.tables import detectors_used, group_by_slice, rings_used
    # -O3 -march=native, built here for this host's CPU
    ora.use_native_build()
    pts, ptr, _ = group_by_slice(w.scatter_points)
    n_sl = min(threads, len(ptr) - 1)
    water = np.load(os.path.join(ROOT, 'opensss_amd', 'data', 'attenuation_water.npy'))
    nru, ndu = w.n_rings_used, w.n_det_used
    lors = nru * (ndu * (ndu - 1) // 2) + nru * (nru - 1) * (ndu * (ndu + 1) // 2)

    dets = detectors_used(w.geometry.shape[1], ndu)

    def timed(per_slice):
        take = [ptr[i] + j for i in range(n_sl) for j in range(min(per_slice, ptr[i + 1] - ptr[i]))]
        sample = pts[take]
        sp, sptr, _ = group_by_slice(sample)
        t0 = time.perf_counter()
        out, cell_of = ora.sss(w.activity, w.attenuation, w.image_size, w.geometry, w.lor_coordinates,
                               w.normals, w.det_size, water, w.energy_resolution, w.energy_threshold,
                               rings_used(w.geometry.shape[0], nru), dets, w.tof_range, w.tof_resolution,
                               w.n_bins, sp, sptr, n_threads=threads, dense=False)
        return sample, time.perf_counter() - t0, out, cell_of

    _, t1, _, _ = timed(1)
    per_slice = max(1, int(round(target_s / max(t1, 1e-3))))
    sample, dt, out, cell_of = timed(per_slice)
    n = len(sample)
    evals = n * lors * w.n_bins
    # oracle result in the GPU's compact layout (B, NRU^2, NDU(NDU+1)/2)
    N = w.geometry.shape[1]
    i1, i2 = np.triu_indices(len(dets))
    ang = w.lor_coordinates[dets[i1], dets[i2], 0].astype(np.int64)
    rad = w.lor_coordinates[dets[i1], dets[i2], 1].astype(np.int64)
    compact = out[:, :, cell_of[rad * (N // 2) + ang]]
    base = {'value': evals / dt, 'unit': 'evals/s', 'cores': threads, 'kind': 'port',
            'host_cpus': os.cpu_count(), 'cpu_model': cpu_model(),
            'sample': f'{n} scatter points (the first {per_slice} of each of {n_sl} z slices) of {w.name}, '
                      f'{evals:.3e} evals in {dt:.2f} s, float64 C oracle (-O3 -march=native), '
                      f'OpenMP over slices'}
    return base, sample, compact


def sinogram_stages(w, out32, out64, oracle_pair=None, accel=2):
    """Config 2's stages after the scatter estimate, at full size on the GPU:
    interpolation of the 15-bin estimate to the 761 x 380 x 159 sinogram
    (span 19, MRD 9), the span tail mask and trues on the 159-ring extended
    geometry, and the span tail fit of every TOF bin on device-resident
    sinograms (AccelerationFactor `accel`).  Prompts per bin: trues / B +
    k_true S64 + uniform randoms (noiseless), S64 the FP64 estimate's
    interpolation; the FP32 estimate's fitted scale is reported against the
    FP64 one's (config 3's tail-scale error).  oracle_pair: (GPU FP32
    estimate, oracle estimate) of one scatter-point sample, whose fits are
    compared the same way.  Times from the host clock around synchronised
    calls."""
    from opensss_amd import _native
    from opensss_amd.backprojection import fit_geometry_buffers, span_entries, tail_fit_device
    from opensss_amd.data_reduction import SinogramLUT
    from opensss_amd.geometry import expandRings
    from opensss_amd.interpolation import Interpolator
    from opensss_amd.prompts import ForwardProject
    from opensss_amd.scaling import scale_to_prompts_device
    from opensss_amd.tails_mask import MaskGeneratorMashing
    ctx = _native.default_context()
    lib = _native.load_library()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    B = out32.shape[0]
    msi = SinogramLUT(NR, 19, 9)[0]
    geom_ext, _ = expandRings(w.geometry, w.normals)
    ip = Interpolator(NR, w.n_rings_used, N, w.n_det_used, w.lor_coordinates, msi, True, context=ctx)
    n_out = ip.n_cells_out * ip.n_out_planes
    bufs = []

    def dev(a):
        b = _native.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a))
        bufs.append(b)
        return b

    def interpolate(compact):
        src = dev(np.asarray(compact, dtype=np.float32))
        out = _native.DeviceBuffer(ctx, 4 * n_out * B)
        tot = _native.DeviceBuffer(ctx, 4 * n_out)
        bufs.extend([out, tot])
        times = []
        for _ in range(4):
            t0 = time.perf_counter()
            ip.run_device(src.ptr.value, B, out.ptr.value, tot.ptr.value)
            ctx.synchronize()
            times.append(time.perf_counter() - t0)
        return out, min(times[1:])

    def bin_host(buf, b):
        host = np.empty(ip.shape_out, np.float32)
        _native.check(lib.sss_buffer_download(ctx.handle, host.ctypes.data_as(_native.ctypes.c_void_p),
                                              _native.ctypes.c_void_p(buf.ptr.value + 4 * n_out * b), host.nbytes),
                      'sss_buffer_download')
        return host

    try:
        i32, t_interp = interpolate(out32)
        i64, _ = interpolate(out64.astype(np.float32))
        t0 = time.perf_counter()
        mask = MaskGeneratorMashing(w.activity, w.attenuation, w.image_size, geom_ext, w.lor_coordinates, msi,
                                    True, 1)
        t_mask = time.perf_counter() - t0
        t0 = time.perf_counter()
        trues = ForwardProject(w.activity, w.attenuation, w.image_size, geom_ext, w.lor_coordinates,
                               MashedSinogramIndices=msi)
        t_fp = time.perf_counter() - t0
        M = dev(np.float32(mask >= 1))
        ent = span_entries(msi)
        gb = fit_geometry_buffers(geom_ext, w.lor_coordinates, ent, context=ctx)
        bufs.extend(gb)
        sums, peaks = [], []
        for b in range(B):
            h = bin_host(i64, b)
            sums.append(float(h.sum()))
            peaks.append(float(h.max()))
        k_true = 0.3 * trues.sum() / max(1e-30, sum(sums))
        # trues and randoms spread evenly over the TOF bins; bins whose scatter
        # peaks below 1% of the estimate's peak are not fitted (their tails
        # hold almost no scatter: the same 1% line as the sinogram errors)
        r_val = np.float32(0.1 * trues.mean() / B)
        R = dev(np.full(ip.shape_out, r_val, np.float32))
        fits, t_fit, p_host = {}, [], []
        for b in range(B):
            if peaks[b] < 0.01 * max(peaks):
                continue
            S64 = bin_host(i64, b)
            p_host.append((trues / B + k_true * S64 + r_val).astype(np.float32))
            P = dev(p_host[-1])
            S64d = dev(S64)
            t0 = time.perf_counter()
            f32 = tail_fit_device(w.attenuation.shape, w.image_size, geom_ext, w.lor_coordinates, ent,
                                  msi.shape[0], accel, True, i32.ptr.value + 4 * n_out * b, P.ptr.value,
                                  R.ptr.value, M.ptr.value, context=ctx, geometry_buffers=gb)[0]
            t_fit.append(time.perf_counter() - t0)
            f64 = tail_fit_device(w.attenuation.shape, w.image_size, geom_ext, w.lor_coordinates, ent,
                                  msi.shape[0], accel, True, S64d.ptr.value, P.ptr.value, R.ptr.value, M.ptr.value,
                                  context=ctx, geometry_buffers=gb)[0]
            fits[b] = (f32, f64)
            for x in (P, S64d):
                x.free()
        fitted = sorted(fits)
        scale_err = [abs(fits[b][0] - fits[b][1]) / abs(fits[b][1]) for b in fitted]
        # every fitted bin in one call (each LOR traced once per 8 bins)
        Pall = dev(np.stack(p_host, axis=-1))
        del p_host
        Sall = dev(np.stack([bin_host(i32, b) for b in fitted]))
        Mu8 = dev((mask >= 1).astype(np.uint8))
        sums_all = _native.DeviceBuffer(ctx, 16 * len(fitted))
        bufs.append(sums_all)
        t_all = []
        for _ in range(2):
            t0 = time.perf_counter()
            scale_to_prompts_device(w.attenuation.shape, w.image_size, geom_ext, w.lor_coordinates, ent,
                                    msi.shape[0], accel, True, len(fitted), Sall.ptr.value, Pall.ptr.value, False,
                                    Mu8.ptr.value, sums_all.ptr.value, randoms_ptr=R.ptr.value,
                                    geometry_buffers=gb, context=ctx)
            ctx.synchronize()
            t_all.append(time.perf_counter() - t0)
        sa = sums_all.download(np.empty((len(fitted), 2), np.float64))
        all_scales = [sa[i, 0] / sa[i, 1] if sa[i, 1] != 0 else 0.0 for i in range(len(fitted))]
        for x in (Pall, Sall, Mu8):
            x.free()
        res = {
            'workload': f'{w.name} -> 761x380x159 span 19 MRD 9, {B} bins',
            'interpolate_ms': t_interp * 1e3,
            'interpolate_GBps': (B + 1) * n_out * 4 / t_interp / 1e9,
            'tail_mask_ms': t_mask * 1e3, 'forward_project_ms': t_fp * 1e3,
            'tail_fit_ms_per_bin': float(np.median(t_fit) * 1e3), 'fit_accel': accel,
            'tail_fit_all_bins_ms': float(min(t_all) * 1e3),
            'tail_fit_all_bins_speedup': float(sum(t_fit) / min(t_all)),
            'tail_fit_all_bins_identical': bool(all(all_scales[i] == fits[b][0] for i, b in enumerate(fitted))),
            'tail_scale_max_rel_err_fp32_vs_fp64': float(max(scale_err)),
            'tail_scale_k_true': float(k_true),
            'tail_scale_fp64_max_rel_err_vs_k_true': float(max(abs(fits[b][1] - k_true) / k_true for b in fitted)),
            'fitted_bins': fitted,
            'tail_scales_fp32': [float(fits[b][0]) for b in fitted],
            'host_clock': True,
        }
        if oracle_pair is not None:
            g32, ora = oracle_pair
            j32, _ = interpolate(g32)
            jora, _ = interpolate(ora)
            b = B // 2
            So = bin_host(jora, b)
            # the sample's randoms level: a tenth of its scatter's mean
            r_o = np.float32(0.1 * k_true * So.mean())
            P = dev((trues / B + k_true * So + r_o).astype(np.float32))
            Ro = dev(np.full(ip.shape_out, r_o, np.float32))
            Sod = dev(So)
            fa = tail_fit_device(w.attenuation.shape, w.image_size, geom_ext, w.lor_coordinates, ent, msi.shape[0],
                                 accel, True, j32.ptr.value + 4 * n_out * b, P.ptr.value, Ro.ptr.value, M.ptr.value,
                                 context=ctx, geometry_buffers=gb)[0]
            fo = tail_fit_device(w.attenuation.shape, w.image_size, geom_ext, w.lor_coordinates, ent, msi.shape[0],
                                 accel, True, Sod.ptr.value, P.ptr.value, Ro.ptr.value, M.ptr.value, context=ctx,
                                 geometry_buffers=gb)[0]
            res['tail_scale_rel_err_vs_oracle'] = abs(fa - fo) / abs(fo)
            res['tail_scale_oracle_sample'] = f'bin {b}, the cpu_baseline sample\'s estimates'
        scales = np.zeros(B, np.float32)
        for i, b in enumerate(fitted):
            scales[b] = all_scales[i]
        res['listmode'] = listmode_leg(ctx, w, SinogramLUT(NR, 19, 9)[2], i32.ptr.value, B, scales)
        res['image_preparation'] = image_leg(ctx, w)
        return res
    finally:
        for b in bufs:
            b.free()


def image_leg(ctx, w, ct_shape=(512, 512, 300), ct_voxel=(0.9765625, 0.9765625, 1.0), seed=8):
    """CropAndDownscale of a CT-sized float32 volume (512 x 512 x 300 of
    ~1 mm) to the workload's grid over its field of view (mm), host arrays in
    and out (the upload of the volume included); best of three."""
    from opensss_amd import images
    rng = np.random.default_rng(seed)
    # x fastest, as LoadImage's transposed view of the interfile data
    ct = np.asfortranarray(rng.uniform(0.0, 0.1, size=ct_shape).astype(np.float32))
    size_mm = [10.0 * (w.image_size[k + 3] - w.image_size[k]) for k in range(3)]
    dims = list(w.attenuation.shape)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        images.CropAndDownscale(ct, ct_voxel, size_mm, dims, zero_below=0.001, context=ctx)
        ts.append(time.perf_counter() - t0)
    return {'workload': f'{ct_shape} float32 of {ct_voxel} mm -> {tuple(dims)} over {[round(v) for v in size_mm]} mm',
            'ms': min(ts) * 1e3, 'input_bytes': int(ct.nbytes), 'host_clock': True}


def listmode_leg(ctx, w, lut, scatter_ptr, n_bins, scales, n_events=50_000_000, seed=7):
    """CASToR list mode at the 760 x 80 scanner: n_events synthetic events
    (ring differences within +-9, opposite detectors +-N/4, TOF normal with
    sigma 0.4 of the range) histogrammed into the span 19 / MRD 9 prompts
    (761 x 380 x 159 x B uint16) and injected with the interpolated FP32
    estimate's scatter rates, device-resident; host clock around the
    synchronous calls (best of three)."""
    from opensss_amd import _native, castor
    lib = _native.load_library()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    lor = np.ascontiguousarray(w.lor_coordinates, dtype=np.int16)
    lut16 = np.ascontiguousarray(lut, dtype=np.int16)
    h = castor.listmode_header('bench_df.Cdf', n_events, duration=600.0, tof_range=float(w.tof_range))
    rng = np.random.default_rng(seed)
    words = castor.event_words(h)
    idw = castor.GenerateEventClass(h)[1][4]
    tofw = castor.GenerateEventClass(h)[1][3]
    ev = np.zeros((n_events, words), np.uint32)
    ev[:, 0] = rng.integers(0, 600000, n_events, dtype=np.uint32)
    ra = rng.integers(0, NR, n_events)
    rb = np.clip(ra + rng.integers(-9, 10, n_events), 0, NR - 1)
    da = rng.integers(0, N, n_events)
    db = (da + N // 2 + rng.integers(-(N // 4), N // 4 + 1, n_events)) % N
    ev[:, idw] = (ra * N + da).astype(np.uint32)
    ev[:, idw + 1] = (rb * N + db).astype(np.uint32)
    del ra, rb, da, db
    ev[:, tofw] = (rng.normal(0.0, 0.4, n_events) * w.tof_range).astype(np.float32).view(np.uint32)
    n_rad, n_ang, n_planes = N + 1, N // 2, int(lut16.max()) + 1
    factors = castor.injection_factors(scales, h['tof_range'], h['duration'])
    lay = castor._layout(h, scatter_word=castor.GenerateEventClass(dict(h, scatter_flag=True))[1][0])
    bufs = [_native.DeviceBuffer.from_array(ctx, a) for a in (ev, lor, lut16, factors)]
    prompts = _native.DeviceBuffer(ctx, 2 * n_rad * n_ang * n_planes * n_bins)
    out = _native.DeviceBuffer(ctx, 4 * (words + 1) * n_events)
    bufs += [prompts, out]
    del ev
    n_out = _native.ctypes.c_int64(0)
    try:
        t_h, t_i = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            _native.check(lib.sss_listmode_histogram(
                ctx.handle, bufs[0].ptr, n_events, _native.ctypes.byref(lay), N, 0, 1, N, bufs[1].ptr, NR,
                bufs[2].ptr, n_rad, n_ang, n_planes, n_bins, h['tof_range'], h['duration'], _native.SSS_LM_DEVICE,
                prompts.ptr, None, None), 'sss_listmode_histogram')
            t_h.append(time.perf_counter() - t0)
        for _ in range(3):
            t0 = time.perf_counter()
            _native.check(lib.sss_listmode_inject(
                ctx.handle, bufs[0].ptr, n_events, _native.ctypes.byref(lay), N, 0, 1, N, bufs[1].ptr, NR,
                bufs[2].ptr, n_rad, n_ang, n_planes, n_bins, h['tof_range'], bufs[3].ptr,
                _native.ctypes.c_void_p(scatter_ptr), _native.SSS_LM_DEVICE, out.ptr, _native.ctypes.byref(n_out)),
                'sss_listmode_inject')
            t_i.append(time.perf_counter() - t0)
    finally:
        for b in bufs:
            b.free()
    th, ti = min(t_h), min(t_i)
    return {
        'workload': f'{n_events} events ({4 * words} B), 760 x 80 scanner, span 19 MRD 9 prompts, {n_bins} TOF bins',
        'histogram_ms': th * 1e3, 'histogram_events_per_s': n_events / th,
        'inject_ms': ti * 1e3, 'inject_events_per_s': n_events / ti, 'injected_events': int(n_out.value),
        'host_clock': True,
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py runs on one GPU only: the scatter estimate of one instance is measured '
              'on one MI355X (the path would shard over z slices across GPUs; not run here)',
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.batch:
        batch_bench(args)
        return
    if args.serve > 0:
        serve_bench(args)
        return
    w = make_workload(args.workload)
    print(f'[bench] workload {w.name}: {len(w.scatter_points)} sp, grid {w.grid}, '
          f'{w.n_rings_used}x{w.n_det_used} sampled, {w.n_bins} bins', file=sys.stderr, flush=True)
    s = subject(w, 'fp32')
    s.set_activity(w.activity)
    s.run()
    s.synchronize()
    info = s.info
    for _ in range(args.warmup):
        s.run()
    s.synchronize()
    print('[bench] warmup done', file=sys.stderr, flush=True)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        s.run()
        s.synchronize()
        lat.append(time.perf_counter() - t0)
    elapsed = time.perf_counter() - t_start
    evals = info['evals']
    value = evals * args.steps / elapsed
    print(f'[bench] {args.steps} steps in {elapsed:.3f} s', file=sys.stderr, flush=True)

    # dominant kernel: the fused scatter sum (HIP events on the context stream)
    kms = []
    for _ in range(max(3, min(args.steps, 10))):
        s.scatter()
        kms.append(s.scatter_kernel_ms())
    k_ms = float(np.mean(kms))
    pairs, samples = s.work_stats()
    flops = 2.0 * w.n_bins * samples + 64.0 * pairs
    achieved = flops / (k_ms * 1e-3) / 1e12
    traffic, traffic_src = profile_traffic(w.name)
    # FP32 vector (VALU) bound, priced against MI355X's 157.3 TFLOP/s FP32 peak
    roofline = {'bound': 'valu', 'pipe': 'valu_fp32', 'achieved': achieved,
                'peak': FP32_VALU_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                'frac': achieved / FP32_VALU_PEAK_TFLOPS, 'traffic': traffic,
                'traffic_source': traffic_src, 'kernel': 'k_scatter', 'kernel_ms': k_ms,
                'flops_per_launch': flops, 'pairs': pairs, 'samples': samples}
    out32 = s.output_compact()
    s.close()

    err = None
    if not args.no_fp64_check:
        s64 = subject(w, 'fp64')
        s64.run(w.activity)
        out64 = s64.output_compact()
        s64.close()
        err = rel_err_stats(out32, out64)
        print(f"[bench] FP32 vs FP64: max rel err {err['max_rel_err_cells_above_1pct_peak']:.3e}",
              file=sys.stderr, flush=True)

    cpu = None
    err_oracle = None
    oracle_pair = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or cpu_threads()
        cpu, sample, want = cpu_baseline(w, threads)
        print(f"[bench] cpu baseline {cpu['value']:.3e} evals/s", file=sys.stderr, flush=True)
        # the FP32 GPU path on the same scatter points, checked against the oracle
        ss = subject(w, 'fp32', points=sample)
        ss.run(w.activity)
        got_sample = ss.output_compact()
        err_oracle = rel_err_stats(got_sample, want)
        oracle_pair = (got_sample, want)
        ss.close()
        err_oracle['sample'] = f'{len(sample)} scatter points of the cpu_baseline sample'
        print(f"[bench] FP32 vs oracle: max rel err {err_oracle['max_rel_err_cells_above_1pct_peak']:.3e}",
              file=sys.stderr, flush=True)

    # The legs below size their work by the free device memory: the device
    # blocks that the config-2 subjects above left in the default context's
    # cache go back first, so each leg starts as it would in a process of its own.
    from opensss_amd import _native
    batched = served = None
    if args.workload == 'config2' and args.default_batch > 0:
        _native.default_context().release_cache()
        batched = batch_reading('random', args.default_batch, args.batch_streams)
        batched['config1_subjects'] = config1_graph_reading()
    if args.workload == 'config2' and args.default_serve > 0:
        _native.default_context().release_cache()
        served = serve_reading('random', args.subjects, args.rate, args.default_serve, args.streams)

    stages = None
    if w.name.startswith('tofsss-760x80') and not args.no_stages and not args.no_fp64_check:
        stages = sinogram_stages(w, out32, out64, oracle_pair)
        print(f"[bench] sinogram stages: interpolate {stages['interpolate_ms']:.2f} ms, mask "
              f"{stages['tail_mask_ms']:.1f} ms, fit {stages['tail_fit_ms_per_bin']:.1f} ms/bin "
              f"({stages['tail_fit_all_bins_ms']:.1f} ms for {len(stages['fitted_bins'])} bins in one call, "
              f"identical {stages['tail_fit_all_bins_identical']}), list mode "
              f"{stages['listmode']['histogram_events_per_s'] / 1e9:.2f} G events/s histogram, "
              f"{stages['listmode']['inject_events_per_s'] / 1e9:.2f} G events/s inject, image preparation "
              f"{stages['image_preparation']['ms']:.1f} ms, tail scale "
              f"FP32 vs FP64 {stages['tail_scale_max_rel_err_fp32_vs_fp64']:.2e}", file=sys.stderr, flush=True)

    lat_ms = np.array(lat) * 1e3
    line = {
        'metric': METRIC, 'value': value, 'unit': 'evals/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': elapsed / args.steps * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': w.name, 'n_scatter_points': int(len(w.scatter_points)),
                   'lors_per_scatter_point': int(info['n_pairs_enumerated'] // max(1, len(w.scatter_points))),
                   'tof_bins': w.n_bins, 'evals_per_step': int(evals),
                   'precision': 'fp32 fold and sums; fp64 crossings, distances, energy index; '
                                '32.32 fixed-point TOF offsets',
                   'parallelism': 'single instance, 1 GPU'},
        'latency_ms': {'p50': float(np.percentile(lat_ms, 50)), 'p99': float(np.percentile(lat_ms, 99))},
        'error_vs_fp64': err,
        'error_vs_oracle': err_oracle,
        'roofline': roofline,
        'cpu_baseline': cpu,
        'sinogram_stages': stages,
        'batched': batched,
        'served': served,
    }
    print(json.dumps(line))


if __name__ == '__main__':
    main()
