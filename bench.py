"""Benchmark: Mvoxels/s skeletonized on one MI355X, device-resident input.

A step is one `brook_amd.skeletonize` call (CCL, EDT, border targets, floods, PDRF, target order,
trace, assembly; packed host result) on the 512^3 uint32 volume of 200 branching-tree labels
(~40% occupancy, anisotropy (8,8,40), scale 1.5, const 300, pdrf_scale 1e5, pdrf_exponent 4,
dust 1000, fix_borders) already resident in HBM as a Fortran-ordered tensor.

Prints one JSON line (rank 0): throughput, p50/p99 step latency, the roofline of the dominant
kernel (algorithmic bytes / event-timed kernel duration on the engine's own stream) and the
CPU oracle timed on the same volume on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mvoxels/s skeletonized (512^3, 200 labels); batched/served vs plain; p99 ms"
PARAMS = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)
ANISO = (8, 8, 40)
HBM_PEAK_GBS = 8000.0


def _device_volume(vol):
    import torch
    flat = torch.from_numpy(np.ravel(vol, order="F")).cuda()
    sx, sy, sz = vol.shape
    return flat.view(sz, sy, sx).permute(2, 1, 0)  # shape (sx, sy, sz), Fortran strides


def _kernel_table(stats, n, b_in, b_cc, F):
    """Per-kernel event time and algorithmic bytes per launch (per-voxel figures of the path's
    stage table; trace and floods per reached/claimed voxel)."""
    reached, claimed, flood = stats.get("railroad_reached_voxels", 0), stats.get("invalidated_voxels", 0), stats.get("flood_voxels", 0)
    algo = {
        "ccl": (["ccl_init", "ccl_union", "ccl_count", "ccl_mark", "ccl_relabel"], (b_in + b_cc) * n),
        "edt": (["edt_x", "edt_y", "edt_z"], (3 * b_cc + 4 + 8 + 8) * n),
        "prepare_labels": (["prepare_labels"], 12 * flood),
        "init_records": (["init_records"], (b_cc + 16) * n),
        "pdrf_alive": (["pdrf_alive"], b_cc * n + (4 + 4 + 4 + 8) * F),
        "trace_labels": (["trace_labels"], 12 * reached + 13 * claimed),
    }
    rows = {}
    for name, (kernels, nbytes) in algo.items():
        ms = sum(stats.get(f"k_{k}_ms", 0.0) for k in kernels)
        launches = max(1.0, stats.get(f"k_{kernels[-1]}_n", 1.0))
        if ms <= 0:
            continue
        achieved = nbytes / (ms / 1e3) / 1e9
        rows[name] = dict(ms=ms, launches=launches, bytes=nbytes, achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS)
    return rows


def _traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def edt_error(dev):
    """The GPU EDT (DBF) of the workload against the digest of the oracle's float32 DBF; the
    oracle's relative error against the float64 EDT (computed on the CPU when the digest was
    written, tests/golden/workload_digests.py) is then the GPU's."""
    import hashlib

    from brook_amd import device
    want = json.load(open(os.path.join(ROOT, "tests", "golden", "workload_digests.json")))["edt"]["trees_512"]
    dbf = device.edt(dev, ANISO)
    got = hashlib.sha256(dbf.permute(2, 1, 0).contiguous().cpu().numpy().tobytes()).hexdigest()
    return dict(equals_oracle_dbf=got == want["sha256"], max_rel_err_f64=want["max_rel_err_f64"],
                mean_rel_err_f64=want["mean_rel_err_f64"], foreground_voxels=want["foreground_voxels"])


def host_io(vol, calls=3):
    """The same workload with host input (Fortran numpy array) and host output: includes the
    host-to-device copy of the labels."""
    import brook_amd
    times = []
    for _ in range(calls):
        t0 = time.perf_counter()
        brook_amd.skeletonize(vol, PARAMS, ANISO, dust_threshold=1000, output="packed", progress=False)
        times.append(time.perf_counter() - t0)
    return dict(value=round(vol.size / float(np.median(times)) / 1e6, 3), unit="Mvoxels/s", calls=calls,
                p50_ms=round(float(np.median(times)) * 1e3, 3))


def parent_paths(dev, n, calls=3):
    """The same workload with fix_branching=False (one vertex-weighted field and parent walks per
    label instead of a railroad per path): Mvoxels/s over `calls` timed calls after a warm one."""
    import torch

    import brook_amd
    from brook_amd import _native as native
    run = lambda: brook_amd.skeletonize(dev, PARAMS, ANISO, dust_threshold=1000, output="packed", progress=False,
                                        fix_branching=False)
    result = run()
    torch.cuda.synchronize()
    ctx = native.context()
    t0 = time.perf_counter()
    for _ in range(calls):
        result = run()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / calls
    return dict(value=round(n / wall / 1e6, 3), unit="Mvoxels/s", calls=calls, ms_per_call=round(wall * 1e3, 3),
                skeletons=len(result), vertices=int(result.vertices.shape[0]),
                stages_ms={k: round(ctx.stat(k), 3) for k in ("prep_floods_ms", "prep_pdrf_order_ms",
                                                              "trace_ms")})


def merge(result, fragments=8, calls=5):
    """Packed merge of the workload's skeletons split into `fragments` fragments (each holds every
    vertex of every skeleton and every `fragments`-th edge: the overlap of chunked results),
    device-resident parts: input vertices merged per second over `calls` timed calls."""
    import torch

    from brook_amd import PackedSkeletons, packed
    h = result.to_host()
    parts = []
    for f in range(fragments):
        pieces = [h.edges[h.e_off[k]:h.e_off[k + 1]][f::fragments] for k in range(len(h))]
        e_off = np.concatenate([[0], np.cumsum([len(p) for p in pieces])]).astype(np.int64)
        parts.append(PackedSkeletons(h.labels, h.v_off, e_off, h.vertices, np.concatenate(pieces), h.radii,
                                     h.anisotropy).to_device("cuda"))
    origins = [(0.0, 0.0, 0.0)] * fragments
    merged = packed.merge_fragments(parts, origins)
    whole = packed.merge_fragments([h], [(0.0, 0.0, 0.0)])
    same = all(np.array_equal(getattr(merged, k), getattr(whole, k)) for k in ("vertices", "edges", "radii", "v_off"))
    times = []
    for _ in range(calls):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        packed.merge_fragments(parts, origins)
        times.append(time.perf_counter() - t0)
    nv_in = fragments * int(h.vertices.shape[0])
    wall = float(np.median(times))
    return dict(value=round(nv_in / wall / 1e6, 3), unit="M input vertices/s", fragments=fragments, calls=calls,
                p50_ms=round(wall * 1e3, 3), input_vertices=nv_in, input_edges=int(h.edges.shape[0]),
                output_vertices=int(merged.vertices.shape[0]), split_equals_whole=bool(same))


def postprocess(result, dust=1500.0, ticks=3500.0, calls=5):
    """The canonical GPU postprocess (dust, loops, joins, ticks; Kimimaro's default thresholds in
    physical units) of the workload's skeletons, device-resident input: input vertices per second
    over `calls` timed calls, with the result's size and whether repeated calls agree."""
    import torch

    from brook_amd import device_post
    part = result.to_host().to_device("cuda")
    first = device_post.postprocess(part, dust, ticks).to_host()
    times = []
    for _ in range(calls):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        device_post.postprocess(part, dust, ticks)
        times.append(time.perf_counter() - t0)
    again = device_post.postprocess(part, dust, ticks).to_host()
    same = all(np.array_equal(getattr(first, k), getattr(again, k)) for k in ("vertices", "edges", "radii", "v_off", "e_off"))
    nv = int(part.vertices.shape[0])
    wall = float(np.median(times))
    return dict(value=round(nv / wall / 1e6, 3), unit="M input vertices/s", calls=calls, p50_ms=round(wall * 1e3, 3),
                dust_threshold=dust, tick_threshold=ticks, input_vertices=nv, input_edges=int(part.edges.shape[0]),
                output_vertices=int(first.vertices.shape[0]), output_edges=int(first.edges.shape[0]),
                repeat_equal=bool(same))


def batched():
    """64 mixed-shape volumes (64..256 per axis, 1-50 tube labels each, ~181 M voxels in all),
    device-resident: one skeletonize_batch call against 64 plain calls; every sample's batch result
    must equal its plain call byte for byte."""
    import torch

    import brook_amd
    from brook_amd import synth
    vols = synth.batch_64()
    dev = [_device_volume(v) for v in vols]
    total = sum(v.size for v in vols)
    params, aniso = PARAMS, (16, 16, 40)
    brook_amd.skeletonize_batch(dev, params, aniso, dust_threshold=1000, output="packed")  # pools warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch = brook_amd.skeletonize_batch(dev, params, aniso, dust_threshold=1000, output="packed")
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    plain = []
    t0 = time.perf_counter()
    for d in dev:
        plain.append(brook_amd.skeletonize(d, params, aniso, dust_threshold=1000, output="packed", progress=False))
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    names = ("labels", "v_off", "e_off", "vertices", "edges", "radii")
    differ = [i for i, p in enumerate(plain)
              if not all(np.array_equal(np.asarray(batch[i].as_dict()[k]), np.asarray(p.as_dict()[k])) for k in names)]
    if differ:
        raise RuntimeError(f"batch results differ from plain calls for samples {differ}")
    return dict(voxels=int(total), samples=len(vols), batch_mvox_s=round(total / t_batch / 1e6, 3),
                plain_mvox_s=round(total / t_plain / 1e6, 3), batch_ms=round(t_batch * 1e3, 3),
                plain_ms=round(t_plain * 1e3, 3), speedup=round(t_plain / t_batch, 3),
                skeletons=int(sum(len(p) for p in plain)), batch_equals_plain=True)


def stress(calls=2):
    """The stress workload: 1024x1024x512 uint64 Voronoi cells (Poisson(5000) seeds under anisotropy
    (4,4,40), eroded by one voxel) holding 1-4 tubes each, device-resident; Mvoxels/s and
    skeletons (objects) per second over `calls` timed calls after one warm call."""
    import torch

    import brook_amd
    from brook_amd import _native as native
    from brook_amd import synth
    t0 = time.perf_counter()
    vol, cells, painted = synth.cells_1024()
    gen_s = time.perf_counter() - t0
    dev = _device_volume(vol.view(np.int64))  # ids < 2^63: the same values as signed integers
    del vol
    aniso = (4, 4, 40)
    run = lambda: brook_amd.skeletonize(dev, PARAMS, aniso, dust_threshold=1000, output="packed", progress=False)
    result = run()
    torch.cuda.synchronize()
    ctx = native.context()
    t0 = time.perf_counter()
    for _ in range(calls):
        result = run()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / calls
    n = dev.numel()
    out = dict(workload="cells_1024: 1024x1024x512 uint64, Poisson(5000) Voronoi cells under (4,4,40) eroded by one "
                        "voxel, 1-4 tubes of radius U[2,6] per cell, ids cell<<32|tube; dust 1000, fix_borders",
               voxels=int(n), cells=cells, painted_voxels=painted, components=int(ctx.stat("components")),
               skeletons=len(result), calls=calls, ms_per_call=round(wall * 1e3, 3),
               mvox_s=round(n / wall / 1e6, 3), objects_s=round(len(result) / wall, 1),
               stages_ms={k: round(ctx.stat(k), 3) for k in ("prep_ccl_ms", "prep_edt_ms", "prep_borders_ms",
                                                              "prep_floods_ms", "prep_pdrf_order_ms", "trace_ms",
                                                              "assembly_ms")},
               generator_s=round(gen_s, 2))
    del dev
    torch.cuda.empty_cache()
    return out


def served(rates=(20.0, 50.0), count=40, seed=5005):
    """Serving form: Poisson request streams; request i is synth.serving_request(i) (a distinct
    host uint32 tube volume, every edge in [128, 512], 5-50 labels). The same arrival sequence is
    served by the continuous-batching server (two slots) and by a serial server (one slot, plain
    calls). Latency = result ready - arrival."""
    from concurrent.futures import ThreadPoolExecutor

    import brook_amd
    from brook_amd import synth
    with ThreadPoolExecutor(16) as pool:
        vols = list(pool.map(synth.serving_request, range(count)))
    params, aniso = PARAMS, (16, 16, 40)
    brook_amd.skeletonize(vols[0], params, aniso, dust_threshold=1000, output="packed", progress=False)
    rng = np.random.default_rng(seed)
    voxels = sum(v.size for v in vols)
    out = {}
    for lam in rates:
        gaps = rng.exponential(1.0 / lam, count)
        row = {}
        for mode, slots, max_batch in (("batched", 2, 64), ("serial", 1, 1)):
            with brook_amd.Server(params, aniso, 1000, max_batch=max_batch, slots=slots) as srv:
                srv.submit(vols[0]).result()  # each slot's context warm
                if slots > 1:  # the other slots' contexts warm
                    warm = [srv.submit(vols[1 + i]) for i in range(slots)]
                    for f in warm:
                        f.result()
                srv.batches.clear()
                t0 = time.perf_counter()
                at = t0
                futures = []
                for i, g in enumerate(gaps):
                    at += g
                    wait = at - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                    futures.append(srv.submit(vols[i]))
                    if (i + 1) % 50 == 0:
                        _progress(f"served {lam:g} req/s, {mode}: {i + 1} of {count} submitted")
                for f in futures:
                    f.result()
                _progress(f"served {lam:g} req/s, {mode} done")
                lat = [(f.request.done - f.request.arrival) * 1e3 for f in futures]
                span = max(f.request.done for f in futures) - t0
                row[mode] = dict(p50_ms=round(float(np.percentile(lat, 50)), 2),
                                 p99_ms=round(float(np.percentile(lat, 99)), 2),
                                 req_per_s=round(count / span, 3), mvox_s=round(voxels / span / 1e6, 3),
                                 mean_batch=round(float(np.mean(srv.batches)), 2),
                                 overlapped=round(float(np.mean([f.request.overlapped for f in futures])), 3))
        out[f"{lam:g}"] = dict(requests=count, **row)
    return out


def _host_cpu():
    """nproc and the CPU model name of this host."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return len(os.sched_getaffinity(0)), model


def cpu_baseline(vol, threads):
    """The oracle (a C++ port of the reference's algorithm, threaded over labels) on the same
    512^3 volume as the GPU steps, and on the 128^3 / 20-label tube volume (anisotropy (16,16,40))."""
    import oracle
    from brook_amd import synth
    t0 = time.perf_counter()
    oracle.skeletonize(vol, PARAMS, ANISO, 1000, threads=threads)
    dt = time.perf_counter() - t0
    small = synth.tubes_128()
    t1 = time.perf_counter()
    oracle.skeletonize(small, PARAMS, (16, 16, 40), 1000, threads=threads)
    dt1 = time.perf_counter() - t1
    nproc, model = _host_cpu()
    return dict(value=round(vol.size / dt / 1e6, 3), unit="Mvoxels/s", cores=threads, kind="port",
                sample=f"oracle on the whole 512^3 / 200-label volume ({vol.size} voxels), {threads} threads, "
                       f"{dt:.2f} s",
                nproc=nproc, cpu_model=model,
                tubes_128=dict(value=round(small.size / dt1 / 1e6, 3), unit="Mvoxels/s", seconds=round(dt1, 3),
                               sample="128^3 uint32, 20 tube labels, anisotropy (16,16,40)"))


def _progress(what):
    """One line on stderr per finished part (the JSON line stays the only stdout output)."""
    print(f"bench: {what}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the host-input, fix_branching=False, merge, batched, stress and served measurements")
    ap.add_argument("--served-sweep", action="store_true",
                    help="serve at 1, 2, 5, 10, 20 and 50 requests/s (default: 20 and 50)")
    ap.add_argument("--served-requests", type=int, default=40, help="requests per arrival rate")
    ap.add_argument("--served-rates", default="1,2,5,10,20,50",
                    help="arrival rates (req/s) of --served-sweep, comma-separated")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the known-size reference "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by volume: batch samples and serving requests per GPU)", file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4,
                          "launches": 3}))
        return 0

    import brook_amd
    from brook_amd import _native as native
    from brook_amd import synth

    t0 = time.perf_counter()
    vol = synth.trees_512()
    gen_s = time.perf_counter() - t0
    n = vol.size
    dev = _device_volume(vol)
    torch.cuda.synchronize()
    ctx = native.context()
    run = lambda: brook_amd.skeletonize(dev, PARAMS, ANISO, dust_threshold=1000, output="packed", progress=False)
    for _ in range(args.warmup):
        run()
    native.check(native.lib().brook_context_set_kernel_timing(ctx.handle, 1))
    names = ["ccl_init", "ccl_union", "ccl_count", "ccl_mark", "ccl_relabel", "edt_x", "edt_y", "edt_z",
             "init_records", "prepare_labels", "pdrf_alive", "trace_labels"]
    stat_names = ["railroad_reached_voxels", "invalidated_voxels", "flood_voxels", "foreground_voxels", "cc_bytes",
                  "paths", "components", "prep_ccl_ms", "prep_analyze_ms", "prep_edt_ms", "prep_borders_ms",
                  "prep_floods_ms", "prep_pdrf_order_ms", "trace_ms", "assembly_ms"]
    acc = {}
    times = []
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        result = run()
        times.append((time.perf_counter() - s0) * 1e3)
        for k in names:
            for suffix in ("_ms", "_n"):
                acc[f"k_{k}{suffix}"] = acc.get(f"k_{k}{suffix}", 0.0) + ctx.stat(f"k_{k}{suffix}")
        for k in stat_names:
            acc[k] = acc.get(k, 0.0) + ctx.stat(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    native.check(native.lib().brook_context_set_kernel_timing(ctx.handle, 0))
    K = args.steps
    per = {k: v / K for k, v in acc.items()}
    b_cc = int(per.get("cc_bytes", 2))
    kernels = _kernel_table(per, n, 4, b_cc, per.get("foreground_voxels", 0))
    dominant = max(kernels, key=lambda k: kernels[k]["ms"]) if kernels else None
    roof = None
    if dominant:
        row = kernels[dominant]
        roof = dict(bound="hbm", kernel=dominant, achieved=round(row["achieved_gbs"], 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(row["frac"], 5), traffic=_traffic(dominant),
                    algorithmic_bytes_per_launch=int(row["bytes"]), ms_per_launch=round(row["ms"], 3))
    out = dict(
        metric=METRIC,
        value=round(n * K / elapsed / 1e6, 3),
        unit="Mvoxels/s",
        n_gpus=1,
        steps=K,
        warmup=args.warmup,
        ms_per_step=round(elapsed / K * 1e3, 3),
        p50_ms=round(float(np.percentile(times, 50)), 3),
        p99_ms=round(float(np.percentile(times, 99)), 3),
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f32+int32",
        data="synthetic",
        config=dict(workload="trees_512: 512^3 uint32 volume, 200 branching-tree labels, ~40% occupancy, "
                             "anisotropy (8,8,40), scale 1.5, const 300, pdrf_scale 1e5, pdrf_exponent 4, "
                             "dust 1000, fix_borders; device-resident input, packed host output",
                    voxels=int(n), skeletons=len(result), vertices=int(result.vertices.shape[0]),
                    paths=int(per.get("paths", 0)), components=int(per.get("components", 0)),
                    parallelism="1 GPU"),
        stages_ms={k: round(per.get(k, 0.0), 3) for k in stat_names if k.endswith("_ms")},
        kernels={k: dict(ms=round(v["ms"], 3), achieved_gbs=round(v["achieved_gbs"], 2), frac=round(v["frac"], 5))
                 for k, v in kernels.items()},
        roofline=roof,
        generator_s=round(gen_s, 2),
    )
    _progress(f"plain steps done ({out['value']} Mvoxels/s)")
    out["edt"] = edt_error(dev)
    if not args.no_extras:
        out["host_io"] = host_io(vol)
        _progress("host_io done")
        out["fix_branching_false"] = parent_paths(dev, n)
        _progress("fix_branching_false done")
        out["merge"] = merge(result)
        _progress("merge done")
        out["postprocess"] = postprocess(result)
        _progress("postprocess done")
        out["batched"] = batched()
        _progress("batched done")
        out["stress"] = stress()
        _progress("stress done")
        rates = tuple(float(r) for r in args.served_rates.split(","))
        out["served"] = served(rates if args.served_sweep else (20.0, 50.0),
                               count=args.served_requests)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(vol, args.cpu_threads)
        _progress("cpu_baseline done")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
