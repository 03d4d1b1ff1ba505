"""Autotuner for the launch shapes and frontier chunk sizes of the persistent label kernels.

The flood (`prepare_labels`) and trace (`trace_labels`) kernels take six compile-time shapes:
threads per label workgroup for each, how many of a voxel's 26 neighbours one thread keeps in
flight together (`for_parts` chunks: 9, 13 or 26; more in flight hides more latency and costs
registers), and how many label workgroups are resident at once (fewer: less contention for the
labels on the critical path, later starts for the rest). They set speed only: every value gives bit-identical skeletons. This tuner searches
them by coordinate descent from the current values in `csrc/tuned.hpp`: for each parameter in
turn it builds the library with each candidate value (the others at their best so far), times it
in a child process on the workloads below, rejects a candidate whose results differ from the
current build's in any byte, keeps the fastest, and finally rewrites `csrc/tuned.hpp`.

Workloads (score = sum of median call times): the 512^3 / 200-label tree volume as one device-
resident `skeletonize` call, and a `skeletonize_batch` call over 16 tube volumes of 128^3.

    python -m brook_amd.tune            # tune on cuda:0, rewrite csrc/tuned.hpp (needs hipcc)
    python -m brook_amd.tune --dry-run  # measure and report only

After it rewrites the header, rebuild (`__graft_entry__.build()`).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

_HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(_HERE, "csrc")
HEADER = os.path.join(SRC, "tuned.hpp")
ROOT = os.path.dirname(_HERE)

SPACE = {  # parameter -> candidate values, in the order coordinate descent visits them
    "BROOK_TRACE_BLOCK": [512, 1024],
    "BROOK_NB_BATCH": [9, 13, 26],
    "BROOK_FLOOD_BLOCK": [512, 1024],
    "BROOK_FLOOD_NB_BATCH": [9, 13, 26],
    "BROOK_TRACE_GRID": [0, 64, 128, 256],
    "BROOK_FLOOD_GRID": [0, 64, 128, 256],
}


def current():
    """The values `csrc/tuned.hpp` holds now."""
    text = open(HEADER).read()
    return {k: int(re.search(rf"#define {k} (\d+)", text).group(1)) for k in SPACE}


def _flags():
    from brook_amd._flags import HIP_FLAGS
    return list(HIP_FLAGS)


def build(values, out_dir, common_dir, jobs=16):
    """Compile the sources with `values` as -D macros into out_dir/libbrook_amd.so. Sources that do
    not include tuned.hpp are compiled once into common_dir and shared by every variant."""
    flags = _flags() + [f"-D{k}={v}" for k, v in values.items()]
    sources = sorted(os.path.join(SRC, f) for f in os.listdir(SRC)
                     if f.endswith((".hip", ".cpp")) and f != "synth.cpp")
    shaped = {src for src in sources if '#include "tuned.hpp"' in open(src).read()}

    def compile_one(src):
        obj = os.path.join(out_dir if src in shaped else common_dir, os.path.basename(src) + ".o")
        if src in shaped or not os.path.exists(obj):
            r = subprocess.run(["hipcc", *flags, "-c", src, "-o", obj + ".tmp"], capture_output=True, text=True)
            if r.returncode:
                raise RuntimeError(f"compile failed for {values}: {src}\n{r.stderr[-2000:]}")
            os.replace(obj + ".tmp", obj)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(compile_one, sources))
    lib = os.path.join(out_dir, "libbrook_amd.so")
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib, *objects], check=True)
    for o in objects:
        if os.path.dirname(o) == out_dir:
            os.remove(o)
    return lib


def measure(repeats=3):
    """Child side: time the workloads with the library BROOK_AMD_LIBRARY names; print one JSON line
    with the median times (ms) and a digest of every result byte."""
    import numpy as np
    import torch

    import brook_amd
    from brook_amd import synth
    digest = hashlib.sha256()

    def absorb(packed):
        d = packed.as_dict()
        for name in ("labels", "v_off", "e_off", "vertices", "edges", "radii"):
            digest.update(np.ascontiguousarray(np.asarray(d[name])).tobytes())

    trees = synth.trees_512()
    flat = torch.from_numpy(np.ravel(trees, order="F")).to("cuda:0")
    device = flat.view(trees.shape[2], trees.shape[1], trees.shape[0]).permute(2, 1, 0)
    params = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)
    plain = lambda: brook_amd.skeletonize(device, params, (8, 8, 40), dust_threshold=1000, output="packed",  # noqa: E731
                                          progress=False)
    tubes = [synth.tubes_128(1001 + i) for i in range(16)]
    batch = lambda: brook_amd.skeletonize_batch(tubes, teasar_params=params, anisotropy=(16, 16, 40),  # noqa: E731
                                                dust_threshold=1000, output="packed")
    out = {}
    for name, run in (("plain_ms", plain), ("batch_ms", batch)):
        result = run()
        times = []
        for _ in range(repeats):
            torch.cuda.synchronize()
            t = time.perf_counter()
            run()
            torch.cuda.synchronize()
            times.append(1e3 * (time.perf_counter() - t))
        out[name] = float(np.median(times))
        for part in (result if name == "batch_ms" else [result]):
            absorb(part)
    out["digest"] = digest.hexdigest()
    out["device"] = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    print(json.dumps(out), flush=True)


def run_child(lib, timeout=300):
    env = dict(os.environ, BROOK_AMD_LIBRARY=lib)
    r = subprocess.run([sys.executable, "-u", "-m", "brook_amd.tune", "--measure"], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=timeout)
    if r.returncode:
        raise RuntimeError(f"measurement failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def write_header(values, scores):
    text = open(HEADER).read()
    for k, v in values.items():
        text = re.sub(rf"#define {k} \d+", f"#define {k} {v}", text)
    note = "// Last tuned: " + json.dumps(scores, sort_keys=True) + "\n"
    text = re.sub(r"// Last tuned: .*\n", "", text)
    text = text.replace("#pragma once\n", "#pragma once\n" + note, 1)
    with open(HEADER, "w") as f:
        f.write(text)


def tune(dry_run=False, log=print):
    best = current()
    work = tempfile.mkdtemp(prefix="brook_tune_")
    common = os.path.join(work, "common")
    os.makedirs(common)
    seen = {}

    def score(values):
        key = tuple(sorted(values.items()))
        if key not in seen:
            out_dir = os.path.join(work, "v%d" % len(seen))
            os.makedirs(out_dir)
            t = time.time()
            lib = build(values, out_dir, common)
            log(f"built {values} in {time.time() - t:.0f} s", flush=True)
            seen[key] = run_child(lib)
            log(f"  {seen[key]}", flush=True)
        return seen[key]

    base = score(best)
    for name, candidates in SPACE.items():
        for value in candidates:
            if value == best[name]:
                continue
            trial = dict(best, **{name: value})
            got = score(trial)
            if got["digest"] != base["digest"]:
                raise RuntimeError(f"{trial} changed the results: the kernels are not shape-independent")
            total = lambda s: s["plain_ms"] + s["batch_ms"]  # noqa: E731
            if total(got) < 0.98 * total(score(best)):  # a 2 % margin over run-to-run noise
                best = trial
        log(f"{name} -> {best[name]}", flush=True)
    final = score(best)
    summary = {"values": best, "plain_ms": round(final["plain_ms"], 2), "batch_ms": round(final["batch_ms"], 2),
               "start_plain_ms": round(base["plain_ms"], 2), "start_batch_ms": round(base["batch_ms"], 2)}
    log(json.dumps(summary), flush=True)
    if not dry_run:
        write_header(best, {"device": final.get("device"), "plain_ms": summary["plain_ms"],
                            "batch_ms": summary["batch_ms"]})
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--measure", action="store_true", help="child mode: time the library BROOK_AMD_LIBRARY names")
    ap.add_argument("--dry-run", action="store_true", help="report the best values without rewriting the header")
    args = ap.parse_args()
    if args.measure:
        measure()
    else:
        tune(dry_run=args.dry_run)


if __name__ == "__main__":
    main()
