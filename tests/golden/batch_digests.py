"""Writes batch_digests.json: for each of the 64 volumes of the batch workload (synth.batch_64:
log-uniform edges in [64, 256], 1-50 tube labels; teasar scale 1.5, const 300, pdrf_scale 1e5,
pdrf_exponent 4; anisotropy (16, 16, 40); dust 1000; fix_borders), one SHA-256 over the oracle's
packed result arrays (workload_digests.digest's normalisation, concatenated in its order) and the
skeleton count. The GPU test of the batch compares each sample of one skeletonize_batch call with
these digests. About a minute on 8 threads.

    python tests/golden/batch_digests.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.golden.workload_digests import PARAMS, digest  # noqa: E402

ANISOTROPY = (16, 16, 40)


def sample_digest(packed):
    """One SHA-256 over the per-array digests (names in workload_digests.NAMES order)."""
    d = digest(packed)
    return hashlib.sha256(json.dumps(d, sort_keys=True).encode()).hexdigest()


def main():
    import oracle
    from brook_amd import synth
    out = []
    for i, vol in enumerate(synth.batch_64()):
        r = oracle.skeletonize(vol, PARAMS, ANISOTROPY, 1000, threads=8)
        out.append(dict(shape=list(vol.shape), skeletons=len(r["labels"]), sha256=sample_digest(r)))
        print(i, vol.shape, len(r["labels"]), flush=True)
    with open(os.path.join(HERE, "batch_digests.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
