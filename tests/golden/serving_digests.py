"""Writes serving_digests.json: SHA-256 digests of the oracle's packed result for the first
requests of the serving stream (synth.serving_request(i): uint32 tube volumes with every edge in
[128, 512] and 5-50 labels; teasar scale 1.5, const 300, pdrf_scale 1e5, pdrf_exponent 4;
anisotropy (16, 16, 40); dust 1000; fix_borders). The GPU serving test checks the server's results
against these digests, so it needs no oracle run on the GPU host. About a minute on 8 threads.

    python tests/golden/serving_digests.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.golden.workload_digests import PARAMS, digest  # noqa: E402

ANISOTROPY = (16, 16, 40)
REQUESTS = 12


def main():
    import oracle
    from brook_amd import synth
    out = {}
    for i in range(REQUESTS):
        vol = synth.serving_request(i)
        r = oracle.skeletonize(vol, PARAMS, ANISOTROPY, 1000, threads=8)
        out[str(i)] = dict(digest(r), shape=list(vol.shape), skeletons=len(r["labels"]))
        print(i, vol.shape, len(r["labels"]), flush=True)
    with open(os.path.join(HERE, "serving_digests.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
