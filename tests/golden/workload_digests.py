"""Writes workload_digests.json: SHA-256 digests of the oracle's packed result on the 512^3 /
200-label tree workload (synth.trees_512, seed 2002; teasar scale 1.5, const 300, pdrf_scale 1e5,
pdrf_exponent 4; anisotropy (8, 8, 40); dust 1000; fix_borders), with fix_branching on and off.
The GPU test of the full workload compares brook_amd's result with these digests, so the check
needs no oracle run on the GPU host. Takes about three minutes on 8 threads.

    python tests/golden/workload_digests.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

NAMES = ("labels", "v_off", "e_off", "vertices", "edges", "radii")
PARAMS = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)


def digest(packed):
    """Per array: dtype-normalised bytes (labels/offsets int64, vertices/radii float32, edges
    uint32, C order) -> sha256, plus the array shapes."""
    kinds = dict(labels=np.int64, v_off=np.int64, e_off=np.int64, vertices=np.float32, edges=np.uint32,
                 radii=np.float32)
    out = {}
    for name in NAMES:
        a = np.ascontiguousarray(np.asarray(packed[name]), dtype=kinds[name])
        out[name] = {"shape": list(a.shape), "sha256": hashlib.sha256(a.tobytes()).hexdigest()}
    return out


def main():
    import oracle
    from brook_amd import synth
    volume = synth.trees_512()
    result = {}
    for fix_branching in (True, False):
        r = oracle.skeletonize(volume, PARAMS, (8, 8, 40), 1000, fix_branching=fix_branching, threads=8)
        result["fix_branching" if fix_branching else "parent_walks"] = dict(digest(r), paths=r["stats"]["paths"])
        print(fix_branching, r["stats"], flush=True)
    with open(os.path.join(HERE, "workload_digests.json"), "w") as f:
        json.dump(result, f, indent=1)


if __name__ == "__main__":
    main()
