"""Writes workload_digests.json: SHA-256 digests of the oracle's packed result on the 512^3 /
200-label tree workload (synth.trees_512, seed 2002; teasar scale 1.5, const 300, pdrf_scale 1e5,
pdrf_exponent 4; anisotropy (8, 8, 40); dust 1000; fix_borders), with fix_branching on and off;
and, under "edt", for that volume and the 128^3 / 20-label tube volume (anisotropy (16, 16, 40)),
the digest of the oracle's float32 EDT of the component ids (the DBF) with its maximum and mean
relative error against the float64 EDT over the foreground. The GPU tests compare brook_amd's
results with these digests, so the checks need no oracle run on the GPU host. Takes about three
minutes on 8 threads (`edt` alone: refresh only the EDT entries).

    python tests/golden/workload_digests.py [edt]
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


def edt_stats(volume, anisotropy):
    """The oracle's float32 DBF (EDT of the component ids) digest and its relative error against
    the float64 EDT over the foreground voxels."""
    import oracle
    cc, _, _ = oracle.ccl(volume)
    f32 = oracle.edt(cc, anisotropy)
    f64 = oracle.edt_f64(cc, anisotropy)
    fg = f64 > 0
    rel = np.abs(f32[fg].astype(np.float64) - f64[fg]) / f64[fg]
    return {"sha256": hashlib.sha256(np.asfortranarray(f32).tobytes(order="F")).hexdigest(),
            "max_rel_err_f64": float(rel.max()), "mean_rel_err_f64": float(rel.mean()),
            "foreground_voxels": int(fg.sum()), "anisotropy": list(anisotropy)}


def main():
    import oracle
    from brook_amd import synth
    path = os.path.join(HERE, "workload_digests.json")
    volume = synth.trees_512()
    edt = {"trees_512": edt_stats(volume, (8, 8, 40)), "tubes_128": edt_stats(synth.tubes_128(), (16, 16, 40))}
    print(edt, flush=True)
    if sys.argv[1:] == ["edt"]:
        result = json.load(open(path))
        result["edt"] = edt
        with open(path, "w") as f:
            json.dump(result, f, indent=1)
        return
    result = {"edt": edt}
    for fix_branching in (True, False):
        r = oracle.skeletonize(volume, PARAMS, (8, 8, 40), 1000, fix_branching=fix_branching, threads=8)
        result["fix_branching" if fix_branching else "parent_walks"] = dict(digest(r), paths=r["stats"]["paths"])
        print(fix_branching, r["stats"], flush=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


if __name__ == "__main__":
    main()
