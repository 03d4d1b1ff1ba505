"""Serving-form parity on the GPU: requests of the serving stream (uint32 tube volumes with every
edge in [128, 512] and 5-50 labels; synth.serving_request) served by brook_amd.Server with two
slots, most of them arriving while another batch is on the GPU, each result checked against the
oracle's digests (tests/golden/serving_digests.py wrote them) and against a plain call alone."""
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def brook():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brook_amd
    return brook_amd


def test_served_requests_equal_oracle_alone_and_in_the_mix(brook):
    from brook_amd import synth
    from tests.golden.serving_digests import ANISOTROPY, REQUESTS
    from tests.golden.workload_digests import PARAMS, digest
    want = json.load(open(os.path.join(HERE, "golden", "serving_digests.json")))
    vols = [synth.serving_request(i) for i in range(REQUESTS)]
    opts = dict(teasar_params=PARAMS, anisotropy=ANISOTROPY, dust_threshold=1000)
    with brook.Server(slots=2, **opts) as srv:
        first = srv.submit(vols[0])
        deadline = time.time() + 60
        while first.request.admitted is None and time.time() < deadline:
            time.sleep(0.001)
        futures = [first] + [srv.submit(v) for v in vols[1:]]  # arrive while request 0 is on the GPU
        got = [f.result(300) for f in futures]
    reqs = [f.request for f in futures]
    assert any(r.overlapped for r in reqs), [(r.slot, r.batch_size) for r in reqs]
    assert len({r.slot for r in reqs}) == 2
    for i, g in enumerate(got):
        d = digest(g.as_dict())
        for name, value in d.items():
            assert value == want[str(i)][name], (i, name, reqs[i].slot, reqs[i].batch_size)
    for i in (0, REQUESTS - 1):  # alone: a plain call gives the same bytes
        alone = brook.skeletonize(vols[i], output="packed", progress=False, **opts).as_dict()
        for name in ("labels", "v_off", "e_off", "vertices", "edges", "radii"):
            assert np.array_equal(np.asarray(alone[name]), np.asarray(got[i].as_dict()[name])), (i, name)
