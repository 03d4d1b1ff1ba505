"""GPU checks of the drop-in boundary beyond the skeletons: DLPack exports consumed by torch (zero
copy, the storage outliving the library's handles), device input fenced on its producer stream
instead of a host synchronization, and C-ordered device and host input gathered on the GPU."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
CFG1_PARAMS = dict(scale=1.5, const=300, pdrf_scale=100000, pdrf_exponent=4)


@pytest.fixture(scope="module")
def torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_dlpack_components_and_edt_equal_the_oracle(torch):
    from brook_amd import device, synth
    labels = synth.volume("tubes", (70, 50, 40), 6, 11)
    ids, mapping, roots = device.connected_components(labels)
    cc, m, r = oracle.ccl(labels)
    assert ids.device.type == "cuda" and tuple(ids.shape) == labels.shape
    assert ids.stride() == (1, labels.shape[0], labels.shape[0] * labels.shape[1])  # Fortran order
    assert np.array_equal(ids.cpu().numpy().astype(np.uint32), cc)
    assert np.array_equal(mapping, m) and np.array_equal(roots, r)
    dist = device.edt(labels, (16, 16, 40))
    want = oracle.edt(labels, (16, 16, 40), False)
    got = dist.cpu().numpy()
    assert dist.dtype == torch.float32 and np.array_equal(got.view(np.uint32), want.view(np.uint32))
    del ids, dist  # the exports' deleters free the library's memory
    torch.cuda.synchronize()


def test_device_input_waits_for_its_producer_stream(torch):
    import brook_amd
    from brook_amd import synth
    labels = synth.tubes_128(1001)
    want = brook_amd.skeletonize(labels, CFG1_PARAMS, (16, 16, 40), output="packed", progress=False)
    host = torch.from_numpy(np.ravel(labels, order="F")).pin_memory()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        dev = torch.zeros(labels.size, dtype=torch.int32, device="cuda")
        torch.cuda._sleep(20_000_000)  # the producer is still busy when skeletonize is called
        dev.copy_(host, non_blocking=True)
    vol = dev.view(labels.shape[::-1]).permute(2, 1, 0)
    got = brook_amd.skeletonize(vol, CFG1_PARAMS, (16, 16, 40), output="packed", progress=False, stream=side)
    for name in ("labels", "v_off", "e_off", "vertices", "edges", "radii"):
        assert np.array_equal(np.asarray(got.as_dict()[name]), np.asarray(want.as_dict()[name])), name


def test_c_ordered_input_gathered_on_the_gpu(torch):
    import brook_amd
    from brook_amd import synth
    labels = synth.volume("tubes", (96, 80, 72), 8, 3)
    want = brook_amd.skeletonize(np.asfortranarray(labels), CFG1_PARAMS, (16, 16, 40), output="packed",
                                 progress=False).as_dict()
    for name, vol in (("host C", np.ascontiguousarray(labels)), ("device C", torch.from_numpy(np.ascontiguousarray(labels)).cuda()),
                      ("host transposed", np.ascontiguousarray(labels.transpose(1, 0, 2)).transpose(1, 0, 2))):
        got = brook_amd.skeletonize(vol, CFG1_PARAMS, (16, 16, 40), output="packed", progress=False).as_dict()
        for key in ("labels", "vertices", "edges", "radii"):
            assert np.array_equal(np.asarray(got[key]), np.asarray(want[key])), (name, key)
