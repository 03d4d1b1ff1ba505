"""The synthetic workload generators (host only): determinism and the stress volume's structure."""
import numpy as np

from brook_amd import synth


def test_cells_do_not_depend_on_thread_count():
    a, na, pa = synth.cells((96, 80, 40), 40, 7, threads=1)
    b, nb, pb = synth.cells((96, 80, 40), 40, 7, threads=5)
    assert (na, pa) == (nb, pb) and np.array_equal(a, b)
    assert pa == int(np.count_nonzero(a))


def test_cells_ids_and_erosion():
    v, cells, painted = synth.cells((128, 112, 48), 60, 11)
    assert v.dtype == np.uint64 and v.flags["F_CONTIGUOUS"] and painted > 0
    ids = np.unique(v[v != 0])
    cell, tube = ids >> np.uint64(32), ids & np.uint64(0xFFFFFFFF)
    assert cell.min() >= 1 and cell.max() <= cells
    assert tube.min() >= 1 and tube.max() <= 4
    # Cells are eroded by one voxel, so painted voxels of different cells never touch across a face.
    c = (v >> np.uint64(32)).astype(np.int64)
    for axis in range(3):
        a = np.take(c, range(c.shape[axis] - 1), axis=axis)
        b = np.take(c, range(1, c.shape[axis]), axis=axis)
        both = (a != 0) & (b != 0)
        assert np.all(a[both] == b[both])


def test_volumes_are_deterministic():
    a = synth.volume("tubes", (48, 40, 32), 5, 3)
    b = synth.volume("tubes", (48, 40, 32), 5, 3)
    assert np.array_equal(a, b) and a.dtype == np.uint32 and a.max() <= 5
