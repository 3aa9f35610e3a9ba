"""The iterator's output contract and on-disk format (code/python/main.py,
code/python/store.py): snapshot prefix, file name, array layout, boundary
names."""

import os
import re

import numpy as np
import pytest


def test_snapshot_prefix_and_file_format(tmp_path):
    from phd_amd import store
    shape = (4, 3, 14)
    full = [np.full(int(np.prod(shape)), float(k)) for k in range(100)]
    assert len(store.snapshots(full, shape)) == 100
    # an early stop leaves the tail empty: the prefix ends at the last state
    early = full[:37] + [np.zeros(0)] * 63
    got = store.snapshots(early, shape)
    assert len(got) == 37 and got[-1].shape == shape and got[-1][0, 0, 0] == 36.0

    def sod():
        pass

    path = store.save_results(got, sod, 3, 0.8, False, 0, directory=str(tmp_path / "results"))
    name = os.path.basename(path)
    assert re.fullmatch(r"sod_3_0\.8_False_0_\d+\.npy", name), name
    arr = np.load(path)
    assert arr.shape == (37,) + shape and arr[5, 0, 0, 0] == 5.0


def test_boundary_names():
    from phd_amd import store
    assert store.boundary_codes("periodic", 2).tolist() == [1, 1, 1, 1]
    assert store.boundary_codes("transitive", 1).tolist() == [0, 0]
    assert store.boundary_codes("lid_driven", 2).tolist() == [3, 3, 3, 4]
    assert store.boundary_codes([5, 0, 0, 0], 2).tolist() == [5, 0, 0, 0]


@pytest.mark.gpu
def test_run_writes_iterator_snapshots(tmp_path):
    """A thermal Sod tube (48 cells, tf = 0.02) through store.run: 100
    snapshots at each 1% of tf, the last one the state at tf, as the
    reference's main.py saves them."""
    from phd_amd import gprpy, problems, store

    def sod48():
        u, MP, _, dX, bcs = problems.sod_1d(48, thermal=True)
        return u, [MP], 0.02, dX, "transitive"

    print("store.run on a 48-cell thermal Sod tube", flush=True)
    path = store.run(sod48, N=3, directory=str(tmp_path / "results"))
    arr = np.load(path)
    assert arr.shape == (100, 48, 17)
    assert np.isfinite(arr).all()
    u0, MP, _, dX, _ = problems.sod_1d(48, thermal=True)
    u = np.ascontiguousarray(u0.ravel())
    G = gprpy.load(3, True)
    G.solvers.iterator(u, 0.02, np.array([48], np.int32), np.array(dX), 0.8,
                       np.array([0, 0], np.int32), False, False, False, 0, [MP], 1.0)
    assert np.array_equal(arr[-1].ravel(), u)
