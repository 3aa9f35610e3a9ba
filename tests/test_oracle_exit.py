"""The oracle's Picard exits (CPU): the early exit the benchmark uses
(bail_from) must leave every cell's values as the reference's rule (Picard
runs to DG_IT = 50) gives them on a state where Picard diverges in part of
the domain, and it must take fewer Picard iterations there."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))

from orcbind import Oracle  # noqa: E402


def test_bail_from_same_states_as_picard_to_dg_it():
    from phd_amd import problems
    n = 48
    u, MP, _, dX, bcs = problems.double_shear_2d(n)
    nX = [n, n]
    ref = Oracle(3, False, 2, bailout=0)
    fast = Oracle(3, False, 2, bailout=0, bail_from=6)
    ur, uf = u.ravel().copy(), u.ravel().copy()
    t, c = 0.0, 0
    seen_fallback = 0
    for step in range(10):
        ur, tr, cr, dtr, rc, sr = ref.step(ur, nX, dX, 0.8, bcs, t, c, 1e300, MP)
        assert rc == 0
        uf, tf_, cf, dtf, rc, sf = fast.step(uf, nX, dX, 0.8, bcs, t, c, 1e300, MP)
        assert rc == 0
        assert np.array_equal(ur, uf), step
        assert sr["fallback_cells"] == sf["fallback_cells"]
        if sr["fallback_cells"]:
            seen_fallback += 1
            assert sf["picard_iters"] < sr["picard_iters"]
        t, c = tr, cr
    assert seen_fallback >= 3

