"""The FP64 operation counts bench.py uses as the roofline numerator are
those phd_amd/csrc/flopcount.cpp prints: that program instantiates the
kernels' own pointwise physics with a counting number type.  Also checks
relations the counts must satisfy: the per-cell totals are the per-line and
per-lane counts times the lines and lanes of a cell, the fallback's
iteration costs more than Picard's, and the solid's relaxation more than the
fluid's."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def counts():
    from phd_amd import build
    exe = build.FLOPCOUNT
    if not os.path.exists(exe):
        build.build_flopcount()
    return json.loads(subprocess.check_output([exe]).decode())


def test_bench_counts_are_the_instrumented_counts():
    import bench
    assert counts() == bench.FLOPS


def test_count_relations():
    c = counts()
    for k, f in c.items():
        assert f["fallback_node"] > f["picard_node"] > 0, k
        assert f["weno_cell"] == 14 * 4 * f["weno_line"], k
        assert f["fv_cell"] == 18 * f["fv_face_lane"] + 27 * f["fv_cell_node"], k
        assert 1 <= f["polar_iterations"] <= 40, k
    # the power-law solid's relaxation rate (a stress norm and a power) costs
    # more per node than the Newtonian fluid's
    assert c["gr_plastic_2d_p2"]["picard_node"] > c["gas_newtonian_2d_p2"]["picard_node"]
