"""The optimized CPU variant of the elastic assembly (a CPU baseline of the
benchmark, oracle/pa_cpu_fast.cpp) against the reference-structured
restatement, which the reference's goldens pin."""
import os

import numpy as np

import oracle


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


import pytest


@pytest.mark.parametrize("dims, threads", [((6, 3, 4), 8), ((16, 6, 6), 4), ((16, 6, 6), 1)])
def test_fast_cpu_variant_matches_restatement(dims, threads):
    """Small meshes take the colouring schedule, larger ones the two-phase chunks."""
    oracle.set_threads(threads)
    coords, conn = oracle.box_mesh(*dims, 2.0, 1.0, 1.0)
    fc, fl = oracle.side_set(coords, conn, 0, 2.0)
    om = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9,
                      face_cells=fc, face_lnodes=fl, flux=(0.0, 0.0, -1.0))
    fast = oracle.FastElastic(coords, conn, fc, fl, (0.0, 0.0, -1.0))
    assert fast.nnzb == om.nnzb and 1 < fast.colors < 64
    assert (fast.colors == 2) == (om.C >= 64 * 2 * threads)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(3 * om.N) * 1e-2
    z = rng.uniform(0.3, 1.0, om.N)
    R, K = fast.value_gradient_u(1.0, 0.3, 3.0, 1e-9, u, z)
    assert rel(R, om.value("elastic", u, z)) < 1e-12
    assert rel(K, om.gradient_u("elastic", u, z)) < 1e-12
    # a second call overwrites, it does not accumulate
    R2, K2 = fast.value_gradient_u(1.0, 0.3, 3.0, 1e-9, u, z)
    assert rel(R2, om.value("elastic", u, z)) < 1e-12 and rel(K2, om.gradient_u("elastic", u, z)) < 1e-12
    oracle.set_threads(os.cpu_count() or 1)
