"""pa_scalar_function_value_noalloc (partial sums in the library's static
device array, no allocation in the call) against pa_scalar_function_value,
through the C ABI: two calls in a row, the second on a smaller mesh, both
physics.  Both run the same kernels in the same order, so the values agree
bit for bit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import platoanalyze_amd
    assert platoanalyze_amd.device_count() >= 1
    return platoanalyze_amd


def _value(pa, fn, sf, u, z):
    out = pa.DeviceArray(1)
    pa._check(fn(sf.handle, u.ptr, z.ptr, out.ptr, None))
    pa.synchronize()
    return float(out.numpy()[0])


@pytest.mark.parametrize("kind", ["Internal Elastic Energy", "Internal Thermal Energy"])
def test_noalloc_value_matches(pa, kind):
    dpn = 3 if "Elastic" in kind else 1
    mat = dict(youngs=1.0, poisson=0.3) if dpn == 3 else dict(conductivity=1.0)
    for dims in [(96, 48, 48), (5, 3, 2)]:  # the second call is on a smaller mesh
        print(f"case: noalloc criterion value, {kind}, {dims}", flush=True)
        mesh = pa.Mesh.box(*dims, 2.0, 1.0, 1.0, side_sets=False)
        rng = np.random.default_rng(sum(dims))
        N = mesh.num_nodes
        u = pa.DeviceArray.from_numpy(rng.standard_normal(dpn * N) * 1e-2)
        z = pa.DeviceArray.from_numpy(rng.uniform(0.1, 1.0, N))
        sf = pa.PhysicsScalarFunction(mesh, kind, penalty_exponent=3.0, penalty_min=1e-9, **mat)
        a = _value(pa, pa.lib.pa_scalar_function_value_noalloc, sf, u, z)
        b = _value(pa, pa.lib.pa_scalar_function_value, sf, u, z)
        assert a == b and np.isfinite(a) and a > 0, (a, b)
