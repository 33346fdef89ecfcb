"""The HIP path reproduces the reference's stored TwoBody Euler run bit for bit.

Same fixture and comparison as test_oracle_golden, with the GPU library as the engine:
730,500 Euler steps, either in fused runs on the device between snapshots (red_run_fixed)
or one red_step at a time, the barycentric transform and unit scaling done by the
library's host helpers, snapshots read back through red_get_state.
"""
import numpy as np
import pytest

from test_oracle_golden import compare_golden, run_twobody

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph,fused", [(True, True), (True, False)])
def test_gpu_twobody_euler_bit_exact(native_mod, graph, fused):
    print(f"case twobody graph={graph} fused={fused}", flush=True)

    def factory(b, prm):
        r = b.r.copy(); v = b.v.copy()
        assert np.all(r[0] == 0) and np.all(v[0] == 0)
        native_mod.host_transform_to_bc(b.mass, r, v)
        native_mod.host_transform_velocity(v)
        d = native_mod.Disk(b, integrator=prm["integrator"], dt_day=prm["dt"], use_graph=graph, r=r, v=v)
        return d, True

    snaps, info = run_twobody(factory, fused=fused)
    assert info["steps"] == 730500
    assert compare_golden(snaps) == 294
