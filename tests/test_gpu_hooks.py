"""External-force hook on the GPU vs the oracle (run on an MI355X).

The GPU side runs the library's HIP counterpart of the tutorial damping helper
(launchAddGlobalDampingForceTorque) on the device views; the oracle side runs the same formula in
numpy on its host arrays.  Forces are compared each step to 1e-5 relative to the step's largest
body force, as in the other parity tests.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_binding import oracle_sim
from parity import step_and_compare
from test_oracle import numpy_damping

pytestmark = pytest.mark.gpu


def test_damping_hook_matches_oracle(L, S):
    sc = S.scene_c1(n=27, lattice=9.8e-3, jitter=0.2e-3)
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg))
    lib = gpu.lib
    f = lib.launchAddGlobalDampingForceTorque
    f.restype = None
    f.argtypes = [C.c_void_p] * 4 + [C.c_double, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p]
    seen = []

    def dev_hook(time, v):
        seen.append(time)
        f(v.force, v.torque, v.velocity, v.angular_velocity, 0.3, v.num, v.grid_dim, v.block_dim, v.stream)
    gpu.set_external_force(dev_hook)
    ref.set_external_force(lambda t, v: numpy_damping(v, 0.3))
    step_and_compare(gpu, ref, 10, lists=(0, 1), log=print)
    assert len(seen) == 10 and seen[0] == 0.0
    assert gpu.time() == ref.time()
    v = gpu.device_view()
    assert v.num == 27 and v.block_dim * v.grid_dim >= 27
