"""The autotuner on the GPU: it times real launch shapes, caches its choice,
and the chosen shape computes the same states as the default one."""
import json

import numpy as np
import pytest

from fvm2d_amd import autotune
from fvm2d_amd.solver import Inputs, Solver
from tests.helpers import case_kwargs, grid

pytestmark = pytest.mark.gpu


def test_autotune_chooses_and_caches(tmp_path):
    name = "euler_roe_lim_129x33"
    g = grid(name)
    inp = Inputs.from_kwargs(**case_kwargs(name))
    with Solver(g, inp) as s:
        geo = s.geo
    path = str(tmp_path / "tune.json")
    c = autotune.tune(g, geo, inp, cache_path=path, device="test", iters=16)
    assert (c.threads, c.patch_size) in autotune.shapes(geo.dims["np"], inp) and c.xcd in (0, 1)
    assert all(t > 0 for _, _, _, t in c.table) and {r[0] for r in c.table} == {256, 512}
    cache = json.load(open(path))
    assert len(cache) == 1 and list(cache.values())[0]["patch_size"] == c.patch_size
    states = []
    for ps, th in ((0, 256), (c.patch_size, c.threads)):
        with Solver(g, Inputs.from_kwargs(patch_size=ps, block_threads=th, **case_kwargs(name)),
                    geometry=geo) as s:
            s.tune(3, 64, c.xcd if ps else 1)
            h, _ = s.iterate(30)
            states.append((h, s.get_state()["prim"]))
    np.testing.assert_array_equal(states[0][1], states[1][1])
    np.testing.assert_array_equal(states[0][0][:, [0, 2, 3, 4, 5]], states[1][0][:, [0, 2, 3, 4, 5]])


def test_stream_copy_bandwidth_is_plausible():
    """The achievable-bandwidth figure the stage kernel is read against: a
    plain copy of 256 MiB reaches a large share of the 8 TB/s peak, and never
    more than the peak."""
    import ctypes
    from fvm2d_amd import gpu
    v = ctypes.c_double(0.0)
    assert gpu.lib().fvm_gpu_stream_copy(256 << 20, 5, ctypes.byref(v)) == 0
    assert 1000.0 < v.value < 8000.0
    assert gpu.lib().fvm_gpu_stream_copy(8, 5, ctypes.byref(v)) != 0     # below one 16-byte item
