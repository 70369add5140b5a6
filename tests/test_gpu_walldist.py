"""fvm_gpu_wall_dist against the host wall distance (src/wall_dist.f
restated in fvm_geometric, itself pinned to the reference): bit for bit."""
import numpy as np
import pytest

from fvm2d_amd import mesh
from fvm2d_amd.solver import geometric
from tests.helpers import fixture

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("compat", [1, 0])
@pytest.mark.parametrize("grid", ["97x25", "513x129"])
def test_wall_distance_gpu_equals_host(compat, grid):
    g = mesh.o_grid(97, 25, seed=11, first_spacing=2e-3) if grid == "97x25" else mesh.naca0012_513x129()
    host = geometric(g.coord, g.elem, g.ptype, iflow=3, sa_compat=compat)
    dev = geometric(g.coord, g.elem, g.ptype, iflow=3, sa_compat=compat, wall_distance="gpu")
    np.testing.assert_array_equal(dev.wd, host.wd)
    np.testing.assert_array_equal(dev.wd1, host.wd1)


def test_wall_distance_gpu_equals_reference():
    fx = fixture("rans_roe_97x25")
    g = mesh.o_grid(97, 25, seed=11, t=0.12, first_spacing=2e-3)
    dev = geometric(g.coord, g.elem, g.ptype, iflow=3, sa_compat=1, wall_distance="gpu")
    np.testing.assert_array_equal(dev.wd, fx["wd"])
