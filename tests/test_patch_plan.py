"""Host-side plan of the LDS-tiled stage kernel: vertex patches, their halos
and every index the kernel will dereference, checked without a GPU."""
import pytest

from fvm2d_amd import mesh
from fvm2d_amd.solver import geometric, patch_plan


@pytest.mark.parametrize("cell", [1, 2])
@pytest.mark.parametrize("iflow", [1, 2, 3])
@pytest.mark.parametrize("ps", [16, 64, 128, 256])
def test_plan_valid(cell, iflow, ps):
    g = mesh.o_grid(97, 25, seed=11, jitter=0.1, first_spacing=2e-3)
    geo = geometric(g.coord, g.elem, g.ptype, cell_type=cell, iflow=iflow)
    st = patch_plan(geo, iflow, ps)
    assert st["failed_check"] == 0
    assert st["max_patch"] <= ps
    assert st["lds_bytes"] <= 58 * 1024
    assert st["patches"] * st["max_patch"] >= geo.dims["np"]


def test_plan_workload_grid():
    g = mesh.naca0012_513x129()
    geo = geometric(g.coord, g.elem, g.ptype)
    st = patch_plan(geo, 1, 64)
    assert st["failed_check"] == 0 and st["patches"] == 66048 // 64
    big = patch_plan(geo, 1, 256, lds_limit=150 * 1024)
    assert big["failed_check"] == 0 and big["max_patch"] == 256
