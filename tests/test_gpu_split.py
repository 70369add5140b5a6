"""Split precision: reconstruction and Riemann flux in FP32, geometry,
accumulation, RK update, SA variable and norms in FP64.  Bar: |dCL|, |dCD|
<= 1e-4 against the FP64 reference (golden histories written by the
reference itself), and the residual within FP32 accuracy of FP64."""
import numpy as np
import pytest

from fvm2d_amd.solver import Inputs, Solver
from tests.golden.cases import CASES
from tests.helpers import case_kwargs, fixture, grid, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["euler_roe_lim_129x33", "euler_kfvs_lim_129x33",
                                  "euler_roe_transonic_65x17", "laminar_roe_97x25",
                                  "rans_roe_97x25"])
def test_split_precision_lift_drag(name):
    fx = fixture(name)
    n = CASES[name]["iters"]
    with Solver(grid(name), Inputs.from_kwargs(precision=1, **case_kwargs(name))) as s:
        hist, _ = s.iterate(n)
    ref = fx["hist"]
    assert hist.shape == ref.shape
    dcl = np.abs(hist[:, 4] - ref[:, 4]).max()
    dcd = np.abs(hist[:, 5] - ref[:, 5]).max()
    assert dcl <= 1e-4 and dcd <= 1e-4, (dcl, dcd)
    # the split path is really different from FP64 (not a silent FP64 run)
    assert np.abs(hist[-1, 4] - ref[-1, 4]) > 0.0 or np.abs(hist[-1, 1] - ref[-1, 1]) > 0.0


@pytest.mark.parametrize("name", ["euler_roe_lim_129x33", "restart_kfvs_129x33"])
def test_split_precision_residual(name):
    """Near convergence the residual is a small difference of O(1) edge
    fluxes; FP32 fluxes are accurate to ~1e-7 of the flux scale
    p_inf*max|ds|, which is what the residual error is measured against."""
    fx = fixture(name)
    prim = fx["primN"] if "primN" in fx else fx["prim0"]
    res_ref = fx["resN"] if "resN" in fx else fx["res0"]
    with Solver(grid(name), Inputs.from_kwargs(precision=1, **case_kwargs(name))) as s:
        res = s.residu(prim, fx["nutN"] if "nutN" in fx else None)
        ds = s.geo.ds
    p_inf = 1.0 / (1.4 * CASES[name]["deck"]["Mach"] ** 2)
    scale = p_inf * np.sqrt((ds ** 2).sum(axis=1)).max()
    err = np.abs(res - res_ref).max() / scale
    assert 1e-14 < err < 1e-5, err
