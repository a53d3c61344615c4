"""Host set-up against the reference: initial conditions, scaled metrics
and reconstruction coefficients (golden vectors from the reference Fortran)."""

import numpy as np
import pytest

from helpers import case, golden


@pytest.mark.parametrize("tc", [1, 2, 5, 6])
def test_initial_conditions(tc):
    g, _ = golden("l3_initial_conditions")
    mesh, phys = case(3, tc)
    for k in ("h", "u", "h_s", "fVertex", "fEdge", "lonCell"):
        np.testing.assert_array_equal(phys[k], g[f"tc{tc}_{k}"], err_msg=k)


def test_scaled_metrics_and_coefficients():
    g, _ = golden("l3_initial_conditions")
    mesh, phys = case(3, 2)
    np.testing.assert_array_equal(phys["areaCell"], g["areaCell"])
    np.testing.assert_array_equal(phys["kiteAreasOnVertex"].T.ravel(), g["kiteAreasOnVertex"])
    mine = np.transpose(phys["coeffs_reconstruct"], (2, 1, 0)).ravel()
    np.testing.assert_array_equal(mine, g["coeffs_reconstruct"])
