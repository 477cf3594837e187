"""Device aperture deposition (hz_deposit_sources behind
physics.source_deposition): bit-identical to the reference's own masses
(fixtures written by its sinc_aperture module) and to the oracle's
restatement at the full workload size (16 x 16 array, orders 64 x 256)."""
import json

import numpy as np
import pytest

from oracle import source as osrc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["array2x2_default_order", "array3x3_coarse_order"])
def test_deposition_matches_reference(golden_dir, name):
    from ultrasound_solver_amd import physics
    f = np.load(golden_dir/"source_fixtures.npz")
    case = json.loads(str(f["meta"]))[name]
    node, col, mass = physics.source_deposition(case["shape_xy"], case["dx"], f[f"{name}/centers"],
                                                case["diameter"], *case["orders"])
    # fixture is CSC over (node, column): same ordering (column, then node)
    np.testing.assert_array_equal(node, f[f"{name}/mass_indices"])
    np.testing.assert_array_equal(mass, f[f"{name}/mass_data"])
    counts = np.bincount(col, minlength=len(f[f"{name}/centers"]))
    np.testing.assert_array_equal(np.concatenate([[0], np.cumsum(counts)]), f[f"{name}/mass_indptr"])


def test_deposition_full_array_bit_identical_to_oracle():
    from ultrasound_solver_amd import physics, scenes
    sc = scenes.config_heterogeneous(128)
    node, col, mass = physics.source_deposition(sc.shape[:2], sc.dx, sc.centers, 2*sc.radius)
    ref = osrc.source_mass_matrix(sc.shape[:2], sc.dx, sc.centers, 2*sc.radius).tocsc()
    assert len(mass) == ref.nnz
    np.testing.assert_array_equal(node, ref.indices)
    np.testing.assert_array_equal(mass, ref.data)
    np.testing.assert_array_equal(np.concatenate([[0], np.cumsum(np.bincount(col, minlength=sc.num_sources))]),
                                  ref.indptr)
    np.testing.assert_allclose(np.bincount(col, weights=mass), 1.0, rtol=1e-13)


def test_deposition_rejects_clipped_apertures():
    from ultrasound_solver_amd import physics
    with pytest.raises(ValueError, match="clearance"):
        physics.source_deposition((20, 20), 1e-3, np.array([[1.5e-3, 10e-3, 0.0]]), 2e-3, 8, 16)
