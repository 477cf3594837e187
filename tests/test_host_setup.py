"""Host-side scene setup of the product against the pinned oracle."""
import json

import numpy as np
import pytest

from oracle import helmholtz as ohz, source as osrc
from ultrasound_solver_amd import physics


@pytest.fixture(scope="module")
def op_fix(golden_dir):
    f = np.load(golden_dir/"operator_fixtures.npz")
    return f, json.loads(str(f["meta"]))


@pytest.mark.parametrize("name", ["mixed_faces", "open_pml_all", "soft_impedance_corner", "reflecting_box"])
def test_axis_and_face_tables_match_oracle(op_fix, name):
    f, meta = op_fix
    s = meta["scenes"][name]
    rho, c = f[f"{name}/rho"], f[f"{name}/c"]
    ref = ohz.OracleOperator(rho, c, meta["dx"], meta["omega"], s["boundaries"], s["pml"])
    w, st = physics.axis_tables(rho.shape, s["boundaries"], s["pml"])
    for a in range(3):
        np.testing.assert_array_equal(w[a], ref.w1d[a])
        np.testing.assert_array_equal(st[a], ref.s1d[a])
    kinds, tables = physics.face_tables(rho, c, meta["dx"], meta["omega"], s["boundaries"], w, st)
    terms = list(ref.face_terms)
    for idx, table in enumerate(tables):
        if table is None:
            continue
        sel, kind, term = terms.pop(0)
        assert kind == kinds[idx]
        np.testing.assert_array_equal(table, np.asarray(term).ravel(order="F"))
    assert not terms
    np.testing.assert_array_equal(physics.soft_node_mask(rho.shape, kinds), ref.fixed)


def test_validation_matches_reference_messages():
    with pytest.raises(ValueError, match="PML leaves no physical interior"):
        physics.axis_tables((6, 6, 6), {"-x": "open", "+x": "open"}, {"cells": 3})
    with pytest.raises(ValueError, match="PML may only be placed on open faces"):
        physics.axis_tables((9, 9, 9), {"-x": "soft"}, {"cells": 2, "faces": ["-x"]})
    with pytest.raises(ValueError, match="Unknown boundary face"):
        physics.axis_tables((9, 9, 9), {"up": "open"})


def test_centers_and_phases_match_archive(golden_dir):
    a = np.load(golden_dir/"archive_fixture.npz")
    cen = physics.ideal_grid_centers(12, 0.01, 0.12, 0.12)
    np.testing.assert_array_equal(cen, a["source_positions"])
    gp = physics.geometric_phases(cen, a["target_points"], 343/40e3, 2*np.pi*40e3/343)
    np.testing.assert_allclose(gp, osrc.geometric_phases(cen, a["target_points"], 343/40e3,
                                                         2*np.pi*40e3/343), atol=0)


def test_surface_axis_matrices_match_the_oracle():
    """Host setup of the surface solver: the 1-D axis operators (PML, open and
    impedance faces) equal the oracle's restatement of the reference's."""
    from oracle import surface as osu
    from ultrasound_solver_amd import surface
    faces = {"-x": "open", "+x": "open", "-y": {"type": "impedance", "impedance": [823.2, 205.8]},
             "+y": "reflecting", "-z": "reflecting", "+z": "open"}
    pml = {"cells": 4, "strength": 1.7, "order": 2}
    shape = (23, 19, 17)
    ops = surface.axis_operators(shape, 1.1e-3, 2*np.pi*40e3, 1.2, 343.0, faces, pml)
    for axis, n in enumerate(shape):
        got, s_got = ops[axis]
        want, s_want = osu.axis_matrix(n, 1.1e-3, 2*np.pi*40e3, 1.2, 343.0, faces, pml, axis)
        np.testing.assert_array_equal(s_got, s_want)
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)
    with pytest.raises(ValueError, match="open/reflecting/impedance"):
        surface.axis_operators((10, 4, 4), 1e-3, 1.0, 1.2, 343.0, {"-x": "soft"}, {})


def test_iterative_solver_options_and_their_checks():
    """The backend's optional solver keys and their validation (no device
    work: GpuIterativeSolver only stores them)."""
    from ultrasound_solver_amd.solvers import GpuIterativeSolver
    s = GpuIterativeSolver(precision="mixed", preconditioner="multigrid", chunk="auto", mixed_restart=12,
                           relative_rtol=1e-6)
    assert (s.chunk, s.mixed_restart, s.restart, s.relative_rtol) == (0, 12, 30, 1e-6)
    assert GpuIterativeSolver().relative_rtol == 0.0
    for bad, msg in ((dict(relative_rtol=-1.0), "relative_rtol"), (dict(precision="fp32"), "gmres_precision"),
                     (dict(preconditioner="multigrid"), "mixed form"), (dict(mixed_restart=0), "restart")):
        with pytest.raises(ValueError, match=msg):
            GpuIterativeSolver(**bad)
