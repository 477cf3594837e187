"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Golden vectors come from the reference's own CPU modules
(``tests/golden/make_fixtures.py``) and from its archived run
(``research_archive/.../phase_oblique_reflecting_x_12x12/adjoint/result.npz``);
closed forms and invariants cover what those leave open.
"""
import json

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import helmholtz, krylov, losses, materials, source


@pytest.fixture(scope="module")
def op_fix(golden_dir):
    f = np.load(golden_dir/"operator_fixtures.npz")
    return f, json.loads(str(f["meta"]))


def _oracle_op(f, meta, name):
    s = meta["scenes"][name]
    return helmholtz.OracleOperator(f[f"{name}/rho"], f[f"{name}/c"], meta["dx"], meta["omega"],
                                    s["boundaries"], s["pml"])


SCENES = ["mixed_faces", "open_pml_all", "soft_impedance_corner", "reflecting_box"]


@pytest.mark.parametrize("name", SCENES)
def test_operator_matches_reference_csr(op_fix, name):
    f, meta = op_fix
    op = _oracle_op(f, meta, name)
    a = op.matrix()
    ref = sp.csr_matrix((f[f"{name}/data"], f[f"{name}/indices"], f[f"{name}/indptr"]), shape=a.shape)
    assert a.nnz == ref.nnz
    assert abs(a - ref).max() == 0.0
    assert np.array_equal(op.fixed.ravel(order="F"), f[f"{name}/fixed"])


@pytest.mark.parametrize("name", SCENES)
def test_material_vjp_matches_reference(op_fix, name):
    f, meta = op_fix
    op = _oracle_op(f, meta, name)
    g_rho, g_c = op.material_vjp(f[f"{name}/vjp_u"], f[f"{name}/vjp_lambda"])
    np.testing.assert_array_equal(g_rho, f[f"{name}/grad_rho"])
    np.testing.assert_array_equal(g_c, f[f"{name}/grad_c"])


def test_planted_midpoint_mistake_is_caught(op_fix, monkeypatch):
    """A plausible slip (stretch taken at the low node instead of the edge
    midpoint) must fail the golden comparison."""
    f, meta = op_fix
    name = "open_pml_all"
    op = _oracle_op(f, meta, name)
    good = op.matrix()
    original = op.edges

    def wrong_edges():
        for axis, (lo, hi, factor) in enumerate(original()):
            s = op._axis_view(op.s1d[axis], axis)
            yield lo, hi, factor*(0.5*(s[lo] + s[hi]))/s[lo]
    monkeypatch.setattr(op, "edges", wrong_edges)
    bad = op.matrix()
    ref = sp.csr_matrix((f[f"{name}/data"], f[f"{name}/indices"], f[f"{name}/indptr"]), shape=good.shape)
    assert abs(good - ref).max() == 0.0
    assert abs(bad - ref).max() > 1e-6*abs(ref).max()


def test_planted_face_sign_mistake_is_caught(op_fix):
    f, meta = op_fix
    name = "mixed_faces"
    op = _oracle_op(f, meta, name)
    op.face_terms = [(sel, kind, -term) for sel, kind, term in op.face_terms]
    ref = sp.csr_matrix((f[f"{name}/data"], f[f"{name}/indices"], f[f"{name}/indptr"]))
    assert abs(op.matrix() - ref).max() > 0


def test_interior_rows_are_textbook_seven_point():
    n, dx, omega, rho, c = 9, 1.07e-3, 2*np.pi*40e3, 1.2, 343.0
    op = helmholtz.OracleOperator(np.full((n, n, n), rho), np.full((n, n, n), c), dx, omega, {})
    a = op.matrix()
    p = 4 + n*(4 + n*4)
    row = a.getrow(p)
    off = 1.0/(rho*dx*dx)
    expected = {p: omega**2/(rho*c*c) - 6*off}
    for step in (1, n, n*n):
        expected[p + step] = off
        expected[p - step] = off
    assert row.nnz == 7
    for col, val in zip(row.indices, row.data):
        assert val == pytest.approx(expected[col], rel=1e-14)


@pytest.mark.parametrize("name", ["open_pml_all", "reflecting_box"])
def test_rows_annihilate_constants_apart_from_mass_and_faces(op_fix, name):
    f, meta = op_fix
    op = _oracle_op(f, meta, name)
    a = op.matrix()
    mass = (op.mass_weight*op.omega**2/(op.rho*op.c**2))
    extra = np.zeros(op.shape, dtype=complex) + mass
    for sel, _, term in op.face_terms:
        extra[sel] += term
    rowsum = np.asarray(a.sum(axis=1)).ravel()
    np.testing.assert_allclose(rowsum, extra.ravel(order="F"), rtol=1e-9, atol=1e-9*abs(a).max())


@pytest.mark.parametrize("name", ["open_pml_all", "reflecting_box"])
def test_complex_symmetric_without_soft_faces(op_fix, name):
    f, meta = op_fix
    a = _oracle_op(f, meta, name).matrix()
    assert abs(a - a.T).max() == 0.0


def test_soft_faces_break_symmetry_only_in_fixed_rows(op_fix):
    f, meta = op_fix
    op = _oracle_op(f, meta, "mixed_faces")
    a = op.matrix()
    diff = (a - a.T).tocoo()
    fixed = op.fixed.ravel(order="F")
    bad = diff.data != 0
    assert np.all(fixed[diff.row[bad]] | fixed[diff.col[bad]])


def test_xy_swap_symmetry_on_cube():
    n = 8
    rng = np.random.default_rng(3)
    rho = rng.uniform(1, 2, (n, n, n))
    rho = 0.5*(rho + rho.transpose(1, 0, 2))
    c = np.full((n, n, n), 343.0)
    bnd = {"-x": "open", "+x": "open", "-y": "open", "+y": "open", "+z": "open"}
    op = helmholtz.OracleOperator(rho, c, 1e-3, 2*np.pi*40e3, bnd, {"cells": 2})
    a = op.matrix()
    perm = np.arange(n**3).reshape((n, n, n), order="F").transpose(1, 0, 2).ravel(order="F")
    pa = a[perm][:, perm]
    assert abs(pa - a).max() <= 1e-12*abs(a).max()


def test_operator_validation_errors():
    g = np.ones((6, 6, 6))
    with pytest.raises(ValueError):
        helmholtz.OracleOperator(g, g, 1e-3, 1.0, {"-q": "open"})
    with pytest.raises(ValueError):
        helmholtz.OracleOperator(g, g, 1e-3, 1.0, {"-x": "open"}, {"cells": 3, "faces": ["-x", "+x"]})
    with pytest.raises(ValueError):
        helmholtz.OracleOperator(g, g, 1e-3, 1.0, {"-x": "impedance"})


# --------------------------------------------------------------------- source
@pytest.fixture(scope="module")
def src_fix(golden_dir):
    f = np.load(golden_dir/"source_fixtures.npz")
    return f, json.loads(str(f["meta"]))


@pytest.mark.parametrize("name", ["array2x2_default_order", "array3x3_coarse_order"])
def test_source_matches_reference(src_fix, name):
    f, meta = src_fix
    case = meta[name]
    cen = f[f"{name}/centers"]
    m = source.source_mass_matrix(case["shape_xy"], case["dx"], cen, case["diameter"], *case["orders"]).tocsc()
    m.sort_indices()
    assert np.array_equal(m.indices, f[f"{name}/mass_indices"])
    np.testing.assert_array_equal(m.data, f[f"{name}/mass_data"])
    rhs = source.flux_rhs_columns(m.tocsr(), (*case["shape_xy"], 5), case["dx"], 1.21).tocsc()
    rhs.sort_indices()
    assert np.array_equal(rhs.indices, f[f"{name}/rhs_indices"])
    np.testing.assert_array_equal(rhs.data, f[f"{name}/rhs_data"])
    off, w = source.aperture_quadrature(case["diameter"], *case["orders"])
    np.testing.assert_array_equal(off, f[f"{name}/quad_offsets"])
    np.testing.assert_array_equal(w, f[f"{name}/quad_weights"])


def test_source_columns_unit_mass_and_centroid():
    nx, ny, dx = 30, 26, 1.07e-3
    cen = source.ideal_grid_centers(3, 6.1e-3, (nx - 1)*dx, (ny - 1)*dx)
    cen[:, 0] += 0.31*dx
    m = source.source_mass_matrix((nx, ny), dx, cen, 4.2e-3, 16, 32).tocsc()
    np.testing.assert_allclose(np.asarray(m.sum(axis=0)).ravel(), 1.0, rtol=0, atol=1e-14)
    ii, jj = np.meshgrid(np.arange(nx)*dx, np.arange(ny)*dx, indexing="ij")
    xs, ys = ii.ravel(order="F"), jj.ravel(order="F")
    np.testing.assert_allclose(m.T @ xs, cen[:, 0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(m.T @ ys, cen[:, 1], rtol=0, atol=1e-15)


def test_quadrature_integrates_sinc_density():
    """f(r) = 1/(2 pi a sqrt(a^2-r^2)) has unit mass and second moment
    (1/a) int_0^a r^3/sqrt(a^2-r^2) dr = 2 a^2/3."""
    a = 2.0e-3
    off, w = source.aperture_quadrature(2*a, 24, 64)
    assert w.sum() == pytest.approx(1.0, abs=1e-14)
    r2 = off[:, 0]**2 + off[:, 1]**2
    assert (w*r2).sum() == pytest.approx(2*a*a/3, rel=1e-12)


def test_source_clearance_error():
    with pytest.raises(ValueError):
        source.source_mass_matrix((10, 10), 1e-3, [[0.5e-3, 4e-3, 0.0]], 2e-3)


# ------------------------------------------------------------------ materials
def test_blend_closed_forms():
    """tests/test_material_blend.py:13-24 of the reference: 1/rho and 1/(rho c^2)
    are linear in h; end values exact."""
    h = np.linspace(0, 1, 11)
    rho, c = materials.blend(h, 1250.0, 2200.0, 1.2, 343.0)
    assert rho[0] == pytest.approx(1250.0, rel=1e-15) and c[0] == pytest.approx(2200.0, rel=1e-15)
    assert rho[-1] == pytest.approx(1.2, rel=1e-15) and c[-1] == pytest.approx(343.0, rel=1e-15)
    np.testing.assert_allclose(np.diff(1/rho, 2), 0, atol=1e-12)
    np.testing.assert_allclose(np.diff(1/(rho*c*c), 2), 0, atol=1e-18)


def test_smoothed_heaviside():
    s = np.array([-5.0, -1.2, 0.0, 1.2, 5.0, 0.6])
    h = materials.smoothed_heaviside(s, 1.2)
    np.testing.assert_allclose(h[:5], [0, 0, 0.5, 1, 1], atol=1e-15)
    t = 0.5
    assert h[5] == pytest.approx(0.5*(1 + t + np.sin(np.pi*t)/np.pi))


# --------------------------------------------------------------------- GMRES
def _small_scene(n=12, seed=0):
    rng = np.random.default_rng(seed)
    rho = 1.2*np.ones((n, n, n))
    c = 343.0*(1 + 0.03*rng.standard_normal((n, n, n)))
    bnd = {"-x": "open", "+x": "reflecting", "-y": "open", "+y": "open", "+z": "open"}
    op = helmholtz.OracleOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, {"cells": 3})
    return op


def test_reference_gmres_matches_direct_solve():
    op = _small_scene()
    a = op.matrix()
    rng = np.random.default_rng(5)
    b = np.zeros(a.shape[0], dtype=complex)
    b[rng.integers(0, a.shape[0], 5)] = rng.standard_normal(5)
    solver = krylov.ReferenceGmres(restart=30, maxiter=3000, rtol=1e-10)
    x = solver.solve(a, b)
    xd = krylov.direct_solve(a, b)
    assert np.linalg.norm(x - xd)/np.linalg.norm(xd) < 1e-8
    assert solver.last_iterations % 30 == 0 and solver.last_iterations > 0
    assert solver.last_scaled_residual <= 1e-10
    assert solver.last_backward_error <= 1e-4


def test_reference_gmres_zero_rhs_and_failure():
    op = _small_scene(8)
    a = op.matrix()
    s = krylov.ReferenceGmres()
    assert np.all(s.solve(a, np.zeros(a.shape[0])) == 0) and s.last_iterations == 0
    b = np.ones(a.shape[0], dtype=complex)
    with pytest.raises(RuntimeError, match="info="):
        krylov.ReferenceGmres(restart=2, maxiter=2, rtol=1e-12).solve(a, b)


def test_adjoint_matrix_is_conjugate_transpose():
    op = _small_scene(8)
    a = op.matrix()
    ah = krylov.adjoint_matrix(a)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(a.shape[0]) + 1j*rng.standard_normal(a.shape[0])
    y = rng.standard_normal(a.shape[0]) + 1j*rng.standard_normal(a.shape[0])
    assert np.vdot(y, a @ x) == pytest.approx(np.vdot(ah @ y, x), rel=1e-12)


# -------------------------------------------------------------------- losses
ARCH_SHAPE = (41, 41, 41)
ARCH_BOX = (0.12, 0.12, 0.12)
ARCH_DX = 0.12/40


@pytest.fixture(scope="module")
def arch(golden_dir):
    return np.load(golden_dir/"archive_fixture.npz")


def test_archived_centers_and_geometric_phases(arch):
    cen = source.ideal_grid_centers(12, 0.01, 0.12, 0.12)
    np.testing.assert_array_equal(cen, arch["source_positions"])
    gp = source.geometric_phases(cen, arch["target_points"], 343/40e3, 2*np.pi*40e3/343)
    # The archived run stored -k d; the reference's current closed form equals it mod 2 pi.
    np.testing.assert_allclose(gp, np.mod(arch["geometric_phases"], 2*np.pi), atol=1e-12)


def test_archived_gaussian_target(arch):
    t, w = losses.gaussian_target(ARCH_SHAPE, ARCH_BOX, (0, 0, 0), ARCH_DX, arch["target_points"],
                                  600.0, 0.006, 0.0, 1.0, 1.3)
    idx = tuple(arch["sample_index"].T)
    np.testing.assert_allclose(t[idx], arch["target_amplitude_sample"], rtol=1e-14, atol=1e-12)
    np.testing.assert_array_equal(w[idx], arch["target_weight_sample"])
    np.testing.assert_allclose(t[24:37, 24:37, 24:37], arch["target_amplitude_patch"], rtol=1e-14, atol=1e-12)


@pytest.mark.parametrize("amp_key,focal_key,loss_key", [
    ("amplitude", "target_value_final", "final_loss"),
    ("geometric_amplitude", "target_value_geometric", "first_loss")])
def test_archived_focal_contrast_loss(arch, amp_key, focal_key, loss_key):
    mask = losses.background_mask(ARCH_SHAPE, ARCH_BOX, arch["target_points"], 0.008, 4)
    idx = [losses.nearest_index(p, (0, 0, 0), ARCH_DX) for p in arch["target_points"]]
    amp = arch[amp_key].astype(np.complex128)
    loss, _ = losses.focal_contrast(amp, idx, 1.0, 1.3, 5.0, mask, False,
                                    focal_field=np.array([arch[focal_key]]))
    assert loss == pytest.approx(float(arch[loss_key]), rel=1e-14)


def test_archived_metrics(arch):
    t, _ = losses.gaussian_target(ARCH_SHAPE, ARCH_BOX, (0, 0, 0), ARCH_DX, arch["target_points"],
                                  600.0, 0.006, 0.0, 1.0, 1.3)
    mask = losses.background_mask(ARCH_SHAPE, ARCH_BOX, arch["target_points"], 0.008, 4)
    m = losses.metrics(arch["geometric_amplitude"], t, [abs(arch["target_value_geometric"])], mask)
    ref = json.loads(str(arch["geometric_metrics"]))
    for key in ("target_mean", "background_max", "contrast", "field_relative_l2", "field_absolute_l2"):
        assert m[key] == pytest.approx(ref[key], rel=1e-13)


def test_worked_cotangent_example():
    """u = 3+4i, target 6, weight 1: c = -0.6-0.8i, dL = Re(c^* du) = -0.022."""
    u = np.array([3 + 4j])
    _, c = losses.field_match(u, np.array([6.0]), np.array([1.0]))
    assert c[0] == pytest.approx(-0.6 - 0.8j, abs=1e-12)
    assert np.real(np.conj(c[0])*(0.01 + 0.02j)) == pytest.approx(-0.022, abs=1e-12)


@pytest.mark.parametrize("kind", ["field_match", "focal_pressure", "focal_contrast", "focal_contrast_ratio"])
def test_phase_vjp_matches_central_differences(kind):
    shape = (7, 6, 5)
    n, m = int(np.prod(shape)), 9
    rng = np.random.default_rng(7)
    g = rng.standard_normal((n, m)) + 1j*rng.standard_normal((n, m))
    target = np.abs(rng.standard_normal(shape))*3
    weight = rng.uniform(0.1, 1.0, shape)
    mask = np.ones(shape, bool)
    mask[3, 3, 2] = False
    idx = [(3, 3, 2), (1, 2, 3)]

    def evaluate(ph, cot=True):
        u = losses.basis_field(g, ph, shape)
        if kind == "field_match":
            return losses.field_match(u, target, weight, cot)
        if kind == "focal_pressure":
            return losses.focal_pressure(u, idx, 1.0, 0.3, cot)
        if kind == "focal_contrast_ratio":
            return losses.focal_contrast_ratio(u, ~mask, mask, cot)
        return losses.focal_contrast(u, idx, 1.0, 1.3, 2.0, mask, cot)

    ph = rng.uniform(0, 2*np.pi, m)
    _, cot = evaluate(ph)
    grad = losses.phase_vjp(g, cot, ph)
    fd = losses.central_difference(lambda p: evaluate(p, False)[0], ph, 1e-5, count=m)
    for i in range(m):
        assert losses.scaled_gradient_error(grad[i], fd[i]) < 1e-7


def test_trilinear_target_response_reproduces_linear_fields():
    shape = (6, 5, 4)
    dx = 1e-3
    i, j, k = np.meshgrid(*[np.arange(s)*dx for s in shape], indexing="ij")
    g = np.stack([(2*i + 3*j - k).ravel(order="F"), (1j*i + 0.5).ravel(order="F")], 1)
    pts = np.array([[1.3e-3, 2.2e-3, 0.7e-3], [5e-3, 4e-3, 3e-3]])
    r = losses.target_response(g, pts, shape, (0, 0, 0), dx)
    np.testing.assert_allclose(r[:, 0], 2*pts[:, 0] + 3*pts[:, 1] - pts[:, 2], atol=1e-15)
    np.testing.assert_allclose(r[:, 1], 1j*pts[:, 0] + 0.5, atol=1e-15)
