"""The drop-in solver, phase basis and autograd path on the GPU against the oracle.

Tolerances: fields within 1e-5 relative L2 of complex128 sparse LU on the
oracle matrix; phase gradients within 1e-4 relative of the oracle basis VJP
and of central finite differences (step 1e-4 rad, the workload's FD
protocol), scaled error as in the reference's gradient check.
"""
import numpy as np
import pytest

from oracle import helmholtz as ohz, krylov as okr, losses as olo, source as osrc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def small_cfg(tmp_path=None, precision="fp64", pre="jacobi", loss="focal_contrast", n=25, obstacle=True,
              top="soft"):
    from ultrasound_solver_amd.config import SimulationConfig
    h = 1.07e-3
    raw = {
        "mode": "phase_optimization",
        "physics": {"sound_speed": 343.0, "medium_density": 1.2},
        "transducer_specs": {"frequency": 40000.0, "diameter": 0.0034, "pitch": 0.0042, "array_n": 2,
                             "source_radial_order": 16, "source_angular_order": 32},
        "domain": {"box_size": [(n - 1)*h]*3, "grid_size": [n]*3},
        "boundary_conditions": {"-x": "open", "+x": "reflecting", "-y": "open", "+y": "open",
                                "-z": "reflecting", "+z": top},
        "targets": [{"point": [(n - 1)*h*0.5, (n - 1)*h*0.5, (n - 1)*h*0.6]}],
        "obstacles": [{"type": "sphere", "center": [(n - 1)*h*0.55, (n - 1)*h*0.5, (n - 1)*h*0.75],
                       "radius": 3*h}] if obstacle else [],
        "training": {"loss_type": loss, "phase_basis_file": str((tmp_path or ".")) + "/G.npy",
                     "phase_basis_batch_size": 3, "background_weight": 0.8, "target_radius": 0.004,
                     "source_exclusion_layers": 3, "sidelobe_temperature": 2.0},
        "solver": {"backend": "gpu_iterative", "gmres_rtol": 1e-9, "gmres_maxiter": 20000,
                   "gmres_precision": precision, "gmres_preconditioner": pre, "gmres_chunk": 4,
                   "pml": {"cells": 4}},
    }
    return SimulationConfig.from_dict(raw)


def oracle_system(solver):
    """The oracle's A and B for the solver's scene (materials from the same inputs)."""
    from ultrasound_solver_amd import physics
    cfg = solver.cfg
    rho, c = solver._materials()
    a = ohz.OracleOperator(rho, c, cfg.domain.dx, cfg.omega, cfg.boundary_conditions,
                           cfg.solver.get("pml")).matrix()
    centers = solver.transducers.centers - solver.origin
    m = osrc.source_mass_matrix(cfg.domain.grid_size[:2], cfg.domain.dx, centers, cfg.specs.diameter,
                                cfg.specs.source_radial_order, cfg.specs.source_angular_order)
    b = osrc.flux_rhs_columns(m, cfg.domain.grid_size, cfg.domain.dx, cfg.physics.medium_density)
    return a, b


@pytest.mark.parametrize("precision,pre", [("fp64", "jacobi"), ("mixed", "jacobi"), ("mixed", "multigrid")])
def test_solver_solve_matches_direct(precision, pre):
    from ultrasound_solver_amd.solvers import HelmholtzDirectSolver
    s = HelmholtzDirectSolver(small_cfg(precision=precision, pre=pre))
    s.set_phases(np.array([0.0, 1.0, 2.0, 3.0]))
    u = s.solve()
    a, b = oracle_system(s)
    ref = okr.direct_solve(a, b @ np.exp(1j*np.array([0.0, 1.0, 2.0, 3.0])))
    got = u.ravel(order="F")
    assert np.linalg.norm(got - ref)/np.linalg.norm(ref) < 1e-5
    lam = s.solve_adjoint(u)
    ref_l = okr.direct_solve(okr.adjoint_matrix(a), got)
    assert np.linalg.norm(lam.ravel(order="F") - ref_l)/np.linalg.norm(ref_l) < 1e-5
    s.close()


def test_basis_build_field_vjp_and_target_response(tmp_path):
    from ultrasound_solver_amd.solvers import HelmholtzDirectSolver
    cfg = small_cfg(tmp_path, precision="mixed", pre="multigrid")
    s = HelmholtzDirectSolver(cfg)
    basis = s.build_phase_response_basis()
    a, b = oracle_system(s)
    g_ref = okr.direct_solve(a, b.toarray())
    g = np.asarray(basis.basis)
    for col in range(g.shape[1]):
        assert np.linalg.norm(g[:, col] - g_ref[:, col])/np.linalg.norm(g_ref[:, col]) < 1e-5
    rng = np.random.default_rng(0)
    ph = rng.uniform(0, 2*np.pi, 4)
    u = basis.field(ph)
    np.testing.assert_allclose(u.ravel(order="F"), g_ref @ np.exp(1j*ph), rtol=0,
                               atol=1e-5*np.abs(g_ref @ np.exp(1j*ph)).max())
    cot = rng.standard_normal(u.shape) + 1j*rng.standard_normal(u.shape)
    grad = basis.phase_vjp(cot, ph)
    ref = olo.phase_vjp(g_ref, cot, ph)
    assert np.max(np.abs(grad - ref))/np.max(np.abs(ref)) < 1e-5
    basis.to_gpu()
    grad_dev = basis.phase_vjp(cot, ph)
    assert np.max(np.abs(grad_dev - ref))/np.max(np.abs(ref)) < 1e-5
    pts = np.asarray(cfg.targets)
    resp_basis = basis.target_response(pts)
    resp_recip = s.target_response(pts)
    assert np.max(np.abs(resp_recip - resp_basis))/np.max(np.abs(resp_basis)) < 1e-5
    basis.close()
    s.close()


def test_homogeneous_target_response_is_exact_and_cached(tmp_path):
    """Homogeneous scene: target_response comes from the surface condensation
    and equals the trilinear rows of the direct-solve basis to 1e-9; a second
    call reads the response cache file; field_at_direct agrees with field_at."""
    from ultrasound_solver_amd.solvers import HelmholtzDirectSolver
    cfg = small_cfg(tmp_path, obstacle=False, top="open")
    cfg.solver["target_response_cache"] = str(tmp_path/"responses.npz")
    s = HelmholtzDirectSolver(cfg)
    a, b = oracle_system(s)
    g_ref = okr.direct_solve(a, b.toarray())
    n = cfg.domain.grid_size[0]
    h = cfg.domain.dx
    pts = np.array([[0.5, 0.5, 0.6], [0.41, 0.55, 0.3], [0.6, 0.45, 0.12]])*(n - 1)*h
    resp = s.target_response(pts)
    assert s._target_surface_solver is not None and s.operator is None  # no 3-D operator was built
    ref = olo.target_response(g_ref, pts, (n, n, n), (0, 0, 0), h)
    assert np.max(np.abs(resp - ref))/np.max(np.abs(ref)) < 1e-9
    s2 = HelmholtzDirectSolver(cfg)
    s2._ensure_target_surface_solver = None  # a cache hit must not need the solver
    np.testing.assert_array_equal(s2.target_response(pts[::-1]), resp[::-1])
    ph = np.random.default_rng(3).uniform(0, 2*np.pi, 4)
    np.testing.assert_allclose(s.field_at_direct(pts, ph), s.field_at(pts, ph), rtol=1e-10)
    s.close()


@pytest.mark.parametrize("loss", ["field_match", "focal_pressure", "focal_contrast"])
def test_adjoint_phase_gradient_matches_basis_and_fd(loss):
    from ultrasound_solver_amd import autograd as hza, training
    from ultrasound_solver_amd.solvers import HelmholtzDirectSolver
    cfg = small_cfg(loss=loss, precision="mixed", pre="multigrid")
    s = HelmholtzDirectSolver(cfg)
    s.assemble_matrix()
    shape = tuple(cfg.domain.grid_size)
    eng = hza.PhaseEngine(s.operator, s.transducers.amplitudes, precision="c64", preconditioner="multigrid",
                          rtol=1e-9, maxiter=20000, chunk=8)
    tgt = training.AmplitudeTarget.from_config(cfg)
    idx = [training.nearest_grid_index(cfg, p) for p in cfg.targets]
    mask = torch.as_tensor(training.build_background_mask(cfg, cfg.targets, 0.004, 3), device="cuda")
    target_t = torch.as_tensor(tgt.target, device="cuda")
    weight_t = torch.as_tensor(tgt.weight, device="cuda")

    def loss_of(u):
        v = hza.grid_view(u, shape)
        if loss == "field_match":
            return hza.field_match_loss(v, target_t, weight_t)
        if loss == "focal_pressure":
            return hza.focal_pressure_loss(v, idx, 1.0, 0.8)
        return hza.focal_contrast_loss(v, idx, 1.0, 0.8, 2.0, mask)

    rng = np.random.default_rng(1)
    ph0 = rng.uniform(0, 2*np.pi, 4)
    ph = torch.tensor(ph0, device="cuda", requires_grad=True)
    val = loss_of(hza.phase_field(ph, eng))
    val.backward()
    grad = ph.grad.cpu().numpy()
    # Oracle: basis VJP with the reference's cotangent.
    a, b = oracle_system(s)
    g_ref = okr.direct_solve(a, b.toarray())
    u_ref = olo.basis_field(g_ref, ph0, shape)
    mask_np = mask.cpu().numpy()
    if loss == "field_match":
        lval, cot = olo.field_match(u_ref, tgt.target, tgt.weight)
    elif loss == "focal_pressure":
        lval, cot = olo.focal_pressure(u_ref, idx, 1.0, 0.8)
    else:
        lval, cot = olo.focal_contrast(u_ref, idx, 1.0, 0.8, 2.0, mask_np)
    assert float(val.detach()) == pytest.approx(lval, rel=1e-5)
    ref = olo.phase_vjp(g_ref, cot, ph0)
    for i in range(4):
        assert olo.scaled_gradient_error(grad[i], ref[i]) < 1e-4
    # Central differences through the device solve (delta = 1e-4 rad).
    with torch.no_grad():
        for i in range(2):
            e = np.zeros(4)
            e[i] = 1e-4
            lp = float(loss_of(hza.phase_field(torch.tensor(ph0 + e, device="cuda"), eng)))
            lm = float(loss_of(hza.phase_field(torch.tensor(ph0 - e, device="cuda"), eng)))
            fd = (lp - lm)/2e-4
            assert olo.scaled_gradient_error(grad[i], fd) < 1e-4
    s.close()


def test_batched_phases_give_per_row_gradients():
    from ultrasound_solver_amd import autograd as hza
    from ultrasound_solver_amd.solvers import HelmholtzDirectSolver
    cfg = small_cfg(precision="mixed", pre="multigrid")
    s = HelmholtzDirectSolver(cfg)
    s.assemble_matrix()
    eng = hza.PhaseEngine(s.operator, s.transducers.amplitudes, preconditioner="multigrid", rtol=1e-9, chunk=8)
    rng = np.random.default_rng(2)
    phs = torch.tensor(rng.uniform(0, 2*np.pi, (3, 4)), device="cuda", requires_grad=True)
    u = hza.phase_field(phs, eng)
    (u.abs()**2).sum().backward()
    for r in range(3):
        p1 = phs[r].detach().clone().requires_grad_(True)
        (hza.phase_field(p1, eng).abs()**2).sum().backward()
        np.testing.assert_allclose(phs.grad[r].cpu().numpy(), p1.grad.cpu().numpy(), rtol=1e-6, atol=1e-9)
    s.close()


def _oracle_objective(s, cfg, loss, g_ref, ph):
    """(loss, phase gradient) of PhaseProblem.evaluate (common.py:275-343) from
    the oracle's basis: exact focal samples through the trilinear target
    response, the grid cotangent's basis VJP plus the focal gradient."""
    shape = tuple(cfg.domain.grid_size)
    box = cfg.domain.box_size
    pts = np.asarray(cfg.targets)
    idx = [olo.nearest_index(p, cfg.domain.origin_m, cfg.domain.dx) for p in pts]
    u = olo.basis_field(g_ref, ph, shape)
    if loss == "field_match":
        from ultrasound_solver_amd import training
        tgt = training.AmplitudeTarget.from_config(cfg)
        val, cot = olo.field_match(u, tgt.target, tgt.weight)
        return val, olo.phase_vjp(g_ref, cot, ph)
    if loss == "focal_contrast_ratio":
        ball, rest = olo.ratio_masks(shape, box, pts, cfg.training.target_radius)
        val, cot = olo.focal_contrast_ratio(u, ball, rest)
        return val, olo.phase_vjp(g_ref, cot, ph)
    resp = olo.target_response(g_ref, pts, shape, cfg.domain.origin_m, cfg.domain.dx)
    focal = resp @ np.exp(1j*ph)
    tr = cfg.training
    if loss == "focal_pressure":
        val, cot = olo.focal_pressure(u, idx, tr.target_weight, tr.background_weight, focal_field=focal)
    else:
        mask = olo.background_mask(shape, box, pts, tr.target_radius, tr.source_exclusion_layers)
        val, cot = olo.focal_contrast(u, idx, tr.target_weight, tr.background_weight, tr.sidelobe_temperature,
                                      mask, focal_field=focal)
    grad = olo.phase_vjp(g_ref, cot, ph) + olo.focal_gradient(resp, focal, ph, tr.target_weight)
    return val, grad


@pytest.mark.parametrize("route", ["basis", "adjoint"])
@pytest.mark.parametrize("loss", ["field_match", "focal_pressure", "focal_contrast", "focal_contrast_ratio"])
def test_phase_objective_matches_oracle_and_gradient_check(tmp_path, loss, route):
    """PhaseObjective.evaluate (common.py:275-343) through the device basis
    (fused loss kernels) and through forward + adjoint solves: loss and phase
    gradient against the oracle's basis composition, including the focal
    gradient of exact samples (common.py:189-195); then the reference's
    central-difference check (adjoint/__init__.py:44-58) on 4 emitters."""
    from ultrasound_solver_amd.solvers import HelmholtzDirectSolver
    from ultrasound_solver_amd.training import PhaseObjective
    cfg = small_cfg(tmp_path, precision="mixed", pre="multigrid", loss=loss)
    s = HelmholtzDirectSolver(cfg)
    if route == "basis":
        obj = PhaseObjective(s, basis=s.build_phase_response_basis())
    else:
        obj = PhaseObjective(s, engine_kwargs=dict(rtol=1e-10, preconditioner="multigrid"))
    a, b = oracle_system(s)
    g_ref = okr.direct_solve(a, b.toarray())
    ph = np.random.default_rng(5).uniform(0, 2*np.pi, 4)
    val, grad = obj.evaluate(ph, gradient=True)
    ref_val, ref_grad = _oracle_objective(s, cfg, loss, g_ref, ph)
    assert val == pytest.approx(ref_val, rel=1e-6)
    for i in range(4):
        assert olo.scaled_gradient_error(grad[i], ref_grad[i]) < 1e-5, (i, grad, ref_grad)
    # The reference's default step (gradient_check_step 1e-3).  Through the
    # basis the loss is exact up to rounding; through solves to 1e-10 the
    # difference quotient carries ~1e-10 |L| / 1e-3 of noise (|L| ~ 1e9 for
    # field_match), hence the workload's 1e-4 bar there.
    checks = obj.gradient_check(ph, step=1e-3, count=4)
    assert [c["index"] for c in checks] == [0, 1, 2, 3]
    for c in checks:
        assert c["scaled_error"] < (1e-5 if route == "basis" else 1e-4), checks
    assert obj.field_evaluations == 10 and obj.vjp_evaluations == 2
    assert obj.history[-1]["stage"] == "gradient_check"
    s.close()


@pytest.mark.parametrize("loss", ["field_match", "focal_pressure", "focal_contrast", "focal_contrast_ratio"])
def test_phase_objective_graph_replay_equals_launches(tmp_path, loss):
    """PhaseObjective(use_graph=True): the device step (e^{i phi}, exact focal
    samples, the allocation-free fused pass, the phase gradient) captured
    once and replayed per evaluation gives the loss and gradient of the same
    step launched op by op, bit for bit, on every replay: replays with and
    without gradient alternate with evaluations of the launched objective
    and host reads in between."""
    from ultrasound_solver_amd.solvers import HelmholtzDirectSolver
    from ultrasound_solver_amd.training import PhaseObjective
    cfg = small_cfg(tmp_path, precision="mixed", pre="multigrid", loss=loss)
    s = HelmholtzDirectSolver(cfg)
    basis = s.build_phase_response_basis()
    eager = PhaseObjective(s, basis=basis)
    graph = PhaseObjective(s, basis=basis, use_graph=True)
    rng = np.random.default_rng(21)
    for q in range(6):
        ph = rng.uniform(0, 2*np.pi, 4)
        gradient = q % 3 != 2
        ve, ge = eager.evaluate(ph, gradient=gradient)
        vg, gg = graph.evaluate(ph, gradient=gradient)
        assert vg == ve, (q, vg, ve)
        if gradient:
            assert np.array_equal(gg, ge), (q, gg, ge)
        else:
            assert gg is None and ge is None
        torch.cuda.synchronize()
    assert len(graph._graphs) == 2
    assert graph.field_evaluations == 6 and graph.vjp_evaluations == 4
    s.close()
