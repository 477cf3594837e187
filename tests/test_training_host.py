"""Host-side layers on the CPU: config, losses and cotangents against the
oracle, torch autograd cotangents, basis storage and resume."""
import json
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import losses as olo
from ultrasound_solver_amd import autograd as hza
from ultrasound_solver_amd import training
from ultrasound_solver_amd.config import SimulationConfig

ARCH = {"shape": (41, 41, 41), "box": (0.12, 0.12, 0.12)}


def archive_cfg(loss="focal_contrast"):
    raw = {
        "mode": "phase_optimization",
        "physics": {"sound_speed": 343.0, "medium_density": 1.21},
        "transducer_specs": {"frequency": 40000.0, "diameter": 0.01, "pitch": 0.01, "array_n": 12},
        "domain": {"box_size": [0.12]*3, "grid_size": [41]*3},
        "boundary_conditions": {"-x": "open", "+x": "reflecting", "-y": "open", "+y": "open",
                                "-z": "reflecting", "+z": "open"},
        "targets": [{"point": [0.09, 0.09, 0.09]}],
        "obstacles": [],
        "training": {"loss_type": loss, "target_peak_pressure": 600.0, "target_sigma": 0.006,
                     "background_weight": 1.3, "target_radius": 0.008, "source_exclusion_layers": 4,
                     "sidelobe_temperature": 5.0},
        "solver": {"backend": "gpu_iterative"},
    }
    return SimulationConfig.from_dict(raw)


def test_config_mirrors_reference_checks(tmp_path):
    cfg = archive_cfg()
    assert cfg.domain.dx == pytest.approx(0.003)
    assert cfg.k0 == pytest.approx(2*np.pi*40000/343)
    bad = dict(box_size=[0.1, 0.12, 0.12], grid_size=[41, 41, 41])
    with pytest.raises(ValueError, match="isotropic"):
        SimulationConfig.from_dict({**_raw_min(), "domain": bad}).domain.dx
    with pytest.raises(ValueError, match="mode"):
        SimulationConfig.from_dict({**_raw_min(), "mode": "nope"})
    p = tmp_path/"c.yaml"
    import yaml
    p.write_text(yaml.safe_dump(_raw_min()))
    assert SimulationConfig.from_yaml(p).specs.array_n == 4


def _raw_min():
    return {"mode": "same_phase", "physics": {}, "transducer_specs": {"array_n": 4},
            "domain": {"box_size": [0.04]*3, "grid_size": [41]*3}, "boundary_conditions": {}}


def test_targets_mask_and_loss_match_archive(golden_dir):
    a = np.load(golden_dir/"archive_fixture.npz")
    cfg = archive_cfg()
    target = training.AmplitudeTarget.from_config(cfg)
    idx = tuple(a["sample_index"].T)
    np.testing.assert_allclose(target.target[idx], a["target_amplitude_sample"], rtol=1e-14, atol=1e-12)
    np.testing.assert_array_equal(target.weight[idx], a["target_weight_sample"])
    loss = training.FocalContrastLoss(cfg, target)
    amp = a["amplitude"].astype(np.complex128)
    val, _ = loss.evaluate(amp, with_cotangent=False, focal_field=np.array([a["target_value_final"]]))
    assert val == pytest.approx(float(a["final_loss"]), rel=1e-14)
    dev, _ = loss.evaluate_device(torch.as_tensor(amp), with_cotangent=False,
                                  focal_field=torch.as_tensor(np.array([a["target_value_final"]])))
    assert float(dev) == pytest.approx(float(a["final_loss"]), rel=1e-13)


def _random_field(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j*rng.standard_normal(shape)


@pytest.mark.parametrize("kind", ["field_match", "focal_pressure", "focal_contrast"])
def test_losses_and_cotangents_match_oracle(kind):
    cfg = archive_cfg(kind)
    tgt = training.AmplitudeTarget.from_config(cfg)
    loss = training.make_loss(cfg, tgt)
    u = _random_field(ARCH["shape"], 1)*200
    idx = [training.nearest_grid_index(cfg, p) for p in cfg.targets]
    mask = olo.background_mask(ARCH["shape"], ARCH["box"], cfg.targets, 0.008, 4)
    if kind == "field_match":
        ref = olo.field_match(u, tgt.target, tgt.weight)
    elif kind == "focal_pressure":
        ref = olo.focal_pressure(u, idx, 1.0, 1.3)
    else:
        ref = olo.focal_contrast(u, idx, 1.0, 1.3, 5.0, mask)
    got = loss.evaluate(u)
    assert got[0] == pytest.approx(ref[0], rel=1e-13)
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-12, atol=1e-12*np.abs(ref[1]).max())
    gd = loss.evaluate_device(torch.as_tensor(u))
    assert float(gd[0]) == pytest.approx(ref[0], rel=1e-12)
    np.testing.assert_allclose(gd[1].numpy(), ref[1], rtol=1e-10, atol=1e-12*np.abs(ref[1]).max())


@pytest.mark.parametrize("kind", ["field_match", "focal_pressure", "focal_contrast"])
def test_autograd_losses_give_reference_cotangents(kind):
    cfg = archive_cfg(kind)
    tgt = training.AmplitudeTarget.from_config(cfg)
    u_np = _random_field(ARCH["shape"], 2)*150
    u = torch.tensor(u_np, requires_grad=True)
    idx = [training.nearest_grid_index(cfg, p) for p in cfg.targets]
    mask = olo.background_mask(ARCH["shape"], ARCH["box"], cfg.targets, 0.008, 4)
    if kind == "field_match":
        val = hza.field_match_loss(u, torch.as_tensor(tgt.target), torch.as_tensor(tgt.weight))
        ref = olo.field_match(u_np, tgt.target, tgt.weight)
    elif kind == "focal_pressure":
        val = hza.focal_pressure_loss(u, idx, 1.0, 1.3)
        ref = olo.focal_pressure(u_np, idx, 1.0, 1.3)
    else:
        val = hza.focal_contrast_loss(u, idx, 1.0, 1.3, 5.0, torch.as_tensor(mask))
        ref = olo.focal_contrast(u_np, idx, 1.0, 1.3, 5.0, mask)
    val.backward()
    assert float(val.detach()) == pytest.approx(ref[0], rel=1e-12)
    np.testing.assert_allclose(u.grad.numpy(), ref[1], rtol=1e-8, atol=1e-10*np.abs(ref[1]).max())


class _FakeSolver:
    """Stands in for the device solver to exercise storage and resume logic."""
    MODEL_VERSION = "fake"

    def __init__(self, n, m, fail_after=None):
        self.cfg = archive_cfg()
        self.total_dofs = n
        self.transducers = SimpleNamespace(num_transducers=m, amplitudes=np.ones(m),
                                           centers=np.zeros((m, 3)))
        self.source_columns = np.ones((n, m))
        self.calls = 0
        self.fail_after = fail_after
        fake = self

        class Op:
            def build_rhs(self, w):
                return w

        class It:
            def solve(self, op, b):
                if fake.fail_after is not None and fake.calls >= fake.fail_after:
                    raise KeyboardInterrupt("simulated kill")
                fake.calls += 1
                cols = torch.nonzero(b)[:, 1]
                return torch.stack([torch.full((n,), float(c) + 1j, dtype=torch.complex128) for c in cols])

        self.operator = Op()
        self.iterative_solver = It()

    def assemble_matrix(self):
        return None

    def scene_fingerprint(self):
        return "scene"


def test_basis_build_resumes_after_interruption(tmp_path):
    path = tmp_path/"G.npy"
    solver = _FakeSolver(30, 10, fail_after=2)
    basis = training.PhaseResponseBasis(solver, path, rhs_batch_size=3)
    with pytest.raises(KeyboardInterrupt):
        basis.build()
    meta = json.loads((tmp_path/"G.npy.json").read_text())
    assert meta["completed_columns"] == 6 and meta["format_version"] == 3
    basis.close()
    solver2 = _FakeSolver(30, 10)
    b2 = training.PhaseResponseBasis(solver2, path, rhs_batch_size=3).build()
    assert solver2.calls == 2  # columns 6..9 only
    g = np.asarray(b2.basis)
    np.testing.assert_array_equal(g, np.broadcast_to(np.arange(10) + 1j, (30, 10)))


def test_basis_rejects_changed_scene(tmp_path):
    path = tmp_path/"G.npy"
    s = _FakeSolver(12, 4)
    training.PhaseResponseBasis(s, path, rhs_batch_size=4).build().close()
    s2 = _FakeSolver(12, 4)
    s2.scene_fingerprint = lambda: "other"
    with pytest.raises(RuntimeError, match="does not match"):
        training.PhaseResponseBasis(s2, path).build()


def _homogeneous_cfg(tmp_path):
    h = 1.07e-3
    n = 25
    return SimulationConfig.from_dict({
        "mode": "phase_optimization",
        "physics": {"sound_speed": 343.0, "medium_density": 1.2},
        "transducer_specs": {"frequency": 40000.0, "diameter": 0.0034, "pitch": 0.0042, "array_n": 2,
                             "source_radial_order": 16, "source_angular_order": 32},
        "domain": {"box_size": [(n - 1)*h]*3, "grid_size": [n]*3},
        "boundary_conditions": {"-x": "open", "+x": "reflecting", "-y": "open", "+y": "open",
                                "-z": "reflecting", "+z": "open"},
        "targets": [{"point": [(n - 1)*h*0.5]*3}],
        "solver": {"backend": "gpu_iterative", "pml": {"cells": 4},
                   "target_response_cache": str(tmp_path/"responses.npz")},
    })


def test_target_points_and_response_cache(tmp_path):
    """Target validation (inside the domain, outside the PML) and the response
    cache file: round trip, and a key that changes with the scene."""
    from ultrasound_solver_amd.response_cache import ResponseCache
    from ultrasound_solver_amd.solvers import HelmholtzSolver
    cfg = _homogeneous_cfg(tmp_path)
    s = HelmholtzSolver(cfg)
    h = cfg.domain.dx
    with pytest.raises(ValueError, match="inside the PDE domain"):
        s._local_points(np.array([[0.0, 0.0, -h]]))
    with pytest.raises(ValueError, match="lies in PML"):
        s._local_points(np.array([[2*h, 12*h, 5*h]]))
    with pytest.raises(ValueError, match="lies in PML"):
        s._local_points(np.array([[12*h, 21*h, 5*h]]))
    pts = np.array([[12*h, 12*h, 8*h], [10*h, 13*h, 3*h]])
    assert np.allclose(s._local_points(pts), pts)
    cache = s._response_cache()
    assert cache.path == tmp_path/"responses.npz" and len(cache.points) == 0
    assert list(cache.lookup(pts)) == [-1, -1]
    resp = np.arange(8).reshape(2, 4)*(1 + 2j)
    cache.store(pts, resp)
    again = s._response_cache()
    np.testing.assert_array_equal(again.points, pts)
    np.testing.assert_array_equal(again.rows, resp)
    assert list(again.lookup(pts[::-1] + 5e-13)) == [1, 0]
    s._ensure_target_surface_solver = None  # every point is cached
    np.testing.assert_array_equal(s.target_response(pts[::-1]), resp[::-1])
    cfg2 = _homogeneous_cfg(tmp_path)
    cfg2.boundary_conditions["+x"] = "open"
    assert len(HelmholtzSolver(cfg2)._response_cache().points) == 0
    # A file with another key or malformed arrays reads as empty.
    np.savez_compressed(tmp_path/"other.npz", fingerprint=np.asarray("x"), points_m=pts, response=resp)
    assert len(ResponseCache(tmp_path/"other.npz", "y", 4).points) == 0
    np.savez_compressed(tmp_path/"bad.npz", fingerprint=np.asarray("y"), points_m=pts, response=resp[:, :3])
    assert len(ResponseCache(tmp_path/"bad.npz", "y", 4).points) == 0


def test_scene_fingerprint_tracks_materials_and_files(tmp_path):
    from ultrasound_solver_amd.solvers import HelmholtzSolver
    cfg = _homogeneous_cfg(tmp_path)
    s = HelmholtzSolver(cfg)
    base = s.scene_fingerprint()
    assert HelmholtzSolver(_homogeneous_cfg(tmp_path)).scene_fingerprint() == base
    rho = np.full(s.shape, 1.2)
    c = np.full(s.shape, 343.0)
    s.set_material_fields(rho, c)
    with_arrays = s.scene_fingerprint()
    assert with_arrays != base
    c[3, 4, 5] = 344.0
    s.set_material_fields(rho, c)
    assert s.scene_fingerprint() not in (base, with_arrays)
    with pytest.raises(ValueError, match="finite and positive"):
        s.set_material_fields(rho, -c)
    path = tmp_path/"materials.npz"
    np.savez(path, density=rho, sound_speed=c)
    cfg3 = _homogeneous_cfg(tmp_path)
    cfg3.physics.material_field_file = str(path)
    s3 = HelmholtzSolver(cfg3)
    first = s3.scene_fingerprint()
    assert first == s3.scene_fingerprint()
    np.savez(path, density=rho, sound_speed=c + 1.0)
    assert s3.scene_fingerprint() != first
    cfg.domain.grid_size = [27]*3
    with pytest.raises(ValueError, match="construct a new solver"):
        s.scene_fingerprint()


def test_focal_contrast_ratio_loss_matches_oracle():
    cfg = archive_cfg("focal_contrast_ratio")
    loss = training.make_loss(cfg)
    ball, rest = olo.ratio_masks(ARCH["shape"], ARCH["box"], cfg.targets, 0.008)
    np.testing.assert_array_equal(loss.ball, ball)
    np.testing.assert_array_equal(loss.rest, rest)
    assert ball.any() and rest.any() and not (ball & rest).any()
    u = _random_field(ARCH["shape"], 3)*40
    ref = olo.focal_contrast_ratio(u, ball, rest)
    got = loss.evaluate(u)
    assert got[0] == pytest.approx(ref[0], rel=1e-13)
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-12, atol=1e-14*np.abs(ref[1]).max())
    gd = loss.evaluate_device(torch.as_tensor(u))
    assert float(gd[0]) == pytest.approx(ref[0], rel=1e-12)
    np.testing.assert_allclose(gd[1].numpy(), ref[1], rtol=1e-10, atol=1e-14*np.abs(ref[1]).max())
