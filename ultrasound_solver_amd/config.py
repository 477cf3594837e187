"""YAML configuration with the reference's keys and checks.

Mirrors ``src/algorithm/config.py`` (dataclasses :9-166, ``from_yaml``
:169-213): the same sections (physics, transducer_specs, domain,
boundary_conditions, targets, obstacles, io, training, solver, algorithm,
algorithm_options), defaults, isotropic-grid check on dx (:68-83) and the
derived omega / k0 / wavelength (:146-160).  The ``solver:`` block stays a
free dict read with ``.get`` defaults, as in the reference.
"""
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Union

import numpy as np
import yaml


@dataclass
class PhysicsConfig:
    sound_speed: float = 343.0
    medium_density: float = 1.21
    material_field_file: str = ""


@dataclass
class TransducerSpecsConfig:
    frequency: float = 40000.0
    diameter: float = 0.0098
    pitch: float = 0.0105
    array_n: int = 6
    geometry: str = "ideal_grid"
    reference_anchor_m: List[float] = field(default_factory=lambda: [0.08, 0.08])
    source_radial_order: int = 64
    source_angular_order: int = 256

    def __post_init__(self):
        orders_ok = (isinstance(self.source_radial_order, int) and self.source_radial_order >= 2
                     and isinstance(self.source_angular_order, int) and self.source_angular_order >= 8)
        if not orders_ok:
            raise ValueError("Invalid source quadrature orders")
        if self.geometry not in ("ideal_grid", "sonicsurface_reference"):
            raise ValueError("geometry must be ideal_grid or sonicsurface_reference")
        if self.geometry == "sonicsurface_reference":
            raise ValueError("sonicsurface_reference geometry is not supported by this backend; use ideal_grid")
        anchor = np.asarray(self.reference_anchor_m, dtype=float)
        if anchor.shape != (2,) or not np.isfinite(anchor).all():
            raise ValueError("reference_anchor_m must contain two finite coordinates")


@dataclass
class DomainConfig:
    box_size: List[float] = field(default_factory=lambda: [0.08, 0.08, 0.08])
    grid_size: List[int] = field(default_factory=lambda: [81, 81, 81])
    origin_m: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])

    def __post_init__(self):
        o = np.asarray(self.origin_m, dtype=float)
        if o.shape != (3,) or not np.isfinite(o).all():
            raise ValueError("domain.origin_m must contain three finite coordinates")

    nx = property(lambda self: self.grid_size[0])
    ny = property(lambda self: self.grid_size[1])
    nz = property(lambda self: self.grid_size[2])
    lx = property(lambda self: self.box_size[0])
    ly = property(lambda self: self.box_size[1])
    lz = property(lambda self: self.box_size[2])

    @property
    def dx(self) -> float:
        steps = [self.box_size[a]/(self.grid_size[a] - 1) for a in range(3)]
        if not (np.isclose(steps[0], steps[1]) and np.isclose(steps[0], steps[2])):
            raise ValueError("The solver needs an isotropic grid (dx = dy = dz); got "
                             f"dx={steps[0]*1e3:.3f}mm, dy={steps[1]*1e3:.3f}mm, dz={steps[2]*1e3:.3f}mm")
        return float(steps[0])


@dataclass
class IOConfig:
    output_file: str = "ultrasound_field.npz"
    auto_visualize: bool = True
    colormap: str = "plasma"


@dataclass
class TrainingConfig:
    loss_type: str = "field_match"
    initial_phase: str = "geometric"
    compare_geometric: bool = False
    phase_basis_file: str = "phase_response_basis.npy"
    phase_basis_batch_size: int = 8
    voxel_chunk_size: int = 262144
    load_basis_to_gpu: bool = True
    release_factor_after_basis: bool = True
    target_field_file: str = ""
    target_peak_pressure: float = 600.0
    target_sigma: float = 0.006
    background_pressure: float = 0.0
    target_weight: float = 1.0
    background_weight: float = 0.05
    target_radius: float = 0.006
    source_exclusion_layers: int = 4
    sidelobe_temperature: float = 25.0
    iterations: int = 100
    seed: int = 0
    max_evaluations: int = 10000
    max_seconds: float = 0.0
    show_loss_curve: bool = True
    loss_curve_update_interval: int = 1
    loss_curve_pause_seconds: float = 0.01


@dataclass
class SimulationConfig:
    mode: str
    physics: PhysicsConfig
    specs: TransducerSpecsConfig
    domain: DomainConfig
    boundary_conditions: Dict[str, Union[str, dict]]
    targets: List[List[float]]
    obstacles: List[Dict]
    io: IOConfig = field(default_factory=IOConfig)
    training: TrainingConfig = field(default_factory=TrainingConfig)
    solver: Dict[str, Union[str, int, float, bool, dict]] = field(default_factory=dict)
    algorithm: str = "adjoint"
    algorithm_options: dict = field(default_factory=dict)
    same_phase_rad: float = 0.0

    def __post_init__(self):
        if self.mode not in ("phase_optimization", "same_phase", "sdf_inverse"):
            raise ValueError("mode must be phase_optimization, same_phase or sdf_inverse")
        if not np.isfinite(self.same_phase_rad):
            raise ValueError("same_phase_rad must be finite")
        if not isinstance(self.algorithm_options, dict):
            raise ValueError("algorithm_options must be a mapping")

    frequency = property(lambda self: self.specs.frequency)
    omega = property(lambda self: 2.0*np.pi*self.specs.frequency)
    k0 = property(lambda self: self.omega/self.physics.sound_speed)
    wavelength = property(lambda self: self.physics.sound_speed/self.specs.frequency)
    lx = property(lambda self: self.domain.lx)
    ly = property(lambda self: self.domain.ly)
    lz = property(lambda self: self.domain.lz)

    @classmethod
    def from_dict(cls, raw, base_dir=None):
        base = Path(base_dir) if base_dir is not None else Path.cwd()
        physics = dict(raw["physics"])
        if physics.get("material_field_file"):
            path = Path(physics["material_field_file"]).expanduser()
            physics["material_field_file"] = str((path if path.is_absolute() else base/path).resolve())
        obstacles = []
        for item in raw.get("obstacles", []):
            obs = dict(item)
            if obs.get("type", "").lower() == "mesh":
                if "file" not in obs:
                    raise ValueError("mesh obstacle requires a 'file' path")
                if "center_m" not in obs:
                    raise ValueError("mesh obstacle requires 'center_m'")
                path = Path(obs["file"]).expanduser()
                obs["file"] = str((path if path.is_absolute() else base/path).resolve())
            obstacles.append(obs)
        return cls(
            mode=str(raw.get("mode", "phase_optimization")).lower(),
            physics=PhysicsConfig(**physics),
            specs=TransducerSpecsConfig(**raw["transducer_specs"]),
            domain=DomainConfig(**raw["domain"]),
            boundary_conditions=raw["boundary_conditions"],
            targets=[t["point"] for t in raw.get("targets", [])],
            obstacles=obstacles,
            io=IOConfig(**raw.get("io", {})),
            training=TrainingConfig(**dict(raw.get("training", {}))),
            solver=raw.get("solver", {}),
            algorithm=str(raw.get("algorithm", "adjoint")).lower(),
            algorithm_options=raw.get("algorithm_options", {}),
            same_phase_rad=float(raw.get("same_phase_rad", 0.0)),
        )

    @classmethod
    def from_yaml(cls, yaml_path):
        path = Path(yaml_path).expanduser().resolve()
        with open(path, "r", encoding="utf-8") as fh:
            raw = yaml.safe_load(fh)
        return cls.from_dict(raw, base_dir=path.parent)
