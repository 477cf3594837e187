"""Write the golden operator/source fixtures from the reference's own CPU code.

Run in the development container only (the reference tree is not present
anywhere else):

    python tests/golden/make_fixtures.py /root/reference

It imports exactly two CPU-only reference modules, both numpy/scipy only:

* ``src/algorithm/physics/conservative_helmholtz.py`` -- ``ConservativeHelmholtz``
  (``__init__`` :18-103, ``assemble`` :119-149, ``material_vjp`` :151-175);
* ``src/algorithm/physics/sinc_aperture.py`` -- ``aperture_quadrature`` :19-33,
  ``source_mass_matrix`` :36-72, ``flux_rhs_columns`` :75-85.

These are the functions whose edge/corner entries (PML meeting impedance and
soft faces) and exact bilinear deposition values no closed form pins
bit-for-bit.  The outputs are stored as plain arrays (CSR triplets, floats) in
``operator_fixtures.npz`` and ``source_fixtures.npz``; nothing of the reference
source travels with them.

The archived run ``research_archive/algorithm_comparison/results/
phase_oblique_reflecting_x_12x12/adjoint/result.npz`` is reference-held data;
``archive_fixture.npz`` keeps the subset of its arrays that pins the loss,
target-field, geometric-phase and transducer-centre computations.
"""
import importlib.util
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _load(reference, relative):
    path = Path(reference) / relative
    spec = importlib.util.spec_from_file_location(path.stem, path)
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


# Scenes chosen so that every face kind and every PML/face intersection occurs.
OPERATOR_SCENES = {
    "mixed_faces": dict(
        shape=(9, 8, 7), seed=11,
        boundaries={"-x": "open", "+x": {"type": "impedance", "impedance": [412.0, 103.0]},
                    "-y": "soft", "+y": "open", "-z": "reflecting", "+z": "open"},
        pml={"cells": 2, "strength": 2.0, "order": 3}),
    "open_pml_all": dict(
        shape=(10, 9, 8), seed=12,
        boundaries={"-x": "open", "+x": "open", "-y": "open", "+y": "open",
                    "-z": "reflecting", "+z": "open"},
        pml={"cells": 3, "strength": 2.5, "order": 2}),
    "soft_impedance_corner": dict(
        shape=(8, 8, 8), seed=13,
        boundaries={"-x": "soft", "+x": "open",
                    "-y": {"type": "impedance", "impedance": [823.2, 205.8]},
                    "+y": {"type": "impedance", "impedance": [823.2, 205.8]},
                    "-z": "reflecting", "+z": "soft"},
        pml={"cells": 2, "strength": 2.0, "order": 3, "faces": ["+x"]}),
    "reflecting_box": dict(
        shape=(6, 5, 4), seed=14,
        boundaries={}, pml=None),
}
DX = 1.07e-3
OMEGA = 2*np.pi*40000.0


def scene_materials(shape, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(1.0, 1.6, shape)
    c = rng.uniform(320.0, 380.0, shape)
    return rho, c


def operator_fixtures(reference):
    ch = _load(reference, "src/algorithm/physics/conservative_helmholtz.py")
    out = {}
    for name, scene in OPERATOR_SCENES.items():
        rho, c = scene_materials(scene["shape"], scene["seed"])
        op = ch.ConservativeHelmholtz(rho, c, DX, OMEGA, scene["boundaries"], (),
                                      pml=scene["pml"], device="cpu")
        a = op.assemble().tocsr()
        a.sort_indices()
        rng = np.random.default_rng(scene["seed"] + 100)
        n = rho.size
        u = rng.standard_normal(n) + 1j*rng.standard_normal(n)
        lam = rng.standard_normal(n) + 1j*rng.standard_normal(n)
        grad_rho, grad_c = op.material_vjp(u, lam)
        out[f"{name}/data"] = a.data
        out[f"{name}/indices"] = a.indices.astype(np.int64)
        out[f"{name}/indptr"] = a.indptr.astype(np.int64)
        out[f"{name}/shape"] = np.asarray(scene["shape"])
        out[f"{name}/rho"] = rho
        out[f"{name}/c"] = c
        out[f"{name}/fixed"] = op.fixed.ravel(order="F")
        out[f"{name}/vjp_u"] = u
        out[f"{name}/vjp_lambda"] = lam
        out[f"{name}/grad_rho"] = grad_rho
        out[f"{name}/grad_c"] = grad_c
    meta = {name: {"boundaries": s["boundaries"], "pml": s["pml"], "seed": s["seed"],
                   "shape": list(s["shape"])}
            for name, s in OPERATOR_SCENES.items()}
    out["meta"] = np.asarray(json.dumps({"dx": DX, "omega": OMEGA, "scenes": meta}))
    np.savez_compressed(HERE/"operator_fixtures.npz", **out)


SOURCE_CASES = {
    "array2x2_default_order": dict(shape_xy=(16, 14), dx=1.0e-3, pitch=4.5e-3,
                                   diameter=4.0e-3, n=2, orders=(64, 256)),
    "array3x3_coarse_order": dict(shape_xy=(21, 23), dx=1.07e-3, pitch=5.3e-3,
                                  diameter=3.1e-3, n=3, orders=(6, 16)),
}


def source_fixtures(reference):
    sa = _load(reference, "src/algorithm/physics/sinc_aperture.py")
    out = {}
    for name, case in SOURCE_CASES.items():
        nx, ny = case["shape_xy"]
        dx, pitch, n = case["dx"], case["pitch"], case["n"]
        span = (n - 1)*pitch*0.5
        xs = (nx - 1)*dx*0.5 - span + np.arange(n)*pitch
        ys = (ny - 1)*dx*0.5 - span + np.arange(n)*pitch
        centers = np.array([[x, y, 0.0] for x in xs for y in ys])
        masses = sa.source_mass_matrix((nx, ny), dx, centers, case["diameter"], *case["orders"]).tocsc()
        masses.sort_indices()
        rhs = sa.flux_rhs_columns(masses.tocsr(), (nx, ny, 5), dx, 1.21).tocsc()
        rhs.sort_indices()
        offsets, weights = sa.aperture_quadrature(case["diameter"], *case["orders"])
        out[f"{name}/centers"] = centers
        out[f"{name}/mass_data"] = masses.data
        out[f"{name}/mass_indices"] = masses.indices.astype(np.int64)
        out[f"{name}/mass_indptr"] = masses.indptr.astype(np.int64)
        out[f"{name}/rhs_data"] = rhs.data
        out[f"{name}/rhs_indices"] = rhs.indices.astype(np.int64)
        out[f"{name}/rhs_indptr"] = rhs.indptr.astype(np.int64)
        out[f"{name}/quad_offsets"] = offsets
        out[f"{name}/quad_weights"] = weights
    out["meta"] = np.asarray(json.dumps(SOURCE_CASES))
    np.savez_compressed(HERE/"source_fixtures.npz", **out)


def archive_fixture(reference):
    folder = Path(reference)/"research_archive/algorithm_comparison/results/phase_oblique_reflecting_x_12x12/adjoint"
    with np.load(folder/"result.npz", allow_pickle=False) as data:
        run = json.loads(str(data["run_metadata"]))
        log = json.loads(str(data["evaluation_log"]))
        target_amp = data["target_amplitude"]
        target_w = data["target_weight"]
        # The Gaussian target is pinned on a patch around the focus plus a seeded sample.
        rng = np.random.default_rng(0)
        sample = rng.integers(0, 41, size=(600, 3))
        out = {
            "amplitude": np.abs(data["u_complex"]),
            "geometric_amplitude": np.abs(data["geometric_u_complex"]),
            "target_value_final": data["u_complex"][30, 30, 30],
            "target_value_geometric": data["geometric_u_complex"][30, 30, 30],
            "phases": data["phases"],
            "geometric_phases": data["geometric_phases"],
            "source_positions": data["source_positions"],
            "target_points": data["target_points"],
            "sample_index": sample,
            "target_amplitude_sample": target_amp[tuple(sample.T)],
            "target_weight_sample": target_w[tuple(sample.T)],
            "target_amplitude_patch": target_amp[24:37, 24:37, 24:37],
            "target_weight_patch": target_w[24:37, 24:37, 24:37],
            "final_loss": np.float64(run["final_loss"]),
            "first_loss": np.float64(log[0]["loss"]),
            "geometric_metrics": np.asarray(str(data["geometric_metrics"])),
        }
    np.savez_compressed(HERE/"archive_fixture.npz", **out)


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    operator_fixtures(ref)
    source_fixtures(ref)
    archive_fixture(ref)
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)
