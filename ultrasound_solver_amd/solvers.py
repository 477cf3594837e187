"""Drop-in solver backend: the reference's duck-typed Helmholtz solver with its
``gpu_iterative`` path running on the HIP library.

* ``TransducerArray``  -- src/algorithm/physics/transducer_array.py:5-77
  (ideal_grid centres, unit amplitudes, geometric phases, set_phases).
* ``GpuIterativeSolver`` -- src/algorithm/solvers/iterative_solver.py:12-104:
  same constructor (restart, maxiter, rtol, backward_error_tol), same
  ``last_*`` attributes, same RuntimeError text on failure; ``solve`` takes
  the device operator in place of a cupyx CSR matrix, and also accepts k
  right-hand sides at once.
* ``HelmholtzSolver`` (alias ``HelmholtzDirectSolver``) --
  src/algorithm/solvers/helmholtz_solver.py:29-740: ``assemble_matrix``,
  ``solve``, ``solve_rhs``, ``solve_adjoint``, ``material_vjp``,
  ``set_material_fields``, ``scene_fingerprint``, ``target_response``
  (exact homogeneous surface condensation with the reference's response
  cache file), ``field_at``, ``field_at_direct``,
  ``build_phase_response_basis``, ``close``; Fortran-order reshapes.

The backend string keeps the reference's ``solver.backend`` key; only
``gpu_iterative`` exists here (the cuDSS direct backends are NVIDIA-only).
Extra optional ``solver:`` keys: ``gmres_precision`` ("fp64" plain form,
"mixed" complex64 + FP64 refinement), ``gmres_preconditioner`` ("jacobi" or
"multigrid", mixed form only), ``gmres_chunk`` (right-hand sides per batch, or
"auto" for the autotuner's choice),
``gmres_max_refinements``, ``gmres_inner_rtol``, ``gmres_mixed_restart`` (restart
length of the mixed form's inner GMRES, 10 by default; ``gmres_restart`` is the
plain form's, 30 as in the reference), ``gmres_relative_rtol`` (> 0: a column
also needs FP64 ||b - Ax|| / ||b|| at most this to count as converged; 0, the
default, keeps the reference's acceptance test alone).
"""
import hashlib
import json
import os
import time
import warnings
from dataclasses import asdict
from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp

from . import physics

BACKENDS = ("gpu_iterative", "cudss_hybrid_direct", "gpu_direct")


class TransducerArray:
    def __init__(self, config):
        self.cfg = config
        self.centers = self._centers()
        self.num_transducers = len(self.centers)
        self.surface_p0 = 1.0
        self.amplitudes = np.ones(self.num_transducers, dtype=np.float64)
        self.phases = np.zeros(self.num_transducers, dtype=np.float64)
        if self.cfg.mode in ("same_phase", "sdf_inverse"):
            self.phases.fill(self.cfg.same_phase_rad)

    def _centers(self):
        spec = self.cfg.specs
        return physics.ideal_grid_centers(spec.array_n, spec.pitch, self.cfg.lx, self.cfg.ly,
                                          self.cfg.domain.origin_m)

    def compute_geometric_phases(self):
        return physics.geometric_phases(self.centers, self.cfg.targets, self.cfg.wavelength, self.cfg.k0)

    def set_phases(self, phases):
        phases = np.asarray(phases, dtype=np.float64)
        if phases.shape != (self.num_transducers,) or not np.isfinite(phases).all():
            raise ValueError("phases must be a finite vector matching the array")
        self.phases = phases.copy()


class GpuIterativeSolver:
    """Row-scaled GMRES on the device (Jacobi by default, as in the reference)."""

    def __init__(self, restart=30, maxiter=3000, rtol=1.0e-6, backward_error_tol=1.0e-4,
                 precision="fp64", chunk=8, preconditioner="jacobi", max_refinements=10, inner_rtol=1.0e-5,
                 verbose=True, mixed_restart=10, relative_rtol=0.0):
        self.restart = int(restart)
        # The mixed form's inner complex64 GMRES restarts every `mixed_restart`
        # steps: with the multigrid preconditioner the step count barely
        # depends on it, while the Gram-Schmidt work per step grows with it.
        self.mixed_restart = int(mixed_restart)
        if self.restart < 1 or self.mixed_restart < 1:
            raise ValueError("GMRES restart lengths must be >= 1")
        self.maxiter = int(maxiter)
        self.rtol = float(rtol)
        self.backward_error_tol = float(backward_error_tol)
        if precision not in ("fp64", "mixed"):
            raise ValueError("gmres_precision must be 'fp64' or 'mixed'")
        if preconditioner not in ("jacobi", "multigrid"):
            raise ValueError("gmres_preconditioner must be 'jacobi' or 'multigrid'")
        if preconditioner == "multigrid" and precision != "mixed":
            raise ValueError("the multigrid preconditioner runs in the mixed form (gmres_precision: mixed)")
        self.precision = precision
        # "auto" (or 0): the first mixed-form solve on an operator runs the
        # autotuner (tile shapes and chunk width) and keeps its choice.
        self.chunk = 0 if chunk in ("auto", 0, None) else int(chunk)
        self.preconditioner = preconditioner
        self.max_refinements = int(max_refinements)
        self.inner_rtol = float(inner_rtol)
        self.relative_rtol = float(relative_rtol)
        if not (self.relative_rtol >= 0.0):
            raise ValueError("gmres_relative_rtol must be >= 0")
        self.verbose = verbose
        self.last_iterations = 0
        self.last_scaled_residual = None
        self.last_backward_error = None
        self.last_stats = []

    def solve(self, operator, rhs, adjoint=False):
        """Solve A x = rhs (A^H x = rhs when adjoint) for each column; rhs is
        (N,) / (N, k) host data or a (k, N) device tensor.  Returns a device
        tensor of the same column count ((N,) stays one row)."""
        b = operator.as_columns(rhs)
        chunk = 1
        restart = self.restart
        if self.precision == "mixed":
            restart = self.mixed_restart
            chunk = self.chunk
            if not chunk:
                if operator.tuning is None:
                    operator.autotune(max(b.shape[0], 64), restart=restart, preconditioner=self.preconditioner)
                chunk = operator.tuning["chunk"]
        x, stats = operator.gmres(
            b, restart=restart, maxiter=self.maxiter, rtol=self.rtol,
            backward_error_tol=self.backward_error_tol,
            precision="c128" if self.precision == "fp64" else "c64", adjoint=adjoint,
            chunk=chunk, max_refinements=self.max_refinements,
            inner_rtol=self.inner_rtol, preconditioner=self.preconditioner, relative_rtol=self.relative_rtol)
        self.last_stats = stats
        self.last_iterations = max(s["iterations"] for s in stats)
        self.last_scaled_residual = max(s["scaled_residual"] for s in stats)
        self.last_backward_error = max(s["backward_error"] for s in stats)
        bad = [s for s in stats if not s["converged"]]
        if bad:
            s = bad[0]
            raise RuntimeError(f"info={s['info']}, iterations={s['iterations']}, "
                               f"scaled residual={s['scaled_residual']:.3e}, "
                               f"backward error={s['backward_error']:.3e}")
        if self.verbose:
            print(f"  --> GMRES iterations: {self.last_iterations} | scaled residual: "
                  f"{self.last_scaled_residual:.3e} | backward error: {self.last_backward_error:.3e}")
        return x


class HelmholtzSolver:
    """Assemble (per-node coefficients on the device) and solve the discrete Helmholtz system."""

    MODEL_VERSION = "helmholtz_conservative_sources_v5"
    field_unit = "Pa"

    def __init__(self, config, device="cuda"):
        self.cfg = config
        self.transducers = TransducerArray(config)
        self.field_unit = "normalized_1_per_m"
        self.origin = np.asarray(config.domain.origin_m, dtype=np.float64)
        if not np.isclose(self.origin[2], 0.0):
            raise ValueError("The transducer plane requires domain.origin_m[2] == 0")
        d = config.domain
        self.shape = (d.nx, d.ny, d.nz)
        self.total_dofs = d.nx*d.ny*d.nz
        self.device = device
        self.operator = None
        self.rhs_cached = None
        self.condensed_solver = None
        self._scene_key = None
        self._material_fields = None
        self._material_digest = None
        self._obs_sdf = None
        self._grid_definition = asdict(config.domain)
        self._source_definition = asdict(config.specs)
        self.source_columns = None
        self.warp_engine = SimpleNamespace(source_columns=None, operator=None)
        self._target_surface_solver = None
        self._target_source_masses = None
        s = config.solver
        self.iterative_solver = GpuIterativeSolver(
            restart=int(s.get("gmres_restart", 30)), maxiter=int(s.get("gmres_maxiter", 3000)),
            rtol=float(s.get("gmres_rtol", 1.0e-6)),
            backward_error_tol=float(s.get("gmres_backward_error_tol", 1.0e-4)),
            precision=str(s.get("gmres_precision", "fp64")), chunk=s.get("gmres_chunk", 8),
            preconditioner=str(s.get("gmres_preconditioner", "jacobi")),
            max_refinements=int(s.get("gmres_max_refinements", 10)),
            inner_rtol=float(s.get("gmres_inner_rtol", 1.0e-5)),
            mixed_restart=int(s.get("gmres_mixed_restart", 10)),
            relative_rtol=float(s.get("gmres_relative_rtol", 0.0)))

    # -- scene ------------------------------------------------------------
    @property
    def obstacle_sdf(self):
        return self._obs_sdf

    def set_material_fields(self, density, sound_speed):
        """Per-node density and sound speed (helmholtz_solver.py:238-247), copied;
        their digest is taken here, once, for the scene fingerprint."""
        fields = []
        for a in (density, sound_speed):
            a = np.array(a, dtype=np.float64, copy=True)
            if a.shape != self.shape or not np.all(np.isfinite(a) & (a > 0)):
                raise ValueError("Material fields must match the grid and be finite and positive")
            fields.append(a)
        if self.cfg.obstacles or self.cfg.physics.material_field_file:
            raise ValueError("Choose one material input: arrays, NPZ file, or obstacle")
        self._material_fields = tuple(fields)
        self._material_digest = _digest_bytes(*(a.tobytes() for a in fields))
        self.invalidate_factorization()

    def scene_fingerprint(self):
        """Key of everything that defines A and B (helmholtz_solver.py:249-271):
        one digest per part -- the configuration blocks, the source layout,
        the content of a material file and the material arrays -- combined.
        A file's digest is reused while its size, inode and modification time
        stay the same, so repeated solves do not re-read it."""
        if asdict(self.cfg.domain) != self._grid_definition or asdict(self.cfg.specs) != self._source_definition:
            raise ValueError("Grid or source specification changed; construct a new solver")
        cfg = self.cfg
        parts = [self.MODEL_VERSION,
                 _digest_json([asdict(cfg.physics), asdict(cfg.domain), asdict(cfg.specs)]),
                 _digest_json([cfg.boundary_conditions, cfg.obstacles, cfg.solver.get("pml", {})]),
                 _digest_json([self.origin.tolist(), (self.transducers.centers - self.origin).tolist()])]
        if cfg.physics.material_field_file:
            parts.append(_file_digest(cfg.physics.material_field_file))
        if self._material_digest is not None:
            parts.append(self._material_digest)
        return _digest_bytes(*(p.encode() for p in parts))

    def source_diagnostics(self):
        centers = (self.transducers.centers - self.origin)[:, :2]
        radius = self.cfg.specs.diameter*0.5
        clipped = np.any((centers - radius < 0) | (centers + radius > np.asarray([self.cfg.lx, self.cfg.ly])), axis=1)
        return {"points_per_wavelength": self.cfg.wavelength/self.cfg.domain.dx,
                "kdx": self.cfg.k0*self.cfg.domain.dx, "clipped_emitters": np.flatnonzero(clipped).tolist()}

    def _materials(self):
        shape, dx = self.shape, self.cfg.domain.dx
        rho0, c0 = self.cfg.physics.medium_density, self.cfg.physics.sound_speed
        if len(self.cfg.obstacles) > 1:
            raise NotImplementedError("The solver currently supports one obstacle at a time")
        obstacle = dict(self.cfg.obstacles[0]) if self.cfg.obstacles else {}
        fields = self._material_fields
        if self.cfg.physics.material_field_file:
            if fields is not None or obstacle:
                raise ValueError("Material NPZ is mutually exclusive with arrays or obstacles")
            with np.load(self.cfg.physics.material_field_file, allow_pickle=False) as data:
                fields = (np.asarray(data["density"], float), np.asarray(data["sound_speed"], float))
        if fields is not None:
            if obstacle:
                raise ValueError("Explicit material fields and obstacle materials are mutually exclusive")
            rho, c = fields
            if rho.shape != shape or c.shape != shape:
                raise ValueError(f"Material fields must match grid {shape}")
            return rho, c
        if not obstacle:
            return np.full(shape, float(rho0)), np.full(shape, float(c0))
        kind = obstacle.get("type", "sphere").lower()
        if kind != "sphere":
            raise ValueError(f"Unsupported obstacle type for this backend: {kind!r} (use material arrays)")
        center = np.asarray(obstacle.get("center", [.04, .04, .02]), float) - self.origin
        width = float(obstacle.get("smoothing_width_cells", 1.2))
        if not np.isfinite(width) or width < 0:
            raise ValueError("smoothing_width_cells must be finite and >= 0")
        mat = obstacle.get("material", {})
        return physics.rasterize_materials(shape, dx, [("sphere", center, float(obstacle.get("radius", .012)))],
                                           rho0, c0, float(mat.get("density", 1250.0)),
                                           float(mat.get("sound_speed", 2200.0)), width)

    def assemble_matrix(self, force=False):
        """Build the device operator and source columns; returns the RHS of the
        current phases (helmholtz_solver.py:273-369)."""
        key = self.scene_fingerprint()
        if self.operator is not None and not force and key == self._scene_key:
            return self.update_rhs()
        self.invalidate_factorization()
        diag = self.source_diagnostics()
        if diag["points_per_wavelength"] < 6:
            raise ValueError("PDE grid has fewer than 6 points per background wavelength; "
                             "refine grid_size before solving")
        if diag["points_per_wavelength"] < 10:
            warnings.warn(f"PDE resolution: {diag['points_per_wavelength']:.2f} points/wavelength; "
                          "small algebraic residual does not establish phase accuracy.", RuntimeWarning)
        if diag["clipped_emitters"]:
            warnings.warn(f"Source discs clipped by the domain: {diag['clipped_emitters']}", RuntimeWarning)
        rho, c = self._materials()
        bcs = self.cfg.boundary_conditions
        if physics.boundary_type(bcs.get("-z", "reflecting")) != "reflecting":
            raise ValueError("The sinc Neumann source requires a reflecting baffled bottom plane")
        from .operator import HelmholtzOperator
        op = HelmholtzOperator(rho, c, self.cfg.domain.dx, self.cfg.omega, bcs, self.cfg.solver.get("pml"),
                               device=self.device)
        centers = self.transducers.centers - self.origin
        node, col, mass = physics.source_deposition(self.shape[:2], self.cfg.domain.dx, centers,
                                                    self.cfg.specs.diameter, self.cfg.specs.source_radial_order,
                                                    self.cfg.specs.source_angular_order)
        value = physics.flux_scale(self.cfg.domain.dx, self.cfg.physics.medium_density)*mass
        if np.any(op.pml_mask[:, :, 0].ravel(order="F")[node]):
            raise ValueError("PML overlaps the deposited sinc aperture")
        if np.any(op.fixed[:, :, 0].ravel(order="F")[node]):
            raise ValueError("Sinc flux conflicts with a pressure-release boundary")
        if c.min()/self.cfg.frequency/self.cfg.domain.dx < 6:
            raise ValueError("PDE grid has fewer than 6 points per shortest local wavelength; "
                             "refine grid_size before solving (10+ points and convergence checks recommended)")
        op.set_sources(node, col, value, self.transducers.num_transducers)
        if self.iterative_solver.preconditioner == "multigrid":
            op.set_multigrid()
        self.operator = op
        self.source_columns = sp.csr_matrix((value, (node, col)),
                                            shape=(self.total_dofs, self.transducers.num_transducers))
        self.warp_engine = SimpleNamespace(source_columns=self.source_columns, operator=op,
                                           fixed_dofs=lambda: np.flatnonzero(op.fixed.ravel(order="F")))
        self._scene_key = key
        return self.update_rhs()

    def update_rhs(self):
        t = self.transducers
        self.rhs_cached = self.source_columns @ (t.amplitudes*np.exp(1j*t.phases))
        return self.rhs_cached

    def set_phases(self, phases):
        self.transducers.set_phases(phases)

    def invalidate_factorization(self):
        if self.operator is not None:
            self.operator.close()
        self.operator = None
        self.rhs_cached = None

    def factorize(self):
        self.assemble_matrix()

    def close(self):
        self._target_surface_solver = None
        self._target_source_masses = None
        self.invalidate_factorization()

    # -- solves -----------------------------------------------------------
    def _requested_backend(self):
        backend = str(self.cfg.solver.get("backend", "gpu_iterative")).lower()
        if backend not in BACKENDS:
            raise ValueError(f"unknown solver.backend={backend!r}; valid values are {sorted(BACKENDS)}")
        if backend != "gpu_iterative":
            raise NotImplementedError(f"solver.backend={backend!r} is an NVIDIA cuDSS backend; "
                                      "this MI355X build provides gpu_iterative")
        return backend

    def _rhs_matrix(self, rhs):
        a = np.asarray(rhs, dtype=np.complex128)
        if a.ndim == 1 and a.size == self.total_dofs:
            return a.reshape(-1, 1), True
        if a.ndim == 2 and a.shape[0] == self.total_dofs:
            return a, False
        raise ValueError(f"rhs must have shape ({self.total_dofs},) or ({self.total_dofs}, nrhs), got {a.shape}")

    def _adjoint_rhs_matrix(self, rhs):
        a = np.asarray(rhs, dtype=np.complex128)
        if a.shape == self.shape:
            return a.reshape((-1, 1), order="F"), True
        if a.ndim == 4 and a.shape[:3] == self.shape:
            return a.reshape((self.total_dofs, -1), order="F"), False
        return self._rhs_matrix(a)

    def _reshape_solution(self, solution):
        a = np.asarray(solution)
        if a.ndim == 1:
            return a.reshape(self.shape, order="F")
        return a.reshape(self.shape + (a.shape[1],), order="F")

    def solve(self):
        start = time.perf_counter()
        rhs = self.assemble_matrix(force=not bool(self.cfg.solver.get("reuse_matrix", True)))
        print(f"[OK] stage 1/2: operator and RHS ready in {(time.perf_counter() - start)*1e3:.2f} ms "
              f"(DOFs: {self.total_dofs:,})")
        start = time.perf_counter()
        solution = self.solve_rhs(rhs)
        print(f"[OK] stage 2/2: linear solve done in {time.perf_counter() - start:.2f} s")
        return self._reshape_solution(solution)

    def solve_rhs(self, rhs, verbose=True):
        """(N,) or (N, k) right-hand sides -> host complex128 solution(s)."""
        if self.operator is None:
            raise RuntimeError("assemble_matrix() must be called first")
        self.assemble_matrix()
        mat, squeeze = self._rhs_matrix(rhs)
        self._requested_backend()
        self.iterative_solver.verbose = verbose
        x = self.iterative_solver.solve(self.operator, mat).cpu().numpy().T
        return x[:, 0] if squeeze else x

    def solve_adjoint(self, adjoint_rhs):
        self.assemble_matrix()
        mat, squeeze = self._adjoint_rhs_matrix(adjoint_rhs)
        self._requested_backend()
        x = self.iterative_solver.solve(self.operator, mat, adjoint=True).cpu().numpy().T
        return self._reshape_solution(x[:, 0] if squeeze else x)

    def material_vjp(self, field, field_cotangent):
        """dL/drho, dL/dc (grid arrays) with one adjoint solve and the device
        material-gradient kernel (conservative_helmholtz.py:151-175)."""
        lam = self.solve_adjoint(field_cotangent)
        u = np.asarray(field).reshape(self.shape, order="F").ravel(order="F")
        g_rho, g_c = self.operator.material_vjp(u, np.asarray(lam).reshape(self.shape, order="F").ravel(order="F"))
        return g_rho.cpu().numpy(), g_c.cpu().numpy()

    # -- target responses ---------------------------------------------------
    def _homogeneous(self):
        return not (self.cfg.obstacles or self.cfg.physics.material_field_file or self._material_fields is not None)

    def _ensure_target_surface_solver(self):
        """Exact homogeneous surface condensation (helmholtz_solver.py:78-103)."""
        if not self._homogeneous():
            raise ValueError("Surface condensation requires a homogeneous medium")
        if physics.boundary_type(self.cfg.boundary_conditions.get("-z", "reflecting")) != "reflecting":
            raise ValueError("Surface condensation requires a reflecting baffled -z face")
        if self._target_surface_solver is not None:
            return
        from .surface import HomogeneousSurfaceSolver
        d, sp_ = self.cfg.domain, self.cfg.specs
        node, col, mass = physics.source_deposition(self.shape[:2], d.dx, self.transducers.centers - self.origin,
                                                    sp_.diameter, sp_.source_radial_order, sp_.source_angular_order)
        self._target_source_masses = sp.csr_matrix((mass, (node, col)),
                                                   shape=(self.shape[0]*self.shape[1], self.transducers.num_transducers))
        self._target_surface_solver = HomogeneousSurfaceSolver(
            self.shape, d.dx, self.cfg.frequency, self.cfg.physics.sound_speed, self.cfg.physics.medium_density,
            self.cfg.boundary_conditions, pml=self.cfg.solver.get("pml", {}), device=self.device)

    def _local_points(self, points):
        """Physical target points -> coordinates relative to the domain origin;
        a point outside the domain, or inside a PML layer (1e-12 m of slack),
        is rejected (helmholtz_solver.py:105-123)."""
        points = np.asarray(points, dtype=float)
        box = np.asarray(self.cfg.domain.box_size, dtype=float)
        local = points - self.origin
        if not (points.ndim == 2 and points.shape[1] == 3 and np.isfinite(points).all()
                and np.all(local >= -1e-12) and np.all(local <= box + 1e-12)):
            raise ValueError("Targets must be finite physical coordinates inside the PDE domain")
        lo, hi = physics.interior_bounds(box, self.cfg.domain.dx, self.cfg.boundary_conditions,
                                         self.cfg.solver.get("pml", {}))
        if np.any(local < lo - 1e-12) or np.any(local > hi + 1e-12):
            raise ValueError("Target lies in PML; enlarge the physical interior")
        return local

    def _response_cache(self):
        """The target_response_cache file for this scene.  Its key covers the
        scene fingerprint, the origin and the code that computes responses
        (this module, the surface solver, the host tables and the device
        surface kernels), so an edit to any of them starts a fresh cache."""
        from .response_cache import ResponseCache
        key = _digest_bytes(self.MODEL_VERSION.encode(), self.scene_fingerprint().encode(),
                            np.asarray(self.origin, np.float64).tobytes(), _implementation_digest().encode())
        name = str(self.cfg.solver.get("target_response_cache", "")).strip()
        return ResponseCache(name or None, key, self.transducers.num_transducers)

    def _adjoint_target_response(self, local):
        """Rows of G at local points by reciprocity, one batched adjoint solve
        (conj(lambda) = A^{-H} conj(t), response = a * B^T lambda).  Serves
        heterogeneous scenes, where the reference raises."""
        from .training import trilinear_rows
        self.assemble_matrix()
        rows, weights = trilinear_rows(local + self.origin, self.shape, self.origin, self.cfg.domain.dx,
                                       np.asarray(self.cfg.domain.box_size, float))
        t = np.zeros((self.total_dofs, len(local)), dtype=np.complex128)
        for q in range(len(local)):
            np.add.at(t[:, q], rows[q], weights[q])
        lam = np.conj(self.iterative_solver.solve(self.operator, t, adjoint=True).cpu().numpy().T)
        return (self.source_columns.T @ lam).T*self.transducers.amplitudes[None, :]

    def _compute_responses(self, local):
        if self._homogeneous():
            self._ensure_target_surface_solver()
            return self._target_surface_solver.flux_target_response(self._target_source_masses, local)
        return self._adjoint_target_response(local)

    def target_response(self, points):
        """(P, M) responses of the discrete PDE at physical points
        (helmholtz_solver.py:179-208).  Points found in the response cache
        file are served from it; the others are computed together -- by the
        exact homogeneous surface condensation on the device, or, for a
        heterogeneous scene (where the reference raises), by one batched
        adjoint solve -- and appended to the file."""
        points = np.asarray(points, dtype=float)
        local = self._local_points(points)
        cache = self._response_cache()
        hit = cache.lookup(points)
        out = np.empty((len(points), self.transducers.num_transducers), dtype=np.complex128)
        out[hit >= 0] = cache.rows[hit[hit >= 0]]
        todo = np.flatnonzero(hit < 0)
        if todo.size:
            fresh = self._compute_responses(local[todo])
            out[todo] = fresh
            cache.store(points[todo], fresh)
        return out

    def _phase_vector(self, phases):
        values = self.transducers.phases if phases is None else np.asarray(phases, dtype=float)
        if values.shape != (self.transducers.num_transducers,) or not np.isfinite(values).all():
            raise ValueError("phases must match the transducer array and be finite")
        return values

    def field_at(self, points, phases=None):
        """Field at physical points for the given (or current) phases, from the
        target responses (helmholtz_solver.py:210-214)."""
        return self.target_response(points) @ np.exp(1j*self._phase_vector(phases))

    def field_at_direct(self, points, phases=None):
        """Field at physical points from one surface solve with these phases
        and trilinear samples of its planes (helmholtz_solver.py:216-222)."""
        values = self._phase_vector(phases)
        local = self._local_points(points)
        self._ensure_target_surface_solver()
        self._target_surface_solver.solve_flux(self._target_source_masses, values)
        return self._target_surface_solver.sample(local)

    def build_phase_response_basis(self):
        from .training import PhaseResponseBasis
        tr = self.cfg.training
        return PhaseResponseBasis(self, tr.phase_basis_file, tr.phase_basis_batch_size, tr.voxel_chunk_size,
                                  tr.release_factor_after_basis).build()


HelmholtzDirectSolver = HelmholtzSolver


def _digest_bytes(*chunks):
    h = hashlib.sha256()
    for c in chunks:
        h.update(len(c).to_bytes(8, "little"))
        h.update(c)
    return h.hexdigest()


def _digest_json(obj):
    return _digest_bytes(json.dumps(obj, sort_keys=True, separators=(",", ":")).encode())


_FILE_DIGESTS = {}


def _file_digest(name):
    """sha256 of a file's content, memoised by (path, inode, size, mtime)."""
    st = os.stat(name)
    key = (os.path.abspath(name), st.st_ino, st.st_size, st.st_mtime_ns)
    if key not in _FILE_DIGESTS:
        h = hashlib.sha256()
        with open(name, "rb") as fh:
            for block in iter(lambda: fh.read(1 << 20), b""):
                h.update(block)
        _FILE_DIGESTS[key] = h.hexdigest()
    return _FILE_DIGESTS[key]


def _implementation_digest():
    """Digest of the sources that compute target responses."""
    here = os.path.dirname(os.path.abspath(__file__))
    names = ("solvers.py", "surface.py", "physics.py", "response_cache.py", os.path.join("csrc", "hz_surface.hpp"))
    return _digest_bytes(*(_file_digest(os.path.join(here, n)).encode() for n in names))
