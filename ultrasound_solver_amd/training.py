"""Phase response basis, losses and the phase objective on the device.

* ``PhaseResponseBasis`` -- src/algorithm/training/phase_response_basis.py:16-416:
  G = A^{-1} B diag(a) built in batches (here by the batched device GMRES,
  which the reference's gpu_iterative backend could not do), stored as the
  same resumable ``.npy`` memmap (N, M) complex128 + ``.json`` metadata
  (format 3), then ``field`` (G e^{i phi}), ``phase_vjp``
  (Re(conj(G^H c) i e^{i phi})) and trilinear ``target_response``.
* Losses -- src/algorithm/training/phase_adjoint_optimizer.py:78-414:
  ``AmplitudeTarget``/Gaussian targets, ``AmplitudeFieldLoss``
  (field_match), ``FocalPressureLoss``, ``FocalContrastLoss`` and the
  background mask, plus ``FocalContrastRatioLoss`` (the ratio form of focal
  contrast), each with a host (numpy) and a device (torch) evaluation
  returning (loss, cotangent) with dL = Re(c^H du), and a one-pass
  evaluation on a device basis returning (loss, G^H c).
* ``PhaseObjective`` -- ``PhaseProblem.evaluate`` (src/algorithm/baselines/
  common.py:172-343) on the device.
"""
import hashlib
import json
from dataclasses import asdict
from pathlib import Path

import numpy as np
import torch

EPS = 1.0e-12


def trilinear_rows(points, shape, origin, dx, box=None):
    """(rows (P, 8) Fortran node indices, weights (P, 8)) of the trilinear stencil;
    the lower corner is clipped to n-2 (phase_response_basis.py:165-174)."""
    points = np.asarray(points, dtype=float)
    shape = np.asarray(shape)
    origin = np.asarray(origin, dtype=float)
    if points.ndim != 2 or points.shape[1] != 3 or not np.isfinite(points).all():
        raise ValueError("Targets must be finite (N,3) coordinates inside the grid")
    upper = origin + (shape - 1)*dx if box is None else origin + np.asarray(box, float)
    if np.any(points < origin) or np.any(points > upper + 1e-12):
        raise ValueError("Targets must be finite (N,3) coordinates inside the grid")
    pos = (points - origin)/dx
    low = np.minimum(np.floor(pos).astype(np.int64), shape - 2)
    frac = pos - low
    rows = np.empty((len(points), 8), dtype=np.int64)
    wts = np.empty((len(points), 8))
    for c, corner in enumerate(np.ndindex(2, 2, 2)):
        corner = np.asarray(corner)
        idx = low + corner
        wts[:, c] = np.prod(np.where(corner, frac, 1.0 - frac), axis=1)
        rows[:, c] = idx[:, 0] + shape[0]*(idx[:, 1] + shape[1]*idx[:, 2])
    return rows, wts


class PhaseResponseBasis:
    FORMAT_VERSION = 3

    def __init__(self, solver, storage_path, rhs_batch_size=8, voxel_chunk_size=262144,
                 release_factor_after_build=True):
        self.solver = solver
        self.storage_path = Path(storage_path)
        self.metadata_path = Path(f"{self.storage_path}.json")
        self.rhs_batch_size = int(rhs_batch_size)
        self.voxel_chunk_size = int(voxel_chunk_size)
        self.release_factor_after_build = bool(release_factor_after_build)
        self.dtype = np.dtype(np.complex128)
        self.basis = None
        self.basis_gpu = None
        self.completed_columns = 0
        self._built_fingerprint = None
        if self.rhs_batch_size < 1:
            raise ValueError("rhs_batch_size must be at least 1")
        if self.voxel_chunk_size < 1:
            raise ValueError("voxel_chunk_size must be at least 1")

    @property
    def num_transducers(self):
        return int(self.solver.transducers.num_transducers)

    @property
    def shape(self):
        return self.solver.total_dofs, self.num_transducers

    @property
    def is_complete(self):
        return self.completed_columns == self.num_transducers

    def build(self):
        backend = str(self.solver.cfg.solver.get("backend", "gpu_iterative")).lower()
        if backend != "gpu_iterative":
            raise ValueError("this backend builds the phase basis with solver.backend='gpu_iterative'")
        self._open_or_create_storage()
        if self.is_complete:
            return self
        self.solver.assemble_matrix()
        counts = np.asarray(abs(self.solver.source_columns).sum(axis=0)).ravel()
        missing = np.flatnonzero(counts == 0)
        if missing.size:
            raise ValueError(f"The grid does not cover transducers {missing.tolist()}")
        op = self.solver.operator
        amps = self.solver.transducers.amplitudes
        it = self.solver.iterative_solver
        for start in range(self.completed_columns, self.num_transducers, self.rhs_batch_size):
            stop = min(start + self.rhs_batch_size, self.num_transducers)
            w = torch.zeros((stop - start, self.num_transducers), dtype=torch.complex128)
            w[torch.arange(stop - start), torch.arange(start, stop)] = torch.as_tensor(amps[start:stop]).to(torch.complex128)
            b = op.build_rhs(w)
            x = it.solve(op, b)
            self.basis[:, start:stop] = x.cpu().numpy().T
            self.basis.flush()
            self.completed_columns = stop
            self._write_metadata()
        return self

    def to_gpu(self, device="cuda"):
        self._require_complete()
        free, _ = torch.cuda.mem_get_info(torch.device(device))
        need = int(self.basis.nbytes)
        if need + 256*1024**2 > free:
            raise MemoryError("Insufficient GPU memory for phase response basis: "
                              f"required={need/1024**3:.2f} GiB, free={free/1024**3:.2f} GiB")
        self.basis_gpu = torch.as_tensor(np.asarray(self.basis), device=device)
        return self

    def release_gpu(self):
        self.basis_gpu = None

    def target_response(self, points):
        """(P, M) trilinear rows of G at the points: the 8 corner rows of each
        point are read (from the device basis, or from the stored file) and
        weighted on the device (phase_response_basis.py:139-175)."""
        self._require_complete()
        cfg = self.solver.cfg
        rows, wts = trilinear_rows(points, cfg.domain.grid_size, cfg.domain.origin_m, cfg.domain.dx,
                                   cfg.domain.box_size)
        self._check_outside_pml(np.asarray(points, float))
        dev = self._device()
        if self.basis_gpu is not None:
            corner = self.basis_gpu.index_select(0, torch.as_tensor(rows.ravel(), device=dev))
        else:
            corner = torch.as_tensor(np.asarray(self.basis[rows.ravel()]), device=dev)
        corner = corner.reshape(len(rows), 8, self.num_transducers)
        out = (torch.as_tensor(wts, device=dev)[:, :, None]*corner).sum(dim=1)
        return out.cpu().numpy()

    def _device(self):
        if self.basis_gpu is not None:
            return self.basis_gpu.device
        op = getattr(self.solver, "operator", None)
        return torch.device(op.device if op is not None else "cuda")

    def _row_chunks(self):
        """(start, device rows) of a basis that stays in its file: chunks of
        voxel_chunk_size rows are uploaded in turn (phase_response_basis.py:373-410)."""
        dev = self._device()
        for s in range(0, self.solver.total_dofs, self.voxel_chunk_size):
            yield s, torch.as_tensor(np.asarray(self.basis[s:s + self.voxel_chunk_size]), device=dev)

    def _check_outside_pml(self, points):
        from .physics import interior_bounds
        cfg = self.solver.cfg
        lo, hi = interior_bounds(cfg.domain.box_size, cfg.domain.dx, cfg.boundary_conditions,
                                 cfg.solver.get("pml", {}))
        local = points - np.asarray(cfg.domain.origin_m, float)
        if np.any(local < lo) or np.any(local > hi):
            raise ValueError("Target lies in PML; enlarge the physical interior")

    def field(self, phases, return_numpy=True):
        """u = G e^{i phi} on the device (phase_response_basis.py:177-206);
        a basis not moved to the device is streamed to it in row chunks."""
        self._require_complete()
        phases = np.asarray(phases) if not isinstance(phases, torch.Tensor) else phases
        if tuple(phases.shape) != (self.num_transducers,):
            raise ValueError(f"phases must have shape ({self.num_transducers},), got {tuple(phases.shape)}")
        from . import basis as dev_basis
        d = self.solver.cfg.domain
        dev = self._device()
        w = torch.exp(1j*torch.as_tensor(phases, dtype=torch.float64, device=dev))
        if self.basis_gpu is not None:
            u = dev_basis.field(self.basis_gpu, w)
        else:
            u = torch.empty(self.solver.total_dofs, dtype=torch.complex128, device=dev)
            for s, rows in self._row_chunks():
                u[s:s + rows.shape[0]] = dev_basis.field(rows, w)
        if return_numpy:
            return u.cpu().numpy().reshape((d.nx, d.ny, d.nz), order="F")
        # (nx, ny, nz) view of the Fortran-ordered vector
        return u.reshape(d.nz, d.ny, d.nx).permute(2, 1, 0)

    def phase_vjp(self, field_cotangent, phases, return_numpy=True):
        """dL/dphi = Re(conj(G^H c) i e^{i phi}) for dL = Re(c^H du), on the
        device (phase_response_basis.py:208-253)."""
        self._require_complete()
        from . import basis as dev_basis
        dev = self._device()
        c = field_cotangent
        if isinstance(c, torch.Tensor):
            c = c.to(dev, torch.complex128)
            c = c.permute(*reversed(range(c.dim()))).reshape(-1) if c.dim() == 3 else c.reshape(-1)
        else:
            c = torch.as_tensor(np.asarray(field_cotangent, np.complex128).reshape(-1, order="F"), device=dev)
        if c.numel() != self.solver.total_dofs:
            raise ValueError("field_cotangent has the wrong size")
        if self.basis_gpu is not None:
            resp = dev_basis.adjoint(self.basis_gpu, c)
        else:
            resp = torch.zeros(self.num_transducers, dtype=torch.complex128, device=dev)
            for s, rows in self._row_chunks():
                resp += dev_basis.adjoint(rows, c[s:s + rows.shape[0]].contiguous())
        w = torch.exp(1j*torch.as_tensor(np.asarray(phases, np.float64), device=dev))
        g = dev_basis.phase_gradient(resp, w)
        return g.cpu().numpy() if return_numpy else g

    def close(self):
        self.release_gpu()
        if isinstance(self.basis, np.memmap):
            self.basis.flush()
            mm = self.basis._mmap
            self.basis = None
            mm.close()

    # -- storage (format 3, resumable) ---------------------------------------
    def _open_or_create_storage(self):
        have_basis, have_meta = self.storage_path.exists(), self.metadata_path.exists()
        if have_basis != have_meta:
            raise RuntimeError("Phase basis data and metadata must exist together")
        if have_basis:
            meta = json.loads(self.metadata_path.read_text(encoding="utf-8"))
            self._validate_metadata(meta)
            self.basis = np.load(self.storage_path, mmap_mode="r+")
            if self.basis.shape != self.shape:
                raise RuntimeError("Stored phase basis shape mismatch")
            if self.basis.dtype != self.dtype:
                raise RuntimeError("Stored phase basis dtype mismatch")
            self.completed_columns = int(meta["completed_columns"])
            self._built_fingerprint = meta["fingerprint"]
            return
        self.storage_path.parent.mkdir(parents=True, exist_ok=True)
        self.basis = np.lib.format.open_memmap(self.storage_path, mode="w+", dtype=self.dtype, shape=self.shape,
                                               fortran_order=False)
        self.completed_columns = 0
        self._write_metadata()

    def _validate_metadata(self, meta):
        if meta.get("format_version") != self.FORMAT_VERSION:
            raise RuntimeError("Unsupported phase basis format version")
        if meta.get("fingerprint") != self._fingerprint():
            raise RuntimeError("Phase basis does not match the current physical scene")
        if meta.get("shape") != list(self.shape):
            raise RuntimeError(f"Phase basis shape mismatch: {meta.get('shape')} != {list(self.shape)}")
        if meta.get("dtype") != self.dtype.name:
            raise RuntimeError("Phase basis dtype mismatch")
        done = int(meta.get("completed_columns", -1))
        if done < 0 or done > self.num_transducers:
            raise RuntimeError("Invalid completed column count")

    def _write_metadata(self):
        fp = self._fingerprint()
        meta = {"format_version": self.FORMAT_VERSION, "fingerprint": fp, "shape": list(self.shape),
                "dtype": self.dtype.name, "completed_columns": self.completed_columns}
        tmp = Path(f"{self.metadata_path}.tmp")
        tmp.write_text(json.dumps(meta, indent=2, sort_keys=True), encoding="utf-8")
        tmp.replace(self.metadata_path)
        self._built_fingerprint = fp

    def _fingerprint(self):
        cfg = self.solver.cfg
        payload = {"format_version": self.FORMAT_VERSION, "propagation_model": self.solver.MODEL_VERSION,
                   "centers_m": self.solver.transducers.centers.tolist(), "physics": asdict(cfg.physics),
                   "transducer_specs": asdict(cfg.specs), "domain": asdict(cfg.domain),
                   "boundary_conditions": cfg.boundary_conditions, "obstacles": cfg.obstacles,
                   "mesh_content_sha256": {}, "amplitudes": self.solver.transducers.amplitudes.tolist(),
                   "pde_scene": self.solver.scene_fingerprint()}
        return hashlib.sha256(json.dumps(payload, sort_keys=True, separators=(",", ":")).encode()).hexdigest()

    def _require_complete(self):
        if self.basis is None or not self.is_complete:
            raise RuntimeError("Phase response basis is not complete")
        if self._built_fingerprint is not None and self._fingerprint() != self._built_fingerprint:
            raise RuntimeError("Physical scene changed; rebuild the phase response basis")


# ------------------------------------------------------------------------ losses
def nearest_grid_index(cfg, point):
    return tuple(np.rint((np.asarray(point, float) - np.asarray(cfg.domain.origin_m))/cfg.domain.dx).astype(int))


def build_gaussian_target_fields(cfg, target_points, peak_pressure, sigma, background_pressure=0.0,
                                 target_weight=1.0, background_weight=0.05):
    pts = np.asarray(target_points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError("target_points must have shape (num_targets, 3)")
    if len(pts) == 0:
        raise ValueError("at least one target point is required")
    if sigma <= 0:
        raise ValueError("target sigma must be positive")
    if peak_pressure < background_pressure:
        raise ValueError("peak_pressure must be >= background_pressure")
    d = cfg.domain
    shape = (d.nx, d.ny, d.nz)
    o = np.asarray(d.origin_m, float)
    axes = [np.linspace(o[a], o[a] + d.box_size[a], shape[a]) for a in range(3)]
    lower, upper = o, o + np.asarray(d.box_size, float)
    target = np.full(shape, float(background_pressure))
    weight = np.full(shape, float(background_weight))
    for p in pts:
        if np.any(p < lower) or np.any(p > upper):
            raise ValueError(f"target point is outside the domain: {p.tolist()} not in "
                             f"[{lower.tolist()}, {upper.tolist()}]")
        r2 = ((axes[0][:, None, None] - p[0])**2 + (axes[1][None, :, None] - p[1])**2
              + (axes[2][None, None, :] - p[2])**2)
        g = np.exp(-0.5*r2/(sigma*sigma))
        g = g/g[nearest_grid_index(cfg, p)]
        target = np.maximum(target, background_pressure + (peak_pressure - background_pressure)*g)
        weight = np.maximum(weight, background_weight + (target_weight - background_weight)*g)
    return target, weight


def build_background_mask(cfg, target_points, target_radius, source_exclusion_layers):
    if target_radius <= 0:
        raise ValueError("training.target_radius must be positive")
    d = cfg.domain
    shape = (d.nx, d.ny, d.nz)
    axes = [np.linspace(0.0, d.box_size[a], shape[a]) for a in range(3)]
    mask = np.ones(shape, dtype=bool)
    layers = max(int(source_exclusion_layers), 0)
    if layers:
        mask[:, :, :layers] = False
    for p in np.asarray(target_points, float).reshape(-1, 3):
        r2 = ((axes[0][:, None, None] - p[0])**2 + (axes[1][None, :, None] - p[1])**2
              + (axes[2][None, None, :] - p[2])**2)
        mask[r2 <= target_radius*target_radius] = False
    return mask


class AmplitudeTarget:
    def __init__(self, target, weight):
        if target.shape != weight.shape:
            raise ValueError("target and weight must have the same shape")
        self.target = np.asarray(target, dtype=np.float64)
        self.weight = np.asarray(weight, dtype=np.float64)

    @classmethod
    def from_config(cls, cfg):
        tr = cfg.training
        if tr.target_field_file:
            path = Path(tr.target_field_file)
            if path.suffix.lower() == ".npy":
                target = np.load(path, allow_pickle=False)
                weight = np.ones_like(target, dtype=np.float64)
            elif path.suffix.lower() == ".npz":
                with np.load(path, allow_pickle=False) as data:
                    target = np.asarray(data["target"], np.float64)
                    weight = np.asarray(data["weight"], np.float64) if "weight" in data else np.ones_like(target)
            else:
                raise ValueError("target_field_file must be .npy or .npz")
            if target.shape != tuple(cfg.domain.grid_size):
                raise ValueError(f"target field shape must be {tuple(cfg.domain.grid_size)}, got {target.shape}")
            return cls(target, weight)
        if not cfg.targets:
            raise ValueError("phase_optimization requires targets or training.target_field_file")
        return cls(*build_gaussian_target_fields(cfg, np.asarray(cfg.targets, float), tr.target_peak_pressure,
                                                 tr.target_sigma, tr.background_pressure, tr.target_weight,
                                                 tr.background_weight))


def _as_device(a, like):
    return torch.as_tensor(a, device=like.device)


def _focal_term(field, indices, weight, with_cotangent, focal_field):
    """Device/host focal part: -w mean |u_t|; exact external samples carry their
    own VJP, so the grid cotangent is then zero (phase_adjoint_optimizer.py:210-223)."""
    dev = isinstance(field, torch.Tensor)
    idx = tuple(np.asarray(indices, dtype=np.intp).T)
    if focal_field is None:
        values = field[idx]
    else:
        values = focal_field
    amp = values.abs() if dev else np.abs(values)
    loss = -weight*(amp.mean() if dev else np.mean(amp))
    cot = None
    if with_cotangent:
        cot = torch.zeros_like(field, dtype=torch.complex128) if dev else np.zeros(field.shape, np.complex128)
        if focal_field is None:
            contrib = -weight/len(indices)*values/(amp + EPS)
            if dev:
                where = tuple(torch.as_tensor(a, device=field.device) for a in idx)
                cot.index_put_(where, contrib.to(torch.complex128), accumulate=True)
            else:
                np.add.at(cot, idx, contrib)
    return loss, cot


def _fortran_rows(indices, shape):
    """Basis rows (Fortran node order) of (i, j, k) grid indices."""
    idx = np.asarray(indices, dtype=np.int64).reshape(-1, 3)
    return idx[:, 0] + shape[0]*(idx[:, 1] + shape[1]*idx[:, 2])


def _basis_focal(G, rows, w, weight, eps, focal_field=None):
    """Focal term -w mean |u_t| on a device basis: with exact samples the grid
    part of the gradient is zero (their gradient is added by the caller, as
    common.py:189-195 does); otherwise u_t = G[rows] w and
    G^H c = G[rows]^H (-w/K u_t/(|u_t| + eps)), repeated rows accumulating."""
    if focal_field is not None:
        return -weight*focal_field.abs().mean(), None
    gt = G.index_select(0, rows)
    vals = gt @ w
    amp = vals.abs()
    return -weight*amp.mean(), gt.conj().transpose(0, 1) @ ((-weight/len(rows))*vals/(amp + eps))


def _device_rows(cache, key, G, make):
    """Per-device cache of a loss's row data (masks, row indices, weights)."""
    hit = cache.get(key)
    if hit is None or hit.device != G.device:
        hit = torch.as_tensor(make(), device=G.device)
        cache[key] = hit
    return hit


class AmplitudeFieldLoss:
    def __init__(self, amplitude_target, eps=EPS):
        self.target = amplitude_target.target
        self.weight = amplitude_target.weight
        self.eps = float(eps)
        self._dev = {}

    def _d(self, name, like):
        if name not in self._dev:
            self._dev[name] = _as_device(getattr(self, name), like)
        return self._dev[name]

    def evaluate(self, field, with_cotangent=True):
        amp = np.abs(field)
        res = amp - self.target
        wres = self.weight*res
        loss = 0.5*float(np.vdot(wres, wres).real)
        if not with_cotangent:
            return loss, None
        return loss, (self.weight*self.weight*res*field/(amp + self.eps)).astype(np.complex128, copy=False)

    def evaluate_device(self, field, with_cotangent=True, focal_field=None):
        target, weight = self._d("target", field), self._d("weight", field)
        amp = field.abs()
        res = amp - target
        wres = weight*res
        loss = 0.5*torch.sum(wres*wres)
        if not with_cotangent:
            return loss, None
        return loss, weight*weight*res*field/(amp + self.eps)

    def evaluate_basis(self, G, w, with_gradient=True, focal_field=None, static=False, focal=True):
        """(loss, G^H c) in one pass over the device basis G (hz_basis_loss)."""
        from . import basis as dev_basis
        tgt = _device_rows(self._dev, "target_rows", G, lambda: self.target.ravel(order="F"))
        wt = _device_rows(self._dev, "weight_rows", G, lambda: self.weight.ravel(order="F"))
        loss, g = dev_basis.fused_loss(G, w, "field_match", target=tgt, weight=wt, eps=self.eps, static=static)
        return loss, (g if with_gradient else None)


class FocalPressureLoss:
    def __init__(self, cfg, amplitude_target, eps=EPS):
        self.cfg = cfg
        self.target = amplitude_target.target
        self.weight = amplitude_target.weight
        self.eps = float(eps)
        self.target_indices = [nearest_grid_index(cfg, p) for p in cfg.targets]
        self.target_weight = float(cfg.training.target_weight)
        self.background_weight = float(cfg.training.background_weight)
        self._dev = {}

    def evaluate(self, field, with_cotangent=True, focal_field=None):
        amp = np.abs(field)
        loss, cot = _focal_term(field, self.target_indices, self.target_weight, with_cotangent, focal_field)
        if self.background_weight > 0:
            loss = loss + 0.5*self.background_weight*float(np.mean(amp*amp))
            if with_cotangent:
                cot = cot + self.background_weight/field.size*field
        return float(loss), cot

    def evaluate_device(self, field, with_cotangent=True, focal_field=None):
        amp = field.abs()
        loss, cot = _focal_term(field, self.target_indices, self.target_weight, with_cotangent, focal_field)
        if self.background_weight > 0:
            loss = loss + 0.5*self.background_weight*torch.mean(amp*amp)
            if with_cotangent:
                cot = cot + self.background_weight/field.numel()*field
        return loss, cot

    def evaluate_basis(self, G, w, with_gradient=True, focal_field=None, static=False, focal=True):
        """(loss, G^H c) on the device basis G: the background energy in one
        fused pass (hz_basis_loss), the focal term on the target rows."""
        from . import basis as dev_basis
        shape = tuple(self.cfg.domain.grid_size)
        if focal:
            rows = _device_rows(self._dev, "rows", G, lambda: _fortran_rows(self.target_indices, shape))
            loss, g = _basis_focal(G, rows, w, self.target_weight, self.eps, focal_field)
        else:
            loss, g = torch.zeros((), dtype=torch.float64, device=G.device), None
        if self.background_weight > 0:
            lb, gb = dev_basis.fused_loss(G, w, "energy", scale=self.background_weight/G.shape[0], static=static)
            loss = loss + lb
            g = gb if g is None else g + gb
        if with_gradient and g is None:
            g = torch.zeros(G.shape[1], dtype=torch.complex128, device=G.device)
        return loss, (g if with_gradient else None)


class FocalContrastLoss:
    def __init__(self, cfg, amplitude_target, eps=EPS):
        tr = cfg.training
        self.cfg = cfg
        self.target = amplitude_target.target
        self.weight = amplitude_target.weight
        self.eps = float(eps)
        self.target_indices = [nearest_grid_index(cfg, p) for p in cfg.targets]
        self.target_weight = float(tr.target_weight)
        self.sidelobe_weight = float(tr.background_weight)
        self.temperature = float(tr.sidelobe_temperature)
        if self.temperature <= 0:
            raise ValueError("training.sidelobe_temperature must be positive")
        self.background_mask = build_background_mask(cfg, np.asarray(cfg.targets, float), float(tr.target_radius),
                                                     int(tr.source_exclusion_layers))
        if not np.any(self.background_mask):
            raise ValueError("focal_contrast background mask is empty")
        self._mask_dev = None
        self._dev = {}

    def evaluate(self, field, with_cotangent=True, focal_field=None):
        amp = np.abs(field)
        loss, cot = _focal_term(field, self.target_indices, self.target_weight, with_cotangent, focal_field)
        bg = amp[self.background_mask]
        top = float(np.max(bg))
        ex = np.exp((bg - top)/self.temperature)
        prob = ex/np.sum(ex)
        loss = float(loss) + self.sidelobe_weight*float(top + self.temperature*(np.log(np.sum(ex)) - np.log(bg.size)))
        if not with_cotangent:
            return loss, None
        cot[self.background_mask] += self.sidelobe_weight*prob*field[self.background_mask]/(bg + self.eps)
        return loss, cot

    def evaluate_device(self, field, with_cotangent=True, focal_field=None):
        if self._mask_dev is None:
            self._mask_dev = torch.as_tensor(self.background_mask, device=field.device)
        mask = self._mask_dev
        amp = field.abs()
        loss, cot = _focal_term(field, self.target_indices, self.target_weight, with_cotangent, focal_field)
        bg = amp[mask]
        top = bg.max()
        ex = torch.exp((bg - top)/self.temperature)
        prob = ex/ex.sum()
        loss = loss + self.sidelobe_weight*(top + self.temperature*(torch.log(ex.sum()) - np.log(bg.numel())))
        if not with_cotangent:
            return loss, None
        cot[mask] += self.sidelobe_weight*prob*field[mask]/(bg + self.eps)
        return loss, cot

    def evaluate_basis(self, G, w, with_gradient=True, focal_field=None, static=False, focal=True):
        """(loss, G^H c) on the device basis G: the smooth-max sidelobe term in
        one fused pass (hz_basis_loss, running-maximum softmax), the focal term
        on the target rows."""
        from . import basis as dev_basis
        shape = tuple(self.cfg.domain.grid_size)
        rows = _device_rows(self._dev, "rows", G, lambda: _fortran_rows(self.target_indices, shape))
        mask = _device_rows(self._dev, "mask_rows", G,
                            lambda: self.background_mask.ravel(order="F").astype(np.uint8))
        loss, g = _basis_focal(G, rows, w, self.target_weight, self.eps, focal_field) if focal else (0.0, None)
        ls, gs = dev_basis.fused_loss(G, w, "softmax", mask=mask, eps=self.eps, scale=self.sidelobe_weight,
                                      temperature=self.temperature,
                                      count_a=int(np.count_nonzero(self.background_mask)), static=static)
        g = gs if g is None else g + gs
        return loss + ls, (g if with_gradient else None)


def build_ratio_masks(cfg, target_points, target_radius):
    """(ball, rest) of the focal contrast ratio: nodes within target_radius of
    a target (the grid from 0 of the reference's background mask,
    phase_adjoint_optimizer.py:384-414) and the nodes of the upper half
    z >= Lz/2 outside every ball."""
    if target_radius <= 0:
        raise ValueError("training.target_radius must be positive")
    d = cfg.domain
    shape = (d.nx, d.ny, d.nz)
    axes = [np.linspace(0.0, d.box_size[a], shape[a]) for a in range(3)]
    ball = np.zeros(shape, dtype=bool)
    for p in np.asarray(target_points, float).reshape(-1, 3):
        r2 = ((axes[0][:, None, None] - p[0])**2 + (axes[1][None, :, None] - p[1])**2
              + (axes[2][None, None, :] - p[2])**2)
        ball |= r2 <= target_radius*target_radius
    rest = np.zeros(shape, dtype=bool)
    rest[:, :, d.nz//2:] = True
    rest &= ~ball
    return ball, rest


class FocalContrastRatioLoss:
    """Focal contrast as a ratio, mean |u|^2 over the target ball over mean
    |u|^2 over the rest of the upper half-domain (loss_type
    "focal_contrast_ratio"; larger is better, so an optimiser minimises its
    negative).  The reference's focal_contrast is the smooth-max form
    (FocalContrastLoss); this is the form a 160^3 gradient workload states.
    Cotangent (dL = Re(c^H du)): 2u/(n_ball mean_rest) on the ball,
    -2 mean_ball u/(n_rest mean_rest^2) on the rest."""

    def __init__(self, cfg, amplitude_target=None, eps=EPS):
        tr = cfg.training
        self.cfg = cfg
        self.eps = float(eps)
        self.target_weight = float(tr.target_weight)
        self.ball, self.rest = build_ratio_masks(cfg, np.asarray(cfg.targets, float), float(tr.target_radius))
        self.n_ball, self.n_rest = int(self.ball.sum()), int(self.rest.sum())
        if not self.n_ball or not self.n_rest:
            raise ValueError("focal_contrast_ratio needs nonempty ball and rest regions")
        self._dev = {}

    def _masks(self, like):
        if "ball" not in self._dev or self._dev["ball"].device != like.device:
            self._dev["ball"] = torch.as_tensor(self.ball, device=like.device)
            self._dev["rest"] = torch.as_tensor(self.rest, device=like.device)
        return self._dev["ball"], self._dev["rest"]

    def evaluate(self, field, with_cotangent=True, focal_field=None):
        p2 = np.abs(field)**2
        na, nb = self.n_ball, self.n_rest
        a, b = float(p2[self.ball].sum())/na, float(p2[self.rest].sum())/nb
        if not with_cotangent:
            return a/b, None
        cot = np.zeros(field.shape, dtype=np.complex128)
        cot[self.ball] = 2.0*field[self.ball]/(na*b)
        cot[self.rest] = -2.0*a*field[self.rest]/(nb*b*b)
        return a/b, cot

    def evaluate_device(self, field, with_cotangent=True, focal_field=None):
        ball, rest = self._masks(field)
        p2 = field.real**2 + field.imag**2
        na, nb = self.n_ball, self.n_rest
        a, b = p2[ball].sum()/na, p2[rest].sum()/nb
        if not with_cotangent:
            return a/b, None
        cot = torch.zeros_like(field, dtype=torch.complex128)
        cot[ball] = 2.0*field[ball]/(na*b)
        cot[rest] = -2.0*a*field[rest]/(nb*b*b)
        return a/b, cot

    def evaluate_basis(self, G, w, with_gradient=True, focal_field=None, static=False, focal=True):
        """(loss, G^H c) in one fused pass over the device basis G (hz_basis_loss)."""
        from . import basis as dev_basis
        cls = _device_rows(self._dev, "class_rows", G, lambda: (self.ball.ravel(order="F").astype(np.uint8)
                                                                + 2*self.rest.ravel(order="F").astype(np.uint8)))
        loss, g = dev_basis.fused_loss(G, w, "ratio", mask=cls, count_a=self.n_ball, count_b=self.n_rest,
                                       static=static)
        return loss, (g if with_gradient else None)


def make_loss(cfg, target=None):
    if cfg.training.loss_type == "focal_contrast_ratio":
        return FocalContrastRatioLoss(cfg)
    target = target or AmplitudeTarget.from_config(cfg)
    kind = cfg.training.loss_type
    if kind == "field_match":
        return AmplitudeFieldLoss(target)
    if kind == "focal_pressure":
        return FocalPressureLoss(cfg, target)
    if kind == "focal_contrast":
        return FocalContrastLoss(cfg, target)
    if kind == "focal_contrast_ratio":
        return FocalContrastRatioLoss(cfg, target)
    raise ValueError("Unknown training.loss_type")


# --------------------------------------------------------------- the objective
def focal_gradient(response, focal, phases, target_weight, eps=EPS):
    """Phase gradient of the focal term -w mean_t |f_t| evaluated on exact
    samples f = R e^{i phi} of a (P, M) target response R
    (src/algorithm/baselines/common.py:189-195): with c_t = -w/P f_t/(|f_t| + eps),
    dL/dphi = Re(conj(R^H c) * i e^{i phi}).  numpy arrays or torch tensors."""
    if isinstance(focal, torch.Tensor):
        ph = torch.as_tensor(phases, dtype=torch.float64, device=focal.device)
        cot = (-target_weight/focal.numel())*focal/(focal.abs() + eps)
        adj = response.conj().transpose(0, 1) @ cot
        return torch.real(adj.conj()*1j*torch.exp(1j*ph))
    focal = np.asarray(focal)
    cot = -target_weight/len(focal)*focal/(np.abs(focal) + eps)
    adj = np.asarray(response).conj().T @ cot
    return np.real(adj.conj()*1j*np.exp(1j*np.asarray(phases, np.float64)))


def wrap_phases(phases):
    phases = np.asarray(phases, dtype=np.float64)
    if not np.isfinite(phases).all():
        raise ValueError("phases must be finite")
    return np.mod(phases, 2*np.pi)


class PhaseObjective:
    """Loss and phase gradient of one scene, evaluated on the device.

    Mirrors ``PhaseProblem.evaluate`` (src/algorithm/baselines/common.py:275-343):
    phases wrapped mod 2 pi, the loss with the exact focal samples (trilinear
    target response) when targets exist, gradient = phase VJP of the grid
    cotangent plus the focal gradient (common.py:189-195), finiteness checks
    and an evaluation history.  Nothing of field size leaves the device: with
    a basis G (moved to the device if needed) each evaluation is one fused
    pass over G that yields the loss and G^H c together
    (``evaluate_basis`` of the loss); without one, a forward solve gives u,
    the torch loss its cotangent c, and an adjoint solve G^H c
    (A^H lambda = c, G^H c = a * B^T lambda) without forming G.  Per
    evaluation only the loss and the M phase gradients reach the host.

    use_graph (basis path): the first evaluation with or without gradient
    captures the whole device step -- e^{i phi}, the exact focal samples,
    the fused pass over G through the allocation-free hz_basis_loss_static,
    the phase gradient -- into a graph on fixed buffers; later evaluations
    copy the phases in, replay it and read loss and gradient back in one
    transfer.  The graph holds the addresses of the device basis and target
    response, which must stay in place.  Evaluations of several objectives
    that replay concurrently on different streams of one GPU would race on
    that entry's shared scratch.
    """

    def __init__(self, solver, basis=None, engine_kwargs=None, device="cuda", use_graph=False):
        self.solver = solver
        self.use_graph = bool(use_graph) and basis is not None
        self._graphs = {}
        self.cfg = solver.cfg
        self.basis = basis
        self.loss_model = make_loss(self.cfg)
        self.size = solver.transducers.num_transducers
        self.shape = tuple(self.cfg.domain.grid_size)
        self.history = []
        self.field_evaluations = 0
        self.vjp_evaluations = 0
        self.engine = None
        if basis is not None:
            if basis.basis_gpu is None:
                basis.to_gpu(device)
            self.device = basis.basis_gpu.device
        else:
            from .autograd import PhaseEngine
            solver.assemble_matrix()
            kw = dict(precision="c64", preconditioner="multigrid" if solver.operator.__dict__.get("hierarchy")
                      else "jacobi", rtol=1e-8, maxiter=100000, chunk=8)
            kw.update(engine_kwargs or {})
            self.engine = PhaseEngine(solver.operator, solver.transducers.amplitudes, **kw)
            self.device = torch.device(solver.operator.device)
        kind = self.cfg.training.loss_type
        if self.cfg.targets and kind not in ("field_match", "focal_contrast_ratio"):
            pts = np.asarray(self.cfg.targets)
            resp = basis.target_response(pts) if basis is not None else solver.target_response(pts)
            self.response = torch.as_tensor(np.asarray(resp, np.complex128), device=self.device)
        else:
            self.response = None

    def _loss_and_adjoint(self, ph, w, gradient, focal):
        """(loss 0-d tensor, G^H c or None) for phases ph, w = e^{i ph}, on the device."""
        kw = {"focal_field": focal} if focal is not None else {}
        if self.basis is not None:
            return self.loss_model.evaluate_basis(self.basis.basis_gpu, w, gradient, **kw)
        from .autograd import grid_view
        u = self.engine.forward(ph[None])[0]
        if not bool(torch.isfinite(torch.view_as_real(u)).all()):
            raise FloatingPointError("Non-finite field or loss")
        loss, cot = self.loss_model.evaluate_device(grid_view(u, self.shape), with_cotangent=gradient, **kw)
        if not gradient:
            return loss, None
        c = cot.permute(2, 1, 0).reshape(1, -1)
        return loss, self.engine.response_adjoint(c)[0]

    def _device_step(self, ph, gradient, static=False):
        """(loss 0-d tensor, phase gradient (M,) tensor or None) for device
        phases ph: the whole evaluation on the device.  With a basis and the
        focal term on exact samples (or none), three launches around the fused
        pass: hz_phase_prologue (e^{i phi}, focal samples), the grid loss and
        G^H c, hz_phase_epilogue (focal term, phase gradient)."""
        from . import basis as dev_basis
        focal_kind = isinstance(self.loss_model, (FocalPressureLoss, FocalContrastLoss))
        exact = self.response is not None and self.response.shape[0] <= dev_basis.PHASE_MAX_TARGETS
        if self.basis is not None and (exact or (self.response is None and not focal_kind)):
            lm = self.loss_model
            w, focal = dev_basis.phase_prologue(ph, self.response)
            loss_t, g = lm.evaluate_basis(self.basis.basis_gpu, w, gradient, static=static, focal=False)
            out = dev_basis.phase_epilogue(w, g, loss_t, self.response, focal,
                                           getattr(lm, "target_weight", 0.0), getattr(lm, "eps", 1.0e-12),
                                           gradient)
            return out[0], (out[1:] if gradient else None)
        w = torch.exp(1j*ph)
        focal = self.response @ w if self.response is not None else None
        if static:
            kw = {"focal_field": focal} if focal is not None else {}
            loss_t, g = self.loss_model.evaluate_basis(self.basis.basis_gpu, w, gradient, static=True, **kw)
        else:
            loss_t, g = self._loss_and_adjoint(ph, w, gradient, focal)
        if not gradient:
            return loss_t, None
        gt = torch.real(torch.conj(g)*(1j*w))
        if focal is not None:
            lm = self.loss_model
            gt = gt + focal_gradient(self.response, focal, ph, lm.target_weight, lm.eps)
        return loss_t, gt

    def _graph_step(self, phases, gradient):
        """Replay of the captured device step (captured on first use)."""
        key = bool(gradient)
        if key not in self._graphs:
            ph = torch.zeros(self.size, dtype=torch.float64, device=self.device)
            out = torch.zeros(self.size + 1, dtype=torch.float64, device=self.device)

            def body():
                loss_t, gt = self._device_step(ph, key, static=True)
                out[:1].copy_(loss_t.reshape(1))
                if gt is not None:
                    out[1:].copy_(gt)

            # Warm-up on a side stream (first-call set-up, such as the
            # static scratch's address, must not happen inside the capture).
            cur = torch.cuda.current_stream(self.device)
            side = torch.cuda.Stream(device=self.device)
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                body()
            cur.wait_stream(side)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                body()
            # Pinned host staging: phases in and results out are asynchronous
            # copies on the replay stream, with one synchronisation.
            ph_host = torch.empty(self.size, dtype=torch.float64, pin_memory=True)
            out_host = torch.empty(self.size + 1, dtype=torch.float64, pin_memory=True)
            self._graphs[key] = (graph, ph, out, ph_host, out_host)
        graph, ph, out, ph_host, out_host = self._graphs[key]
        ph_host.numpy()[:] = phases
        ph.copy_(ph_host, non_blocking=True)
        graph.replay()
        out_host.copy_(out, non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        res = out_host.numpy()
        return float(res[0]), (res[1:].copy() if gradient else None)

    def evaluate(self, phases, gradient=False, stage="optimization"):
        phases = wrap_phases(phases)
        if phases.shape != (self.size,):
            raise ValueError(f"Expected {self.size} phases")
        self.field_evaluations += 1
        if gradient:
            self.vjp_evaluations += 1
        if self.use_graph:
            loss, grad = self._graph_step(phases, gradient)
        else:
            loss_t, gt = self._device_step(torch.as_tensor(phases, device=self.device), gradient)
            grad = gt.cpu().numpy() if gt is not None else None
            loss = float(loss_t)
        if not np.isfinite(loss):
            raise FloatingPointError("Non-finite field or loss")
        if grad is not None and not np.isfinite(grad).all():
            raise FloatingPointError("Non-finite phase gradient")
        self.history.append({"evaluation": self.field_evaluations, "vjp_evaluations": self.vjp_evaluations,
                             "loss": loss, "stage": stage,
                             "gradient_norm": float(np.linalg.norm(grad)) if grad is not None else None})
        return loss, grad

    def gradient_check(self, phases, step=1.0e-3, count=8):
        """Central differences on the first min(count, M) emitters with the
        scaled error |g - fd| / max(|g|, |fd|, 1) (baselines/adjoint/__init__.py:44-58)."""
        _, grad = self.evaluate(phases, gradient=True)
        checks = []
        for index in range(min(count, self.size)):
            e = np.zeros(self.size)
            e[index] = step
            plus, _ = self.evaluate(phases + e, stage="gradient_check")
            minus, _ = self.evaluate(phases - e, stage="gradient_check")
            fd = (plus - minus)/(2*step)
            err = abs(grad[index] - fd)/max(abs(grad[index]), abs(fd), 1.0)
            checks.append({"index": index, "analytic": float(grad[index]), "finite_difference": float(fd),
                           "scaled_error": float(err)})
        return checks
