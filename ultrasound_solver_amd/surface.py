"""Homogeneous surface condensation on the device: the drop-in for the
reference's ``HomogeneousSurfaceSolver``
(src/algorithm/solvers/homogeneous_surface.py:113-259).

In a homogeneous medium over a reflecting -z baffle the finite-volume
operator is the Kronecker sum of three 1-D axis operators.  The x and y
operators are diagonalised on the host (LAPACK); every (x, y) mode then owns a tridiagonal z system whose solution is
u[j] = H[j-1] u[0], with the transfer ratios H from a backward continued
fraction and u[0] from the surface Schur complement.  The device holds the
eigenvector matrices, H and the Schur complement (``hz_surface_*`` in
``include/hz.h``): the recurrence, the modal transforms, the emitter
responses and the field samples all run there.

Same constructor, validation, attributes and method results as the
reference class, with two differences: ``device`` must name the HIP device
(there is no host path; the reference's numpy branch lives only in the
test oracle), and ``transform`` / ``inverse_transform`` / ``surface_rhs`` /
``surface_modes`` / ``surface`` / ``h`` / ``schur`` are torch device
tensors where the reference returns CuPy arrays.
"""
import ctypes
import functools
import hashlib

import numpy as np
import scipy.linalg as la
import scipy.sparse as sps
import torch

from . import _capi, physics

FACES = physics.FACES


def axis_operators(shape, dx, omega, density, sound_speed, boundaries, pml):
    """The three 1-D operators whose Kronecker sum, plus k^2, is the
    homogeneous finite-volume operator divided by (1/rho) prod_a w_a s_a, and
    each axis' PML stretch s.

    Along axis a: L u_i = (1/(w_i s_i)) sum_{edges i-j} e_ij (u_j - u_i) with
    e_ij = 1/(0.5 (s_i + s_j) h^2), plus i omega rho Y/(h w_i s_i) at an open
    (Y = 1/(rho c)) or impedance (Y = 1/Z) end.  The weights w, stretches s
    and admittances come from the production tables (physics.axis_tables,
    physics.face_admittance) of the full grid, so the PML and face checks of
    the 3-D operator apply.  Soft faces have no such separable form."""
    kinds = [physics.boundary_type(boundaries.get(f, "reflecting")) for f in FACES]
    if any(k not in ("open", "reflecting", "impedance") for k in kinds):
        raise ValueError("Surface diagnostic supports open/reflecting/impedance faces")
    weights, stretch = physics.axis_tables(shape, boundaries, pml or {})
    out = []
    for axis in range(3):
        w, s = weights[axis], stretch[axis]
        e = 1.0/(0.5*(s[:-1] + s[1:])*dx*dx)
        diag = -np.concatenate((e, [0.0])) - np.concatenate(([0.0], e))
        for end, face in ((0, FACES[2*axis]), (-1, FACES[2*axis + 1])):
            y = physics.face_admittance(boundaries.get(face, "reflecting"), density, sound_speed)
            if y is not None:
                diag[end] += 1j*omega*density*y/dx
        band = sps.diags([e, diag, e], [-1, 0, 1], shape=(len(w), len(w))).toarray()
        out.append((band/(w*s)[:, None], s))
    return out


@functools.lru_cache(maxsize=8)
def _axis_modes(shape, dx, omega, density, sound_speed, faces_key, pml_key, axis):
    """Eigen-decomposition (values, vectors, inverse, 1-norm condition number
    of the vectors) of one axis operator, cached by the definition of the
    axis, so that the x and y axes of a square grid, and repeated solvers of
    one scene, decompose once."""
    boundaries = {f: (dict(v) if isinstance(v, tuple) else v) for f, v in faces_key}
    matrix = axis_operators(shape, dx, omega, density, sound_speed, boundaries, dict(pml_key))[axis][0]
    values, vectors = la.eig(matrix)
    inverse = la.inv(vectors)
    return values, vectors, inverse, float(la.norm(vectors, 1)*la.norm(inverse, 1))


def _freeze(mapping):
    """Hashable form of a boundary or PML mapping (dict values become sorted tuples)."""
    def item(v):
        if isinstance(v, dict):
            return tuple(sorted((k, item(x)) for k, x in v.items()))
        if isinstance(v, list):
            return tuple(item(x) for x in v)
        return v
    return tuple(sorted((k, item(v)) for k, v in (mapping or {}).items()))


def _fortran(a):
    """Host complex array -> contiguous float64 view in column-major order."""
    return np.asfortranarray(np.asarray(a, dtype=np.complex128)).ravel(order="K").view(np.float64)


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class HomogeneousSurfaceSolver:
    def __init__(self, shape, dx, frequency, sound_speed, density, boundaries, pml=None, device="cuda"):
        if not all(np.ndim(v) == 0 and np.isfinite(v) and v > 0 for v in (dx, frequency, sound_speed, density)):
            raise ValueError("This diagnostic requires positive scalar homogeneous materials")
        if set(boundaries) - set(FACES):
            raise ValueError("Unknown boundary face")
        if not str(device).startswith("cuda"):
            raise ValueError("device must be cuda: the surface solver runs on the HIP device only")
        self.shape = tuple(int(v) for v in shape)
        if len(self.shape) != 3 or min(self.shape) < 2:
            raise ValueError("shape must have three edges >= 2")
        self.dx = float(dx)
        self.device = torch.device(device)
        self.boundaries = boundaries.copy()
        omega = 2*np.pi*frequency
        axes = axis_operators(self.shape, self.dx, omega, density, sound_speed, boundaries, pml or {})
        self.axes = [a[0] for a in axes]
        self.surface_pml = (axes[0][1].imag[:, None] + axes[1][1].imag[None, :] + axes[2][1].imag[0]) > 0
        self.surface_rhs = self.surface_modes = self.surface = None
        key = (self.shape, self.dx, float(omega), float(density), float(sound_speed), _freeze(boundaries),
               _freeze(pml))
        ex, vx, ix, cond_x = _axis_modes(*key, 0)
        ey, vy, iy, cond_y = _axis_modes(*key, 1)
        self.eigen_condition = [cond_x, cond_y]
        shift = ex[:, None] + ey[None, :] + (omega/sound_speed)**2
        z = self.axes[2]
        keep = [_fortran(vx), _fortran(ix), _fortran(vy), _fortran(iy), _fortran(shift),
                _fortran(np.diag(z)), _fortran(np.diag(z, 1)), _fortran(np.diag(z, -1))]
        nx, ny, nz = self.shape
        desc = _capi.SurfaceDesc(nx, ny, nz, self.dx, *[_dptr(a) for a in keep])
        handle = ctypes.c_void_p()
        _capi.check(_capi.lib().hz_surface_create(ctypes.byref(desc), ctypes.byref(handle)), "hz_surface_create")
        self.handle = handle
        dev = self.device
        self.vx, self.ix = (torch.as_tensor(a, device=dev) for a in (vx, ix))
        self.vy, self.iy = (torch.as_tensor(a, device=dev) for a in (vy, iy))
        self._masses_key = None
        self.diagnostics = {}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            _capi.lib().hz_surface_destroy(h)
            self.handle = None

    # -- device planes (nx x ny, x fastest) -------------------------------
    def _plane_buffer(self, count=None):
        nx, ny = self.shape[:2]
        shape = (ny, nx) if count is None else (count, ny, nx)
        return torch.empty(shape, dtype=torch.complex128, device=self.device)

    @staticmethod
    def _as_planes(buf):
        """(.., ny, nx) device buffer -> (.., nx, ny) view."""
        return buf.transpose(-1, -2)

    def _to_device_planes(self, plane):
        t = torch.as_tensor(plane).to(self.device, torch.complex128)
        if t.shape[-2:] != self.shape[:2] or t.ndim not in (2, 3):
            raise ValueError("planes must be (nx, ny) or (count, nx, ny)")
        return t.transpose(-1, -2).contiguous()

    def _transform(self, plane, inverse):
        src = self._to_device_planes(plane)
        count = 1 if src.ndim == 2 else src.shape[0]
        out = torch.empty_like(src)
        _capi.check(_capi.lib().hz_surface_transform(self.handle, int(inverse), ctypes.c_void_p(src.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), count, _stream()),
                     "hz_surface_transform")
        return self._as_planes(out)

    def transform(self, plane):
        """ix @ plane @ iy.T (homogeneous_surface.py:166-167)."""
        return self._transform(plane, False)

    def inverse_transform(self, plane):
        """vx @ plane @ vy.T (homogeneous_surface.py:169-170)."""
        return self._transform(plane, True)

    @property
    def h(self):
        """Transfer ratios H[j] = u[j+1]/u[0], (nz-1, nx, ny) on the device."""
        nx, ny, nz = self.shape
        out = torch.empty((nz - 1, ny, nx), dtype=torch.complex128, device=self.device)
        _capi.check(_capi.lib().hz_surface_tables(self.handle, ctypes.c_void_p(out.data_ptr()), None, _stream()),
                    "hz_surface_tables")
        return self._as_planes(out)

    @property
    def schur(self):
        """Surface Schur complement shift + z[0,0] + z[0,1] H[0], (nx, ny)."""
        out = self._plane_buffer()
        _capi.check(_capi.lib().hz_surface_tables(self.handle, None, ctypes.c_void_p(out.data_ptr()), _stream()),
                    "hz_surface_tables")
        return self._as_planes(out)

    # -- sources ----------------------------------------------------------
    def _validate_flux_masses(self, masses):
        """homogeneous_surface.py:178-187."""
        if physics.boundary_type(self.boundaries.get("-z", "reflecting")) != "reflecting":
            raise ValueError("Sinc Neumann source requires a reflecting/baffled bottom")
        if (masses.shape[0] != np.prod(self.shape[:2]) or masses.shape[1] == 0
                or not np.isfinite(masses.data).all()):
            raise ValueError("Invalid source masses")
        active = np.asarray(abs(masses).sum(axis=1)).ravel().reshape(self.shape[:2], order="F") > 0
        if np.any(active & self.surface_pml) or any(np.any(a) for a in
                                                   (active[0], active[-1], active[:, 0], active[:, -1])):
            raise ValueError("Flux source overlaps PML or side boundaries")

    def _set_masses(self, masses):
        coo = masses.tocoo()
        key = (coo.shape, coo.nnz, hashlib.sha1(np.ascontiguousarray(coo.row).tobytes()
                                                 + np.ascontiguousarray(coo.col).tobytes()
                                                 + np.ascontiguousarray(coo.data, dtype=np.float64).tobytes()).digest())
        if key == self._masses_key:
            return
        node = np.ascontiguousarray(coo.row, dtype=np.int64)
        col = np.ascontiguousarray(coo.col, dtype=np.int32)
        val = np.ascontiguousarray(coo.data, dtype=np.float64)
        _capi.check(_capi.lib().hz_surface_set_sources(
            self.handle, coo.shape[1], len(val), node.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dptr(val)), "hz_surface_set_sources")
        self._masses_key = key

    def _sample_coordinates(self, points):
        """homogeneous_surface.py:238-248."""
        points = np.asarray(points, float)
        coordinates = points/self.dx
        if (points.ndim != 2 or points.shape[1] != 3 or not np.isfinite(points).all()
                or np.any(coordinates < -1e-9)
                or np.any(coordinates > np.asarray(self.shape) - 1 + 1e-9)):
            raise ValueError("Points must lie inside the local grid")
        coordinates = np.clip(coordinates, 0, np.asarray(self.shape) - 1)
        low = np.minimum(np.floor(coordinates).astype(int), np.asarray(self.shape) - 2)
        return low, coordinates - low

    def _point_args(self, points):
        low, frac = self._sample_coordinates(points)
        low32 = np.ascontiguousarray(low, dtype=np.int32)
        frac = np.ascontiguousarray(frac, dtype=np.float64)
        return low32, frac, (low32, frac)

    # -- solves -----------------------------------------------------------
    def flux_target_response(self, masses, points):
        """(points x channels) responses of the discrete PDE with ordinary
        transposes, p = response @ w (homogeneous_surface.py:189-205)."""
        self._validate_flux_masses(masses)
        self._set_masses(masses)
        low, frac, _ = self._point_args(points)
        out = torch.empty((len(low), masses.shape[1]), dtype=torch.complex128, device=self.device)
        _capi.check(_capi.lib().hz_surface_target_response(
            self.handle, low.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dptr(frac), len(low),
            ctypes.c_void_p(out.data_ptr()), _stream()), "hz_surface_target_response")
        return out.cpu().numpy()

    def solve_flux(self, masses, phases, amplitudes=None):
        """Normalized baffled Neumann aperture (homogeneous_surface.py:207-229)."""
        self._validate_flux_masses(masses)
        phases = np.asarray(phases, dtype=float)
        amplitudes = np.ones_like(phases) if amplitudes is None else np.asarray(amplitudes)
        if (phases.ndim != 1 or amplitudes.shape != phases.shape
                or masses.shape != (np.prod(self.shape[:2]), len(phases))
                or not np.isfinite(phases).all() or not np.isfinite(amplitudes).all()
                or not np.isfinite(masses.data).all()):
            raise ValueError("Invalid source masses or phases")
        self._set_masses(masses)
        w = torch.as_tensor(amplitudes*np.exp(1j*phases), dtype=torch.complex128, device=self.device)
        rhs, modes, surface = self._plane_buffer(), self._plane_buffer(), self._plane_buffer()
        _capi.check(_capi.lib().hz_surface_solve_flux(
            self.handle, ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(rhs.data_ptr()),
            ctypes.c_void_p(modes.data_ptr()), ctypes.c_void_p(surface.data_ptr()), _stream()),
            "hz_surface_solve_flux")
        self.surface_rhs, self.surface_modes, self.surface = (self._as_planes(t) for t in (rhs, modes, surface))
        self.diagnostics = {"iterations": 0, "eigenvector_condition": self.eigen_condition,
                            "source_integral_min_max": [float(masses.sum(axis=0).min()),
                                                       float(masses.sum(axis=0).max())]}
        return self.diagnostics

    def plane(self, index):
        """Plane `index` of the last solved field (homogeneous_surface.py:231-236)."""
        if not 0 <= index < self.shape[2]:
            raise ValueError("z index outside grid")
        if self.surface is None:
            raise RuntimeError("call solve_flux first")
        if index == 0:
            return self.surface.cpu().numpy()
        out = self._plane_buffer()
        z = (ctypes.c_int32*1)(int(index))
        _capi.check(_capi.lib().hz_surface_planes(self.handle, z, 1, ctypes.c_void_p(out.data_ptr()), _stream()),
                    "hz_surface_planes")
        return self._as_planes(out).cpu().numpy()

    def sample(self, points):
        """Trilinear physical-grid interpolation of the last solved field,
        with no target snapping (homogeneous_surface.py:250-259)."""
        if self.surface is None:
            raise RuntimeError("call solve_flux first")
        low, frac, _ = self._point_args(points)
        out = torch.empty(len(low), dtype=torch.complex128, device=self.device)
        _capi.check(_capi.lib().hz_surface_sample(self.handle, low.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                  _dptr(frac), len(low), ctypes.c_void_p(out.data_ptr()), _stream()),
                    "hz_surface_sample")
        return out.cpu().numpy()
