"""FP64 restatement of the homogeneous surface-condensation solver.

Follows ``src/algorithm/solvers/homogeneous_surface.py``: the 1-D axis
operators (``axis_matrix`` :85-110), the x/y eigendecomposition with its
1-norm condition estimate (:20-38), the backward continued fraction
H[j] = u[j+1]/u[0] and its prefix products (:151-163, the same recurrence
as the CUDA ``z_transfer`` kernel :41-82), the surface Schur complement
(:164), the modal transforms (:169-173), ``flux_target_response``
(:189-205), ``solve_flux`` (:207-229), ``plane`` (:231-236) and ``sample``
(:250-259).

Test infrastructure only (see ``oracle/__init__.py``).  Pinned by the
mathematics it condenses: for a homogeneous medium with a reflecting -z face
the surface solve must reproduce the full 3-D finite-volume solve of
``oracle.helmholtz`` (itself pinned by the reference's CSR fixtures) with
the flux right-hand side of ``oracle.source``, plane by plane;
``flux_target_response`` must agree with ``solve_flux`` + ``sample`` by
linearity; each mode's transfer ratios must solve its tridiagonal z system.
"""
import numpy as np
import scipy.linalg as la

FACES = ("-x", "+x", "-y", "+y", "-z", "+z")


def _kind(value):
    return value.get("type", "reflecting") if isinstance(value, dict) else value


def _axis_weights_stretch(n, faces, pml):
    """Half-volume weights and PML stretch of one axis, as the operator's
    constructor builds them for x (conservative_helmholtz.py:40-64)."""
    pml = pml or {}
    cells = int(pml.get("cells", 0))
    strength = float(pml.get("strength", 2.0))
    order = int(pml.get("order", 3))
    w = np.ones(n)
    w[[0, -1]] = 0.5
    sigma = np.zeros(n)
    active = [f for f in ("-x", "+x") if f in pml.get("faces", [])]
    if cells and active:
        if cells*len(active) >= n - 1:
            raise ValueError("PML leaves no physical interior on an axis")
        ramp = strength*np.maximum(1.0 - np.arange(n)/cells, 0.0)**order
        for f in active:
            if _kind(faces.get(f, "reflecting")) != "open":
                raise ValueError("PML may only be placed on open faces")
            sigma = sigma + (ramp if f == "-x" else ramp[::-1])
    return w, 1.0 + 1j*sigma


def axis_matrix(n, dx, omega, density, sound_speed, faces, pml, axis):
    """(n x n) complex 1-D operator of one axis and its stretch
    (homogeneous_surface.py:85-110)."""
    translated = {"-x": faces.get(FACES[2*axis], "reflecting"), "+x": faces.get(FACES[2*axis + 1], "reflecting")}
    if any(_kind(v) not in {"open", "reflecting", "impedance"} for v in translated.values()):
        raise ValueError("Surface diagnostic supports open/reflecting/impedance faces")
    pml = pml or {}
    original = pml.get("faces", [f for f in FACES if _kind(faces.get(f, "reflecting")) == "open"])
    local = {**pml, "faces": [("-x", "+x")[i] for i in range(2) if FACES[2*axis + i] in original]}
    w, s = _axis_weights_stretch(n, translated, local)
    edge = 1/(.5*(s[:-1] + s[1:])*dx*dx)
    diagonal = np.zeros(n, complex)
    diagonal[:-1] -= edge
    diagonal[1:] -= edge
    matrix = np.diag(diagonal) + np.diag(edge, 1) + np.diag(edge, -1)
    matrix /= (w*s)[:, None]
    for i, face in ((0, "-x"), (-1, "+x")):
        kind = _kind(translated[face])
        if kind in {"open", "impedance"}:
            impedance = density*sound_speed if kind == "open" else complex(*translated[face]["impedance"])
            matrix[i, i] += 1j*omega*density/(impedance*dx*w[i]*s[i])
    return matrix, s


def eigendecomposition(matrix):
    """(eigenvalues, eigenvectors, inverse, 1-norm condition) (homogeneous_surface.py:20-38)."""
    e, v = la.eig(matrix)
    inv = la.inv(v)
    return e, v, inv, float(la.norm(v, 1)*la.norm(inv, 1))


def z_transfer(shift, z):
    """H (nz-1, *shift.shape): H[j] = u[j+1]/u[0] of every mode by the
    backward continued fraction, then prefix products
    (homogeneous_surface.py:151-163)."""
    nz = z.shape[0]
    h = np.empty((nz - 1, *shift.shape), dtype=np.complex128)
    ratio = np.zeros(shift.shape, dtype=np.complex128)
    for j in range(nz - 1, 0, -1):
        diagonal = shift + z[j, j]
        if j < nz - 1:
            diagonal = diagonal + z[j, j + 1]*ratio
        ratio = -z[j, j - 1]/diagonal
        h[j - 1] = ratio
    for j in range(1, nz - 1):
        h[j] *= h[j - 1]
    return h


class SurfaceSolver:
    """CPU path of HomogeneousSurfaceSolver (homogeneous_surface.py:113-259)."""

    def __init__(self, shape, dx, frequency, sound_speed, density, boundaries, pml=None):
        self.shape = tuple(shape)
        self.dx = dx
        self.boundaries = dict(boundaries)
        omega = 2*np.pi*frequency
        axes = [axis_matrix(n, dx, omega, density, sound_speed, boundaries, pml or {}, a)
                for a, n in enumerate(self.shape)]
        self.axes = [a[0] for a in axes]
        self.surface_pml = (axes[0][1].imag[:, None] + axes[1][1].imag[None, :] + axes[2][1].imag[0]) > 0
        ex, self.vx, self.ix, cx = eigendecomposition(self.axes[0])
        ey, self.vy, self.iy, cy = eigendecomposition(self.axes[1])
        self.eigen_condition = [cx, cy]
        self.shift = ex[:, None] + ey[None, :] + (omega/sound_speed)**2
        z = self.axes[2]
        self.h = z_transfer(self.shift, z)
        self.schur = self.shift + z[0, 0] + z[0, 1]*self.h[0]

    def transform(self, plane):
        return self.ix @ plane @ self.iy.T

    def inverse_transform(self, plane):
        return self.vx @ plane @ self.vy.T

    def solve_flux(self, masses, phases, amplitudes=None):
        phases = np.asarray(phases, float)
        amplitudes = np.ones_like(phases) if amplitudes is None else np.asarray(amplitudes, float)
        self.surface_rhs = (-4*np.pi/self.dx**3*(masses @ (amplitudes*np.exp(1j*phases)))).reshape(
            self.shape[:2], order="F")
        self.surface_modes = self.transform(self.surface_rhs)/self.schur
        self.surface = self.inverse_transform(self.surface_modes)

    def plane(self, index):
        if index == 0:
            return self.surface
        return self.inverse_transform(self.h[index - 1]*self.surface_modes)

    def sample_coordinates(self, points):
        c = np.clip(np.asarray(points, float)/self.dx, 0, np.asarray(self.shape) - 1)
        low = np.minimum(np.floor(c).astype(int), np.asarray(self.shape) - 2)
        return low, c - low

    def sample(self, points):
        low, frac = self.sample_coordinates(points)
        out = np.zeros(len(low), complex)
        for q, ((x, y, z), (fx, fy, fz)) in enumerate(zip(low, frac)):
            for dz in (0, 1):
                pl = self.plane(z + dz)
                for a in (0, 1):
                    for b in (0, 1):
                        w = (fx if a else 1 - fx)*(fy if b else 1 - fy)*(fz if dz else 1 - fz)
                        out[q] += w*pl[x + a, y + b]
        return out

    def flux_target_response(self, masses, points):
        low, frac = self.sample_coordinates(points)
        out = np.empty((len(low), masses.shape[1]), complex)
        for q, ((x, y, z), (fx, fy, fz)) in enumerate(zip(low, frac)):
            vx = (1 - fx)*self.vx[x] + fx*self.vx[x + 1]
            vy = (1 - fy)*self.vy[y] + fy*self.vy[y + 1]
            hz = (1 - fz)*(1 if z == 0 else self.h[z - 1]) + fz*self.h[z]
            observation = vx[:, None]*vy[None, :]*hz/self.schur
            kernel = self.ix.T @ observation @ self.iy
            out[q] = -4*np.pi/self.dx**3*(masses.T @ kernel.ravel(order="F"))
        return out
