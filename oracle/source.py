"""FP64 restatement of the sinc-Neumann aperture source and array geometry.

Follows ``src/algorithm/physics/sinc_aperture.py`` (aperture quadrature
:19-33, bilinear deposition with unit column mass :36-72, flux RHS columns
:75-85), ``src/algorithm/physics/warp_utils.py:118-120`` (b = B (a * e^{i phi}))
and ``src/algorithm/physics/transducer_array.py`` (ideal-grid centres
:32-44, geometric phases :46-70).

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
import scipy.sparse as sp


def aperture_quadrature(diameter, radial_order=32, angular_order=128):
    """Gauss-Legendre in mu = cos(theta) x uniform azimuth (sinc_aperture.py:19-33)."""
    if not np.isfinite(diameter) or diameter <= 0:
        raise ValueError("Aperture diameter must be positive")
    if radial_order < 2 or angular_order < 8:
        raise ValueError("Aperture quadrature orders are too small")
    nodes, w = np.polynomial.legendre.leggauss(radial_order)
    mu = 0.5*(nodes + 1.0)
    w = 0.5*w
    r = 0.5*diameter*np.sqrt(1.0 - mu*mu)
    phi = 2.0*np.pi*np.arange(angular_order)/angular_order
    px = (r[:, None]*np.cos(phi)[None, :]).ravel()
    py = (r[:, None]*np.sin(phi)[None, :]).ravel()
    offsets = np.stack([px, py, np.zeros_like(px)], axis=1)
    weights = np.repeat(w/angular_order, angular_order)
    return offsets, weights


def source_mass_matrix(shape_xy, dx, centers, diameter, radial_order=64, angular_order=256):
    """(nx*ny) x M CSR of deposition weights, unit mass per column (sinc_aperture.py:36-72)."""
    shape_xy = tuple(int(v) for v in shape_xy)
    centers = np.asarray(centers, dtype=float)
    if (len(shape_xy) != 2 or min(shape_xy) < 3 or not np.isfinite(dx) or dx <= 0
            or centers.ndim != 2 or centers.shape[1] != 3 or len(centers) == 0
            or not np.isfinite(centers).all() or np.any(centers[:, 2] != 0)):
        raise ValueError("Sources require finite centers on z=0 and a valid xy grid")
    offsets, weights = aperture_quadrature(diameter, radial_order, angular_order)
    extent = (np.asarray(shape_xy) - 1)*dx
    if np.any(centers[:, :2] - diameter/2 < dx) or np.any(centers[:, :2] + diameter/2 > extent - dx):
        raise ValueError("Sinc aperture needs one grid cell clearance from side boundaries")
    nx = shape_xy[0]
    all_rows, all_cols, all_vals = [], [], []
    for col, centre in enumerate(centers):
        g = (centre[None, :2] + offsets[:, :2])/dx
        base = np.floor(g).astype(np.int64)
        t = g - base
        rows = []
        vals = []
        for ox in (0, 1):
            fx = t[:, 0] if ox else 1.0 - t[:, 0]
            for oy in (0, 1):
                fy = t[:, 1] if oy else 1.0 - t[:, 1]
                rows.append(base[:, 0] + ox + nx*(base[:, 1] + oy))
                vals.append(weights*fx*fy)
        # Accumulate per distinct node in the order (00, 01, 10, 11) of the reference.
        rows = np.concatenate([rows[0], rows[1], rows[2], rows[3]])
        vals = np.concatenate([vals[0], vals[1], vals[2], vals[3]])
        uniq, inv = np.unique(rows, return_inverse=True)
        all_rows.append(uniq)
        all_cols.append(np.full(len(uniq), col))
        all_vals.append(np.bincount(inv, weights=vals))
    m = sp.coo_matrix((np.concatenate(all_vals), (np.concatenate(all_rows), np.concatenate(all_cols))),
                      shape=(int(np.prod(shape_xy)), len(centers))).tocsr()
    m.sum_duplicates()
    m.eliminate_zeros()
    return m


def flux_rhs_columns(masses, shape, dx, reference_density):
    """B = [-2 pi/(rho_ref h^3) masses; 0] (sinc_aperture.py:75-85)."""
    if not np.isfinite(reference_density) or reference_density <= 0:
        raise ValueError("reference_density must be positive")
    top = -2.0*np.pi/(reference_density*dx**3)*masses
    rest = sp.csr_matrix((int(np.prod(shape)) - masses.shape[0], masses.shape[1]))
    return sp.vstack((top, rest), format="csr")


def rhs_for_phases(columns, amplitudes, phases):
    """b = B (a e^{i phi}) (warp_utils.py:118-120)."""
    return columns @ (np.asarray(amplitudes)*np.exp(1j*np.asarray(phases)))


def ideal_grid_centers(array_n, pitch, lx, ly, origin=(0.0, 0.0, 0.0)):
    """n x n centres about the domain centre on z = 0 (transducer_array.py:32-44)."""
    origin = np.asarray(origin, dtype=float)
    cx = origin[0] + 0.5*lx
    cy = origin[1] + 0.5*ly
    half = 0.5*(array_n - 1)*pitch
    out = np.empty((array_n*array_n, 3))
    idx = 0
    for i in range(array_n):
        for j in range(array_n):
            out[idx] = (cx - half + i*pitch, cy - half + j*pitch, 0.0)
            idx += 1
    return out


def geometric_phases(centers, targets, wavelength, k0):
    """Single target: 2 pi (ceil(d/lambda) - d/lambda); several: arg of the
    time-reversed superposition (transducer_array.py:46-70)."""
    centers = np.asarray(centers, dtype=float)
    if len(targets) == 0:
        return np.zeros(len(centers))
    if len(targets) == 1:
        d = np.linalg.norm(centers - np.asarray(targets[0], dtype=float), axis=-1)
        cyc = d/wavelength
        return 2.0*np.pi*(np.ceil(cyc) - cyc)
    acc = np.zeros(len(centers), dtype=np.complex128)
    for t in targets:
        d = np.linalg.norm(centers - np.asarray(t, dtype=float), axis=-1)
        acc += np.exp(-1j*k0*d)/d
    return np.angle(acc)
