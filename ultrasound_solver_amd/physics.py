"""Host-side scene setup for the HIP operator: per-axis tables, face terms,
materials, transducer centres and the sinc-Neumann source deposition.

Everything here is O(n^2) or O(M * quadrature) preprocessing done once per
scene; the O(N) operator itself never exists on the host.  Conventions
follow the reference so that the device operator reproduces its matrix A:

* per-axis half-volume weights and PML stretch s = 1 + i sigma,
  sigma = strength * max(1 - d/cells, 0)^order
  (src/algorithm/physics/conservative_helmholtz.py:48-64);
* open/impedance face terms i omega (w_b w_c s_b s_c / h) Y, Y = 1/(rho c)
  or 1/Z (conservative_helmholtz.py:89-103);
* sinc aperture quadrature and bilinear unit-mass deposition, the masses
  computed on the device (src/algorithm/physics/sinc_aperture.py:19-72), flux scale
  -2 pi / (rho_ref h^3) (sinc_aperture.py:75-85);
* material blend of 1/rho and 1/(rho c^2) across a smoothed SDF, rasterised
  on the device (src/algorithm/physics/warp_utils.py:11-23, 44-79).
"""
import numpy as np

FACES = ("-x", "+x", "-y", "+y", "-z", "+z")
KINDS = ("reflecting", "open", "soft", "impedance")


def boundary_type(value):
    return value.get("type", "reflecting") if isinstance(value, dict) else value


def pml_settings(pml, boundaries):
    pml = dict(pml or {})
    cells = int(pml.get("cells", 0))
    strength = float(pml.get("strength", 2.0))
    order = int(pml.get("order", 3))
    if cells < 0 or order < 1 or not np.isfinite(strength) or strength < 0:
        raise ValueError("Invalid PML cells, strength or order")
    faces = pml.get("faces")
    if faces is None:
        faces = [f for f in FACES if boundary_type(boundaries.get(f, "reflecting")) == "open"]
    if set(faces) - set(FACES):
        raise ValueError("Unknown PML face")
    return cells, strength, order, list(faces)


def axis_tables(shape, boundaries, pml=None):
    """Per-axis weights (real) and stretches (complex), validated like the reference."""
    if set(boundaries) - set(FACES):
        raise ValueError("Unknown boundary face")
    cells, strength, order, faces = pml_settings(pml, boundaries)
    weights, stretch = [], []
    for axis, n in enumerate(shape):
        w = np.ones(n)
        w[[0, -1]] = 0.5
        sigma = np.zeros(n)
        here = [f for f in FACES[2*axis:2*axis + 2] if f in faces]
        if cells and here:
            if cells*len(here) >= n - 1:
                raise ValueError("PML leaves no physical interior on an axis")
            ramp = strength*np.maximum(1.0 - np.arange(n)/cells, 0.0)**order
            for f in here:
                if boundary_type(boundaries.get(f, "reflecting")) != "open":
                    raise ValueError("PML may only be placed on open faces")
                sigma = sigma + (ramp if f[0] == "-" else ramp[::-1])
        weights.append(w)
        stretch.append(1.0 + 1j*sigma)
    return weights, stretch


def interior_bounds(box_size, dx, boundaries, pml=None):
    """(lo, hi): per axis, the local coordinates between which a point is
    outside every PML layer (a face with PML moves its bound inward by
    cells * dx; helmholtz_solver.py:105-123 and phase_response_basis.py:153-164
    reject target points beyond it)."""
    cells, _, _, faces = pml_settings(pml, boundaries)
    box = np.asarray(box_size, dtype=float)
    lo, hi = np.zeros(3), box.copy()
    if cells:
        for f in faces:
            axis = "xyz".index(f[1])
            if f[0] == "-":
                lo[axis] = cells*dx
            else:
                hi[axis] = box[axis] - cells*dx
    return lo, hi


def pml_node_mask(stretch, shape):
    sx, sy, sz = (np.imag(s) for s in stretch)
    return (sx[:, None, None] + sy[None, :, None] + sz[None, None, :]) > 0


def face_admittance(spec, rho, c):
    """Admittance Y of a face: 1/(rho c) for an open face (rho, c scalars or
    the face's node arrays), 1/Z for an impedance face given as
    {type: impedance, impedance: [Re Z, Im Z]} in Pa s/m, None otherwise
    (conservative_helmholtz.py:89-103)."""
    kind = boundary_type(spec)
    if kind == "open":
        return 1.0/(rho*c)
    if kind != "impedance":
        return None
    if not isinstance(spec, dict):
        raise ValueError("impedance requires {type: impedance, impedance: [real, imag]}")
    z = np.asarray(spec.get("impedance"), dtype=float)
    if z.shape != (2,) or not np.isfinite(z).all() or z[0] < 0 or not np.any(z):
        raise ValueError("impedance must be finite [real>=0, imag], nonzero, in Pa s/m")
    return 1.0/complex(z[0], z[1])


def face_tables(rho, c, dx, omega, boundaries, weights, stretch):
    """(kinds[6], tables[6]); tables are complex (n_b, n_c) arrays flattened with
    the lower remaining axis fastest, or None for reflecting/soft faces."""
    kinds, tables = [], []
    for index, face in enumerate(FACES):
        axis = index//2
        spec = boundaries.get(face, "reflecting")
        kind = boundary_type(spec)
        if kind not in KINDS:
            raise ValueError(f"Unsupported boundary {face}: {spec!r}")
        kinds.append(kind)
        if kind not in ("open", "impedance"):
            tables.append(None)
            continue
        b, cc = [a for a in range(3) if a != axis]
        area = (weights[b][:, None]*weights[cc][None, :]*stretch[b][:, None]*stretch[cc][None, :])/dx
        sel = [slice(None)]*3
        sel[axis] = 0 if index % 2 == 0 else -1
        sel = tuple(sel)
        term = 1j*omega*area*face_admittance(spec, rho[sel], c[sel])
        tables.append(np.ascontiguousarray(term.ravel(order="F"), dtype=np.complex128))
    return kinds, tables


def soft_node_mask(shape, kinds):
    mask = np.zeros(shape, dtype=bool)
    for index, kind in enumerate(kinds):
        if kind == "soft":
            sel = [slice(None)]*3
            sel[index//2] = 0 if index % 2 == 0 else -1
            mask[tuple(sel)] = True
    return mask


# ----------------------------------------------------------------- materials
PRIMITIVES = {"sphere": 0, "box": 1, "ellipsoid": 2}


def rasterize_materials(shape, dx, primitives, rho0, c0, rho_obstacle=1250.0, c_obstacle=2200.0,
                        smoothing_cells=1.2, rho_medium=None, c_medium=None, want_sdf=False):
    """Density and sound speed of a union of obstacles in a medium, on the
    device (``hz_rasterize_materials``; warp_utils.py:11-23, 52-77).

    primitives: ``(kind, centre, size)`` with kind "sphere" (size = radius),
    "box" (half extents) or "ellipsoid" (semi-axes, first-order distance);
    metres, node (i, j, k) at (i dx, j dx, k dx).  The SDF is the union's
    (minimum); the blend variable is the smoothed Heaviside over
    ``smoothing_cells * dx`` (a sharp step for 0), and 1/rho and 1/(rho c^2)
    are linear in it between the obstacle material and the medium
    (``rho0``/``c0``, or the per-node ``rho_medium``/``c_medium`` arrays).
    Returns host (nx, ny, nz) arrays (rho, c), and the SDF when asked.
    """
    import ctypes
    import torch
    from . import _capi
    shape = tuple(int(v) for v in shape)
    if len(primitives) > 16:
        raise ValueError("at most 16 obstacle primitives")
    width = float(smoothing_cells)*dx
    if not np.isfinite(width) or width < 0:
        raise ValueError("smoothing_width_cells must be finite and >= 0")
    prims = (_capi.Primitive*max(1, len(primitives)))()
    for q, (kind, centre, size) in enumerate(primitives):
        if kind not in PRIMITIVES:
            raise ValueError(f"unknown obstacle primitive {kind!r}")
        prims[q].kind = PRIMITIVES[kind]
        size = np.broadcast_to(np.asarray(size, float), (3,)) if kind == "sphere" else np.asarray(size, float)
        for e in range(3):
            prims[q].center[e] = float(centre[e])
            prims[q].size[e] = float(size[e])
    dev = torch.device("cuda", torch.cuda.current_device())
    fortran = shape[::-1]  # a C-contiguous (nz, ny, nx) tensor holds a Fortran-order grid

    def field(a):
        if a is None or np.ndim(a) == 0:
            return None
        a = np.broadcast_to(np.asarray(a, float), shape)
        return torch.as_tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device=dev)
    fr, fc = field(rho_medium), field(c_medium)
    if rho_medium is not None and np.ndim(rho_medium) == 0:
        rho0 = float(rho_medium)
    if c_medium is not None and np.ndim(c_medium) == 0:
        c0 = float(c_medium)
    rho = torch.empty(fortran, dtype=torch.float64, device=dev)
    c = torch.empty_like(rho)
    sdf = torch.empty_like(rho) if want_sdf else None
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    _capi.check(_capi.lib().hz_rasterize_materials(
        shape[0], shape[1], shape[2], float(dx), len(primitives), prims, width, float(rho_obstacle),
        float(c_obstacle), float(rho0), float(c0), ptr(fr), ptr(fc), ptr(rho), ptr(c), ptr(sdf),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "hz_rasterize_materials")
    out = [t.cpu().numpy().transpose(2, 1, 0) for t in (rho, c) + ((sdf,) if want_sdf else ())]
    return tuple(out)


# ------------------------------------------------------------- transducers
def ideal_grid_centers(array_n, pitch, lx, ly, origin=(0.0, 0.0, 0.0)):
    origin = np.asarray(origin, dtype=float)
    cx, cy = origin[:2] + np.asarray([lx, ly])*0.5
    half = (array_n - 1)*pitch*0.5
    steps = np.arange(array_n)*pitch
    gx, gy = np.meshgrid(cx - half + steps, cy - half + steps, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)


def geometric_phases(centers, targets, wavelength, k0):
    if len(targets) == 0:
        return np.zeros(len(centers))
    if len(targets) == 1:
        cyc = np.linalg.norm(centers - np.asarray(targets[0], float), axis=-1)/wavelength
        return 2.0*np.pi*(np.ceil(cyc) - cyc)
    field = np.zeros(len(centers), dtype=np.complex128)
    for t in targets:
        d = np.linalg.norm(centers - np.asarray(t, float), axis=-1)
        field += (1.0/d)*np.exp(-1j*k0*d)
    return np.angle(field)


# ------------------------------------------------------------------ sources
def aperture_quadrature(diameter, radial_order=32, angular_order=128):
    if not np.isfinite(diameter) or diameter <= 0:
        raise ValueError("Aperture diameter must be positive")
    if radial_order < 2 or angular_order < 8:
        raise ValueError("Aperture quadrature orders are too small")
    x, w = np.polynomial.legendre.leggauss(radial_order)
    mu, w = (x + 1)/2, w/2
    rad = diameter/2*np.sqrt(1 - mu*mu)
    ang = 2*np.pi*np.arange(angular_order)/angular_order
    offsets = np.column_stack(((rad[:, None]*np.cos(ang)).ravel(), (rad[:, None]*np.sin(ang)).ravel(),
                               np.zeros(radial_order*angular_order)))
    return offsets, np.repeat(w/angular_order, angular_order)


def source_deposition(shape_xy, dx, centers, diameter, radial_order=64, angular_order=256):
    """Unit-mass bilinear deposition of every aperture on the z = 0 plane.

    Returns (node, column, mass) triplets, node = i + nx*j, sorted by column
    then node; each column sums to 1 (sinc_aperture.py:36-72).  Validation,
    quadrature and each emitter's node box are host work; the masses come
    from the device (``hz_deposit_sources``), one sequential sum per node in
    the reference's order, so they are bit-identical to the reference's.
    """
    import ctypes
    from . import _capi
    nx, ny = (int(v) for v in shape_xy)
    centers = np.asarray(centers, float)
    if (min(nx, ny) < 3 or not np.isfinite(dx) or dx <= 0 or centers.ndim != 2
            or centers.shape[1] != 3 or not len(centers) or not np.isfinite(centers).all()
            or np.any(centers[:, 2] != 0)):
        raise ValueError("Sources require finite centers on z=0 and a valid xy grid")
    offsets, weights = aperture_quadrature(diameter, radial_order, angular_order)
    extent = np.array([nx - 1, ny - 1])*dx
    if np.any(centers[:, :2] - diameter/2 < dx) or np.any(centers[:, :2] + diameter/2 > extent - dx):
        raise ValueError("Sinc aperture needs one grid cell clearance from side boundaries")
    m = len(centers)
    # Node box of each aperture, one cell of margin on each side: every
    # corner node of every quadrature point lies inside.
    r = diameter/2
    lo = np.maximum(np.floor((centers[:, :2] - r)/dx).astype(np.int64) - 1, 0)
    hi = np.minimum(np.floor((centers[:, :2] + r)/dx).astype(np.int64) + 2, np.array([nx - 1, ny - 1]))
    size = hi - lo + 1
    boxes = np.ascontiguousarray(np.column_stack((lo, size)), dtype=np.int32)
    cap = int((size[:, 0]*size[:, 1]).max())
    mass = np.empty((m, cap))
    xy = np.ascontiguousarray(centers[:, :2])
    off = np.ascontiguousarray(offsets[:, :2])
    w = np.ascontiguousarray(weights, dtype=np.float64)
    dptr = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _capi.check(_capi.lib().hz_deposit_sources(m, dptr(xy), len(w), dptr(off), dptr(w), float(dx),
                                               boxes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap,
                                               dptr(mass)), "hz_deposit_sources")
    t = np.arange(cap)
    width = size[:, 0]
    keep = (t[None, :] < (size[:, 0]*size[:, 1])[:, None]) & (mass != 0)
    col, tt = np.nonzero(keep)
    node = (lo[col, 0] + tt % width[col]) + nx*(lo[col, 1] + tt//width[col])
    return node.astype(np.int64), col.astype(np.int32), mass[col, tt]


def flux_scale(dx, reference_density):
    if not np.isfinite(reference_density) or reference_density <= 0:
        raise ValueError("reference_density must be positive")
    return -2.0*np.pi/(reference_density*dx**3)
