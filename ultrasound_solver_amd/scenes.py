"""Seeded synthetic scenes for the benchmark workloads.

Each builder returns a ``Scene`` (materials, faces, PML, transducer centres,
aperture) in the reference's conventions (h = 1.07 mm, f = 40 kHz, air
rho = 1.2 kg/m^3, c = 343 m/s, sinc aperture on the reflecting z = 0 baffle).

Deviations from the literal workload text, each forced by the reference's own
validity checks (sinc_aperture.py:52-53, warp_utils.py:93-94 and 104-113,
conservative_helmholtz.py:59-61, helmholtz_solver.py:299-302):

* array too wide for the grid: pitch and radius are scaled together by
  s = min(1, (interior - 2h) / span) so the array clears the PML by one cell;
* "PML on all faces": PML only on open faces, -z stays the reflecting baffle;
* normalised impedance z is multiplied by rho0 c0 (the reference takes Pa s/m);
* "rigid" obstacles use the reference default material (rho 1250, c 2200)
  with the 1.2-cell blend; several obstacles use the union SDF min(sdf_i)
  blended into rho/c fields (the material-array route);
* the ellipsoid uses the first-order SDF k0 (k0 - 1) / k1.
"""
from dataclasses import dataclass, field

import numpy as np

from . import physics

H = 1.07e-3
FREQ = 40.0e3
C0 = 343.0
RHO0 = 1.2
OMEGA = 2*np.pi*FREQ


@dataclass
class Scene:
    name: str
    rho: np.ndarray
    c: np.ndarray
    boundaries: dict
    pml: dict
    centers: np.ndarray
    radius: float
    dx: float = H
    omega: float = OMEGA
    rho_ref: float = RHO0
    notes: list = field(default_factory=list)

    @property
    def shape(self):
        return self.rho.shape

    @property
    def num_sources(self):
        return len(self.centers)


def gaussian_random_field(shape, corr_cells, seed, clip=3.0, truncate=4.0):
    """White noise (seeded, host) filtered on the device by a periodic
    Gaussian of `corr_cells` standard deviation, truncated at `truncate`
    of them (the weights and truncation of scipy.ndimage.gaussian_filter,
    mode "wrap"), normalised to unit standard deviation and clipped."""
    import ctypes
    import torch
    from . import _capi
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    radius = int(truncate*float(corr_cells) + 0.5)
    x = np.arange(-radius, radius + 1)
    w = np.exp(-0.5/(corr_cells*corr_cells)*x**2)
    w = np.ascontiguousarray(w/w.sum())
    dev = torch.as_tensor(np.ascontiguousarray(noise.transpose(2, 1, 0)), device=torch.device("cuda"))
    _capi.check(_capi.lib().hz_gaussian_filter(shape[0], shape[1], shape[2], radius,
                                               w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               ctypes.c_void_p(dev.data_ptr()),
                                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                "hz_gaussian_filter")
    g = dev.cpu().numpy().transpose(2, 1, 0)
    g = (g - g.mean())/g.std()
    return np.clip(g, -clip, clip)


def fitted_array(n_grid, array_n, pitch, radius, pml_cells, dx=H):
    """Scale pitch and radius so the array (plus one deposition cell) stays out of the PML."""
    interior = (n_grid - 1 - 2*pml_cells)*dx
    span = (array_n - 1)*pitch + 2*radius
    s = min(1.0, (interior - 2*dx)/span)
    return pitch*s, radius*s, s


def _array_centers(shape, array_n, pitch, dx=H):
    lx, ly = (shape[0] - 1)*dx, (shape[1] - 1)*dx
    return physics.ideal_grid_centers(array_n, pitch, lx, ly)


OPEN5 = {"-x": "open", "+x": "open", "-y": "open", "+y": "open", "-z": "reflecting", "+z": "open"}


def config_small(n=64, array_n=4):
    """64^3 homogeneous air, PML on the open faces, 4x4 array, 10.5 mm pitch, r = 5 mm."""
    shape = (n, n, n)
    pml = {"cells": 8, "strength": 2.0, "order": 3}
    pitch, radius, s = fitted_array(n, array_n, 10.5e-3, 5.0e-3, pml["cells"])
    sc = Scene("homogeneous_air_%d" % n, np.full(shape, RHO0), np.full(shape, C0), dict(OPEN5), pml,
               _array_centers(shape, array_n, pitch), radius)
    if s < 1:
        sc.notes.append(f"array scaled by {s:.3f}: pitch {pitch*1e3:.2f} mm, radius {radius*1e3:.2f} mm")
    return sc


def config_heterogeneous(n=128, array_n=16, seed=1, sigma=0.05, corr=5.0, sphere_cells=15,
                         offset_cells=(10, 0, 20)):
    """128^3, c = 343 (1 + 0.05 g) with a seeded Gaussian random field g (corr 5 cells,
    clipped to +-3 sigma), rigid sphere r = 15 cells at centre + (10, 0, 20) cells,
    PML on -x, +-y, +z, reflecting +x and -z, 16x16 array."""
    shape = (n, n, n)
    g = gaussian_random_field(shape, corr, seed)
    c_med = C0*(1 + sigma*g)
    rho_med = np.full(shape, RHO0)
    centre = np.array([(n - 1)/2 + offset_cells[0], (n - 1)/2 + offset_cells[1],
                       (n - 1)/2 + offset_cells[2]])*H
    rho, c = physics.rasterize_materials(shape, H, [("sphere", centre, sphere_cells*H)], RHO0, C0,
                                         rho_medium=rho_med, c_medium=c_med)
    bnd = {"-x": "open", "+x": "reflecting", "-y": "open", "+y": "open", "-z": "reflecting", "+z": "open"}
    pml = {"cells": 8, "strength": 2.0, "order": 3}
    pitch, radius, s = fitted_array(n, array_n, 10.5e-3, 5.0e-3, pml["cells"])
    sc = Scene("heterogeneous_sphere_%d" % n, rho, c, bnd, pml, _array_centers(shape, array_n, pitch), radius)
    sc.notes.append(f"array scaled by {s:.3f}: pitch {pitch*1e3:.2f} mm, radius {radius*1e3:.2f} mm")
    return sc


def config_two_layer(n=256, array_n=16, semi_axes_cells=(40, 25, 15)):
    """Two-layer medium at z = Lz/2 (rho 1.2 -> 1.0, c 343 -> 360) plus a rigid
    ellipsoid; PML on the open faces."""
    shape = (n, n, n)
    k = np.arange(n)[None, None, :]
    rho_med = np.where(k < n//2, 1.2, 1.0)*np.ones(shape)
    c_med = np.where(k < n//2, 343.0, 360.0)*np.ones(shape)
    centre = np.array([(n - 1)/2]*3)*H
    rho, c = physics.rasterize_materials(shape, H, [("ellipsoid", centre, np.array(semi_axes_cells)*H)], RHO0, C0,
                                         rho_medium=rho_med, c_medium=c_med)
    pml = {"cells": 8, "strength": 2.0, "order": 3}
    pitch, radius, s = fitted_array(n, array_n, 10.5e-3, 5.0e-3, pml["cells"])
    return Scene("two_layer_ellipsoid_%d" % n, rho, c, dict(OPEN5), pml, _array_centers(shape, array_n, pitch),
                 radius)


def config_uneven(index, seed=5, array_n=8):
    """Scene `index` of the uneven batch: edge from {64, 96, 128}, 1-3 sphere/box
    obstacles, c perturbation sigma 0.03, impedance walls (2.0+0.5i) rho0 c0 on +-y."""
    rng = np.random.default_rng([seed, index])
    n = int(rng.choice([64, 96, 128]))
    shape = (n, n, n)
    g = gaussian_random_field(shape, 5.0, seed*1000 + index)
    c_med = C0*(1 + 0.03*g)
    prims = []
    for _ in range(int(rng.integers(1, 4))):
        ctr = rng.uniform(0.3, 0.7, 3)*(n - 1)*H
        ctr[2] = rng.uniform(0.45, 0.8)*(n - 1)*H
        if rng.random() < 0.5:
            prims.append(("sphere", ctr, rng.uniform(6, 16)*H))
        else:
            prims.append(("box", ctr, rng.uniform(4, 12, 3)*H))
    rho, c = physics.rasterize_materials(shape, H, prims, RHO0, C0, rho_medium=RHO0, c_medium=c_med)
    z = [2.0*RHO0*C0, 0.5*RHO0*C0]
    bnd = {"-x": "open", "+x": "open", "-y": {"type": "impedance", "impedance": z},
           "+y": {"type": "impedance", "impedance": z}, "-z": "reflecting", "+z": "open"}
    pml = {"cells": 8, "strength": 2.0, "order": 3}
    pitch, radius, s = fitted_array(n, array_n, 10.5e-3, 5.0e-3, pml["cells"])
    return Scene("uneven_%02d_n%d" % (index, n), rho, c, bnd, pml, _array_centers(shape, array_n, pitch), radius)


def source_triplets(scene, radial_order=64, angular_order=256):
    """(node, column, value) of the flux RHS columns B on the z = 0 plane."""
    node, col, mass = physics.source_deposition(scene.shape[:2], scene.dx, scene.centers, 2*scene.radius,
                                                radial_order, angular_order)
    return node, col, physics.flux_scale(scene.dx, scene.rho_ref)*mass
