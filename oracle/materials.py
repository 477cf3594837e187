"""FP64 restatement of the obstacle material blend.

Follows ``src/algorithm/physics/warp_utils.py``: ``blend_obstacle_material``
(:11-23, 1/rho and 1/(rho c^2) linear in the blend variable h), the sphere
SDF (:58-61) and the smoothed Heaviside h = (1 + t + sin(pi t)/pi)/2,
t = clip(sdf/w, -1, 1), w = 1.2 h by default (:64-72).

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np


def blend(hval, rho_obs, c_obs, rho_med, c_med):
    inv_rho = 1.0/rho_obs + (1.0/rho_med - 1.0/rho_obs)*hval
    kappa_obs = 1.0/(rho_obs*c_obs**2)
    kappa = kappa_obs + (1.0/(rho_med*c_med**2) - kappa_obs)*hval
    return 1.0/inv_rho, np.sqrt(inv_rho/kappa)


def smoothed_heaviside(sdf, width):
    if width:
        t = np.clip(sdf/width, -1.0, 1.0)
        return 0.5*(1.0 + t + np.sin(np.pi*t)/np.pi)
    return (sdf > 0).astype(float)


def sphere_sdf(shape, dx, center, radius):
    i, j, k = np.ogrid[:shape[0], :shape[1], :shape[2]]
    coords = (i, j, k)
    dist = np.sqrt(sum((coords[a]*dx - center[a])**2 for a in range(3)))
    return dist - radius


def sphere_materials(shape, dx, rho0, c0, center, radius, rho_obs=1250.0, c_obs=2200.0,
                     smoothing_cells=1.2):
    sdf = sphere_sdf(shape, dx, center, radius)
    h = smoothed_heaviside(sdf, smoothing_cells*dx)
    return blend(h, rho_obs, c_obs, rho0, c0)
