"""Pins the surface-condensation oracle (oracle/surface.py) to what the
mathematics fixes: the condensed solve is the full 3-D finite-volume solve of
a homogeneous medium over a reflecting -z baffle.

* every z-plane of solve_flux equals the sparse direct solve of the 3-D
  operator (oracle.helmholtz, pinned by the reference's CSR fixtures) with
  the flux right-hand side (oracle.source), to 1e-9 relative;
* each mode's transfer ratios H[:, m] solve rows 1..nz-1 of its tridiagonal
  z system with u[0] = 1;
* flux_target_response(points) @ (a e^{i phi}) equals solve_flux + sample;
* transform and inverse_transform are inverse to each other.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import helmholtz as ohz, source as osrc, surface as osu

SHAPE = (22, 20, 15)
DX = 1.0e-3
FREQ = 40e3
C0, RHO0 = 343.0, 1.2
BOUNDARIES = {"-x": "open", "+x": "open", "-y": {"type": "impedance", "impedance": [823.2, 205.8]},
              "+y": "reflecting", "-z": "reflecting", "+z": "open"}
PML = {"cells": 3, "strength": 2.0, "order": 3}
CENTERS = np.array([[8.5e-3, 9.2e-3, 0.0], [13.1e-3, 10.4e-3, 0.0]])
DIAMETER = 4.2e-3


def masses():
    return osrc.source_mass_matrix(SHAPE[:2], DX, CENTERS, DIAMETER, 8, 16)


def direct_field(phases, amps):
    a = ohz.OracleOperator(np.full(SHAPE, RHO0), np.full(SHAPE, C0), DX, 2*np.pi*FREQ, BOUNDARIES, PML).matrix()
    b = osrc.rhs_for_phases(osrc.flux_rhs_columns(masses(), SHAPE, DX, RHO0), amps, phases)
    return spla.spsolve(a.tocsc(), b).reshape(SHAPE, order="F")


@pytest.fixture(scope="module")
def solver():
    return osu.SurfaceSolver(SHAPE, DX, FREQ, C0, RHO0, BOUNDARIES, PML)


def test_condensed_planes_equal_the_3d_direct_solve(solver):
    phases, amps = np.array([0.3, -1.9]), np.array([1.0, 0.7])
    u = direct_field(phases, amps)
    solver.solve_flux(masses(), phases, amps)
    scale = np.abs(u).max()
    for z in range(SHAPE[2]):
        assert np.abs(solver.plane(z) - u[:, :, z]).max() <= 1e-9*scale, z


def test_transfer_ratios_solve_each_modes_z_system(solver):
    z = solver.axes[2]
    nz = z.shape[0]
    for a, b in ((0, 0), (5, 11), (21, 19), (13, 2)):
        t = z + solver.shift[a, b]*np.eye(nz)
        u = np.linalg.solve(t[1:, 1:], -t[1:, 0])
        np.testing.assert_allclose(solver.h[:, a, b], u, rtol=1e-11, atol=1e-13*np.abs(u).max())
    # Row 0 closes the system: the Schur complement times u[0] is the surface load.
    a, b = 7, 9
    t = z + solver.shift[a, b]*np.eye(nz)
    u = np.concatenate(([1.0], solver.h[:, a, b]))
    assert abs(t[0] @ u - solver.schur[a, b]) <= 1e-12*abs(solver.schur[a, b])


def test_target_response_is_the_sampled_field(solver):
    m = masses()
    phases, amps = np.array([1.1, 2.5]), np.array([0.9, 1.3])
    pts = np.array([[10.3e-3, 9.9e-3, 6.2e-3], [4.0e-3, 15.5e-3, 0.0], [12.0e-3, 7.0e-3, 11.75e-3]])
    resp = solver.flux_target_response(m, pts)
    solver.solve_flux(m, phases, amps)
    direct = solver.sample(pts)
    np.testing.assert_allclose(resp @ (amps*np.exp(1j*phases)), direct, rtol=1e-10)
    # And the trilinear sample of the 3-D direct solve.
    u = direct_field(phases, amps)
    low, frac = solver.sample_coordinates(pts)
    for q in range(len(pts)):
        v = 0
        for dx_ in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = ((frac[q, 0] if dx_ else 1 - frac[q, 0])*(frac[q, 1] if dy else 1 - frac[q, 1])
                         * (frac[q, 2] if dz else 1 - frac[q, 2]))
                    v += w*u[low[q, 0] + dx_, low[q, 1] + dy, low[q, 2] + dz]
        assert abs(direct[q] - v) <= 1e-9*np.abs(u).max()


def test_modal_transforms_are_inverse(solver):
    rng = np.random.default_rng(2)
    p = rng.standard_normal(SHAPE[:2]) + 1j*rng.standard_normal(SHAPE[:2])
    np.testing.assert_allclose(solver.inverse_transform(solver.transform(p)), p, rtol=0, atol=1e-10*np.abs(p).max())
