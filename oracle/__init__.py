"""CPU FP64 oracle for the gpu_iterative Helmholtz path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker of the HIP path (or as
the timed CPU baseline), never as the thing measured or shipped.

Each function restates the reference algorithm in numpy/scipy and cites the
reference ``file:line`` it follows.  How each part is pinned:

* ``helmholtz``  -- operator A, face terms, PML, soft rows, material VJP:
  golden CSR/VJP fixtures written by the reference's own CPU module
  (``tests/golden/operator_fixtures.npz``), plus closed forms (textbook
  7-point interior rows, constant annihilation, complex symmetry).
* ``source``     -- aperture quadrature, bilinear deposition, flux RHS,
  ideal-grid centres, geometric phases: golden fixtures from the reference
  module (``source_fixtures.npz``), the archived centres/phases
  (``archive_fixture.npz``), unit column mass and centroid preservation.
* ``materials``  -- sphere SDF + smoothed Heaviside + 1/rho, 1/(rho c^2) blend:
  closed forms (end values exact, linear in the blend variable).
* ``krylov``     -- row-scaled Jacobi GMRES with CuPy's restart semantics:
  solution parity against scipy ``splu`` (complex128) and the reference's
  acceptance invariants.  The CuPy iteration trajectory itself is
  parity unpinned (CuPy is not available anywhere in this build).
* ``surface``    -- homogeneous surface condensation (axis operators,
  transfer ratios, Schur complement, modal transforms, target response):
  planes equal the sparse direct solve of ``helmholtz``'s 3-D operator with
  ``source``'s flux RHS; per-mode tridiagonal solves; linearity between the
  target response and the sampled field.
* ``losses``     -- field_match / focal_pressure / focal_contrast, the
  Gaussian target, background mask and trilinear target response: the
  archived run's stored loss values and target fields; the phase VJP by
  central finite differences and the worked example u = 3+4i.
  focal_pressure against the reference's derived closed form
  (docs/GRADIENT_DERIVATION.html:914-930) on a constant field and a field
  with one nonzero node; focal_gradient by a two-emitter closed form; the
  ratio form of focal contrast (no reference code) by its value on
  piecewise-constant fields, scale invariance and central differences.
"""
