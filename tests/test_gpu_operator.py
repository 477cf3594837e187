"""HIP operator and GMRES against the CPU oracle (runs on an MI355X).

Tolerances: complex128 applies match the oracle CSR SpMV to 1e-13
(relative, max-norm over each vector); complex64 applies to 2e-6; GMRES
solutions match the complex128 sparse-LU solution to 1e-5 relative L2 (the
path's field tolerance) at the requested residuals.
"""
import json

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import helmholtz as ohz, krylov as okr, losses as olo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SCENES = ["mixed_faces", "open_pml_all", "soft_impedance_corner", "reflecting_box"]


@pytest.fixture(scope="module")
def fix(golden_dir):
    f = np.load(golden_dir/"operator_fixtures.npz")
    return f, json.loads(str(f["meta"]))


def _op(f, meta, name):
    from ultrasound_solver_amd.operator import HelmholtzOperator
    s = meta["scenes"][name]
    op = HelmholtzOperator(f[f"{name}/rho"], f[f"{name}/c"], meta["dx"], meta["omega"], s["boundaries"],
                           s["pml"], device="cuda:0")
    a = sp.csr_matrix((f[f"{name}/data"], f[f"{name}/indices"], f[f"{name}/indptr"]), shape=(op.n, op.n))
    return op, a


def _rand(n, k, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, k)) + 1j*rng.standard_normal((n, k))


def _rel(got, ref):
    return np.max(np.abs(got - ref))/np.max(np.abs(ref))


@pytest.mark.parametrize("name", SCENES)
def test_apply_c128_matches_reference_matrix(fix, name):
    f, meta = fix
    op, a = _op(f, meta, name)
    x = _rand(op.n, 3, 1)
    scale = okr.row_scale(a)
    minv = okr.jacobi_inverse(a, scale)
    ah = okr.adjoint_matrix(a)
    scale_h = okr.row_scale(ah)
    minv_h = okr.jacobi_inverse(ah, scale_h)
    refs = {"A": a @ x, "AH": ah @ x,
            "RAM": scale[:, None]*(a @ (minv[:, None]*x)),
            "RAMH": scale_h[:, None]*(ah @ (minv_h[:, None]*x)),
            "RA": scale[:, None]*(a @ x), "RAH": scale_h[:, None]*(ah @ x)}
    for mode, ref in refs.items():
        got = op.apply(x, mode=mode).cpu().numpy().T
        assert _rel(got, ref) < 1e-13, mode


@pytest.mark.parametrize("name", SCENES)
def test_apply_c64_matches_reference_matrix(fix, name):
    f, meta = fix
    op, a = _op(f, meta, name)
    x = _rand(op.n, 2, 2)
    scale = okr.row_scale(a)
    minv = okr.jacobi_inverse(a, scale)
    ref = scale[:, None]*(a @ (minv[:, None]*x))
    got = op.apply(x, mode="RAM", precision="c64").cpu().numpy().T
    assert _rel(got, ref) < 2e-6
    got = op.apply(x, mode="A", precision="c64").cpu().numpy().T
    assert _rel(got, a @ x) < 2e-6
    got = op.apply(x, mode="RA", precision="c64").cpu().numpy().T
    assert _rel(got, scale[:, None]*(a @ x)) < 2e-6
    # The adjoint modes against the conjugate transpose of the same matrix.
    ah = okr.adjoint_matrix(a)
    scale_h = okr.row_scale(ah)
    minv_h = okr.jacobi_inverse(ah, scale_h)
    refs = {"AH": ah @ x, "RAMH": scale_h[:, None]*(ah @ (minv_h[:, None]*x)), "RAH": scale_h[:, None]*(ah @ x)}
    for mode, ref in refs.items():
        got = op.apply(x, mode=mode, precision="c64").cpu().numpy().T
        assert _rel(got, ref) < 2e-6, mode


def _scene(n, seed=0, soft=False):
    rng = np.random.default_rng(seed)
    rho = 1.2*(1 + 0.1*rng.random((n, n, n)))
    c = 343.0*(1 + 0.03*rng.standard_normal((n, n, n)))
    bnd = {"-x": "open", "+x": "reflecting", "-y": "open", "+y": "open", "+z": "open"}
    if soft:
        bnd["+x"] = "soft"
    pml = {"cells": 3}
    return rho, c, bnd, pml


def _solver_pair(n, seed=0, soft=False):
    from ultrasound_solver_amd.operator import HelmholtzOperator
    rho, c, bnd, pml = _scene(n, seed, soft)
    dx, omega = 1.07e-3, 2*np.pi*40e3
    op = HelmholtzOperator(rho, c, dx, omega, bnd, pml, device="cuda:0")
    a = ohz.OracleOperator(rho, c, dx, omega, bnd, pml).matrix()
    return op, a


def _point_rhs(n_total, k, seed):
    rng = np.random.default_rng(seed)
    b = np.zeros((n_total, k), dtype=complex)
    for r in range(k):
        b[rng.integers(0, n_total, 3), r] = rng.standard_normal(3) + 1j*rng.standard_normal(3)
    return b


def test_plain_gmres_follows_reference_semantics():
    op, a = _solver_pair(12)
    b = _point_rhs(op.n, 1, 3)[:, 0]
    x, st = op.gmres(b, restart=30, maxiter=3000, rtol=1e-8, precision="c128")
    ref_solver = okr.ReferenceGmres(restart=30, maxiter=3000, rtol=1e-8)
    ref_x = ref_solver.solve(a, b)
    direct = okr.direct_solve(a, b)
    got = x.cpu().numpy()[0]
    assert st[0]["converged"] == 1 and st[0]["info"] == 0
    # Same algorithm in FP64: same number of restart cycles.
    assert st[0]["iterations"] == ref_solver.last_iterations
    assert np.linalg.norm(got - ref_x)/np.linalg.norm(ref_x) < 1e-6
    assert np.linalg.norm(got - direct)/np.linalg.norm(direct) < 1e-5
    assert st[0]["scaled_residual"] <= 1e-8
    assert abs(st[0]["scaled_residual"] - okr.scaled_residual(a, got, b)) < 1e-3*1e-8
    assert abs(st[0]["backward_error"] - okr.backward_error(a, got, b)) <= 1e-6*okr.backward_error(a, got, b) + 1e-18


def test_plain_gmres_batched_columns_match_single():
    op, a = _solver_pair(10, seed=4)
    b = _point_rhs(op.n, 3, 5)
    xb, stb = op.gmres(b, rtol=1e-8, chunk=3)
    direct = okr.direct_solve(a, b)
    for r in range(3):
        xs, sts = op.gmres(b[:, r], rtol=1e-8)
        np.testing.assert_allclose(xb[r].cpu().numpy(), xs[0].cpu().numpy(), rtol=0, atol=1e-12*np.abs(direct[:, r]).max())
        assert stb[r]["iterations"] == sts[0]["iterations"]


@pytest.mark.parametrize("soft", [False, True])
def test_mixed_gmres_refines_to_fp64_accuracy(soft):
    op, a = _solver_pair(14, seed=1, soft=soft)
    b = _point_rhs(op.n, 4, 7)
    x, st = op.gmres(b, restart=30, maxiter=6000, rtol=1e-8, precision="c64", chunk=4)
    direct = okr.direct_solve(a, b)
    for r in range(4):
        got = x[r].cpu().numpy()
        assert st[r]["converged"] == 1, st[r]
        assert st[r]["refinements"] >= 1
        assert np.linalg.norm(got - direct[:, r])/np.linalg.norm(direct[:, r]) < 1e-5
        assert okr.scaled_residual(a, got, b[:, r]) <= 1e-8


@pytest.mark.parametrize("preconditioner", ["jacobi", "multigrid"])
def test_mixed_gmres_relative_residual_target(preconditioner):
    """relative_rtol: the mixed form refines until ||b - A x||/||b|| meets it
    as well (FP64, recomputed here on the oracle matrix); a target the
    refinement budget cannot reach is reported as not converged."""
    op, a = _solver_pair(14, seed=3)
    if preconditioner == "multigrid":
        op.set_multigrid()
    b = _point_rhs(op.n, 3, 11)
    x, st = op.gmres(b, restart=10, maxiter=6000, rtol=1e-6, precision="c64", chunk=3,
                     preconditioner=preconditioner, relative_rtol=1e-10)
    for r in range(3):
        got = x[r].cpu().numpy()
        rel = np.linalg.norm(b[:, r] - a @ got)/np.linalg.norm(b[:, r])
        assert st[r]["converged"] == 1, st[r]
        assert rel <= 1e-10
        assert st[r]["relative_residual"] == pytest.approx(rel, rel=1e-6, abs=1e-16)
        assert okr.scaled_residual(a, got, b[:, r]) <= 1e-6
    _, st = op.gmres(b[:, :1], restart=10, maxiter=6000, rtol=1e-6, precision="c64", chunk=1,
                     preconditioner=preconditioner, relative_rtol=1e-10, max_refinements=0)
    assert st[0]["converged"] == 0 and st[0]["relative_residual"] > 1e-10


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_adjoint_solve_matches_direct(precision):
    op, a = _solver_pair(12, seed=2, soft=True)
    c = _point_rhs(op.n, 2, 9)
    lam, st = op.gmres(c, rtol=1e-8, precision=precision, adjoint=True, chunk=2, maxiter=6000)
    direct = okr.direct_solve(okr.adjoint_matrix(a), c)
    for r in range(2):
        got = lam[r].cpu().numpy()
        assert st[r]["converged"] == 1, st[r]
        assert np.linalg.norm(got - direct[:, r])/np.linalg.norm(direct[:, r]) < 1e-5


def test_zero_rhs_and_nonconvergence_report():
    op, a = _solver_pair(10)
    x, st = op.gmres(np.zeros(op.n), rtol=1e-8)
    assert st[0]["converged"] == 1 and st[0]["iterations"] == 0
    assert torch.count_nonzero(x).item() == 0
    b = _point_rhs(op.n, 1, 1)[:, 0]
    with pytest.raises(RuntimeError, match="info="):
        op.gmres(b, restart=2, maxiter=2, rtol=1e-12, raise_on_failure=True)


def test_sources_scatter_gather_and_sampling():
    from ultrasound_solver_amd import physics
    from oracle import source as osrc
    n, dx = 24, 1.07e-3
    op, a = _solver_pair(n)
    centers = osrc.ideal_grid_centers(2, 6e-3, (n - 1)*dx, (n - 1)*dx)
    node, col, mass = physics.source_deposition((n, n), dx, centers, 4e-3, 16, 64)
    scale = physics.flux_scale(dx, 1.21)
    op.set_sources(node, col, scale*mass, len(centers))
    masses = osrc.source_mass_matrix((n, n), dx, centers, 4e-3, 16, 64)
    bref = osrc.flux_rhs_columns(masses, (n, n, n), dx, 1.21)
    rng = np.random.default_rng(3)
    w = rng.standard_normal((2, 4)) + 1j*rng.standard_normal((2, 4))
    got = op.build_rhs(w).cpu().numpy()
    ref = (bref @ w.T).T
    assert _rel(got, ref) < 1e-14
    lam = _rand(op.n, 2, 4)
    g = op.gather_sources(lam).cpu().numpy()
    assert _rel(g, (bref.T @ lam).T) < 1e-13
    pts = np.array([[3.3e-3, 7.1e-3, 5.5e-3], [(n - 1)*dx, 0.0, 2.0e-3]])
    rows, wts = olo.trilinear_rows(pts, (n, n, n), (0, 0, 0), dx)
    s = op.sample(rows, wts, lam).cpu().numpy()
    ref_s = olo.target_response(lam, pts, (n, n, n), (0, 0, 0), dx).T
    assert _rel(s, ref_s) < 1e-14


@pytest.mark.parametrize("adjoint,soft", [(False, False), (False, True), (True, True)])
def test_multigrid_mixed_gmres_matches_direct(adjoint, soft):
    """Shifted-Laplacian multigrid preconditioning changes the Krylov path, not
    the answer: solutions match sparse LU to the field tolerance."""
    from ultrasound_solver_amd.operator import HelmholtzOperator
    n = 21
    rho, c, bnd, pml = _scene(n, seed=3, soft=soft)
    bnd["+y"] = {"type": "impedance", "impedance": [823.2, 205.8]}
    dx, omega = 1.07e-3, 2*np.pi*40e3
    op = HelmholtzOperator(rho, c, dx, omega, bnd, pml, device="cuda:0")
    op.set_multigrid()
    a = ohz.OracleOperator(rho, c, dx, omega, bnd, pml).matrix()
    b = _point_rhs(op.n, 3, 11)
    x, st = op.gmres(b, rtol=1e-8, precision="c64", chunk=3, maxiter=4000, adjoint=adjoint,
                     preconditioner="multigrid")
    mat = okr.adjoint_matrix(a) if adjoint else a
    direct = okr.direct_solve(mat, b)
    for r in range(3):
        got = x[r].cpu().numpy()
        assert st[r]["converged"] == 1, st[r]
        assert np.linalg.norm(got - direct[:, r])/np.linalg.norm(direct[:, r]) < 1e-5


def test_multigrid_needs_fewer_iterations_than_jacobi():
    from ultrasound_solver_amd.operator import HelmholtzOperator
    n = 32
    rho, c, bnd, pml = _scene(n, seed=5)
    pml = {"cells": 6}
    op = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml, device="cuda:0")
    op.set_multigrid()
    b = _point_rhs(op.n, 2, 13)
    _, sj = op.gmres(b, rtol=1e-6, precision="c64", chunk=2, maxiter=20000)
    _, sm = op.gmres(b, rtol=1e-6, precision="c64", chunk=2, maxiter=20000, preconditioner="multigrid")
    for r in range(2):
        assert sj[r]["converged"] and sm[r]["converged"]
        assert sm[r]["iterations"]*3 < sj[r]["iterations"], (sm[r], sj[r])


def test_tile_shapes_are_bitwise_identical_and_autotune_picks_one():
    """Every slab tile shape evaluates the same arithmetic per node: applies
    and multigrid-preconditioned solves agree bit for bit across shapes.  The
    autotuner returns valid shapes and a chunk width within its bounds."""
    from ultrasound_solver_amd.operator import HelmholtzOperator
    n = 37  # odd sizes leave partial tiles in x, y and z
    rho, c, bnd, pml = _scene(n, seed=7, soft=True)
    bnd["+y"] = {"type": "impedance", "impedance": [823.2, 205.8]}
    op = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml, device="cuda:0")
    op.set_multigrid()
    nshape = 1 + 3*len(op.hierarchy.levels)
    x = torch.as_tensor(_rand(op.n, 5, 21).T.copy(), device="cuda:0").to(torch.complex64)
    b = _point_rhs(op.n, 3, 17)
    ref_apply, ref_x, ref_it = None, None, None
    nvar = len(HelmholtzOperator.SLAB_SHAPES)
    for shape in range(nvar):
        op.set_tile_shapes([shape]*nshape)
        ys = [op.apply(x, mode=m, precision="c64") for m in ("A", "RAM", "AH", "RAMH")]
        sol, st = op.gmres(b, rtol=1e-8, precision="c64", chunk=3, maxiter=4000, preconditioner="multigrid")
        its = [s["iterations"] for s in st]
        if ref_apply is None:
            ref_apply, ref_x, ref_it = ys, sol, its
            continue
        for y, y0 in zip(ys, ref_apply):
            assert torch.equal(y, y0), f"shape {shape}: apply differs"
        assert its == ref_it, (shape, its, ref_it)
        assert torch.equal(sol, ref_x), f"shape {shape}: solution differs"
    tune = op.autotune(64, max_chunk=32, restart=30, preconditioner="multigrid")
    assert tune["chunk"] in (4, 8, 16, 32)
    assert len(tune["shapes"]) == nshape and all(0 <= v < nvar for v in tune["shapes"])
    sol, st = op.gmres(b, rtol=1e-8, precision="c64", chunk=3, maxiter=4000, preconditioner="multigrid")
    assert torch.equal(sol, ref_x)


def test_tile_shapes_bitwise_identical_on_a_wide_even_grid():
    """An even width above two 128-node tile rows: the paired tiles (two nodes
    per lane, neighbours across tile edges loaded by lanes 0 and 63, a partial
    last tile) against every other shape, for all six apply modes and a
    multigrid-preconditioned solve (whose finest level is even, the coarser
    ones odd)."""
    from ultrasound_solver_amd.operator import HelmholtzOperator
    rng = np.random.default_rng(11)
    shape3 = (260, 9, 8)
    rho = 1.2*(1 + 0.1*rng.random(shape3))
    c = 343.0*(1 + 0.03*rng.standard_normal(shape3))
    bnd = {"-x": "open", "+x": "soft", "-y": "open", "+y": "reflecting", "-z": "reflecting", "+z": "open"}
    op = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, {"cells": 3}, device="cuda:0")
    op.set_multigrid()
    nshape = 1 + 3*len(op.hierarchy.levels)
    x = torch.as_tensor(_rand(op.n, 3, 5).T.copy(), device="cuda:0").to(torch.complex64)
    b = _point_rhs(op.n, 2, 9)
    modes = ("A", "RAM", "AH", "RAMH", "RA", "RAH")
    ref = None
    for shape in range(len(HelmholtzOperator.SLAB_SHAPES)):
        op.set_tile_shapes([shape]*nshape)
        ys = [op.apply(x, mode=m, precision="c64") for m in modes]
        sol, st = op.gmres(b, rtol=1e-6, precision="c64", chunk=2, maxiter=3000, preconditioner="multigrid")
        its = [s["iterations"] for s in st]
        if ref is None:
            ref = (ys, sol, its)
            continue
        for m, y, y0 in zip(modes, ys, ref[0]):
            assert torch.equal(y, y0), f"shape {shape}, mode {m}: apply differs"
        assert its == ref[2], (shape, its, ref[2])
        assert torch.equal(sol, ref[1]), f"shape {shape}: solution differs"


@pytest.mark.parametrize("shape3", [(64, 32, 24), (256, 16, 16)])
def test_xcd_block_tiles_bitwise_identical(shape3):
    """Grids whose tile counts divide into XCD blocks (the 8 XCDs take 4 x 2
    blocks of the tile grid in y and z for most rectangular and paired shapes
    and for the FP64 slab, with one tile column in x at 64 x 32 x 24 and two
    to four at 256 x 16 x 16; flat and z-fastest shapes keep contiguous
    ranges): every shape's applies and multigrid solve equal those of a flat
    shape bit for bit, and the FP64 applies match the oracle's matrix."""
    from ultrasound_solver_amd.operator import HelmholtzOperator
    rng = np.random.default_rng(13)
    rho = 1.2*(1 + 0.1*rng.random(shape3))
    c = 343.0*(1 + 0.03*rng.standard_normal(shape3))
    bnd = {"-x": "open", "+x": "soft", "-y": "open", "+y": {"type": "impedance", "impedance": [823.2, 205.8]},
           "-z": "reflecting", "+z": "open"}
    dx, omega = 1.07e-3, 2*np.pi*40e3
    op = HelmholtzOperator(rho, c, dx, omega, bnd, {"cells": 3}, device="cuda:0")
    op.set_multigrid()
    nshape = 1 + 3*len(op.hierarchy.levels)
    x = torch.as_tensor(_rand(op.n, 3, 23).T.copy(), device="cuda:0").to(torch.complex64)
    b = _point_rhs(op.n, 2, 29)
    modes = ("A", "RAM", "AH", "RAMH", "RA", "RAH")
    flat = 8  # flat tiles: never XCD blocks
    assert HelmholtzOperator.SLAB_SHAPES[flat][4] == 1
    ref = None
    for shape in [flat] + [s for s in range(len(HelmholtzOperator.SLAB_SHAPES)) if s != flat]:
        op.set_tile_shapes([shape]*nshape)
        ys = [op.apply(x, mode=m, precision="c64") for m in modes]
        sol, st = op.gmres(b, rtol=1e-6, precision="c64", chunk=2, maxiter=3000, preconditioner="multigrid")
        its = [s["iterations"] for s in st]
        if ref is None:
            ref = (ys, sol, its)
            continue
        for m, y, y0 in zip(modes, ys, ref[0]):
            assert torch.equal(y, y0), f"shape {shape}, mode {m}: apply differs"
        assert its == ref[2], (shape, its, ref[2])
        assert torch.equal(sol, ref[1]), f"shape {shape}: solution differs"
    a = ohz.OracleOperator(rho, c, dx, omega, bnd, {"cells": 3}).matrix()
    ah = okr.adjoint_matrix(a)
    x128 = _rand(op.n, 3, 31)
    xt = torch.as_tensor(x128.T.copy(), device="cuda:0")
    for mode, m in (("A", a), ("AH", ah)):
        got = op.apply(xt, mode=mode, precision="c128").cpu().numpy().T
        assert _rel(got, m @ x128) <= 1e-13, mode


@pytest.mark.parametrize("precision,pre", [("c128", "jacobi"), ("c64", "multigrid")])
def test_graph_replay_matches_eager(precision, pre):
    """A captured cycle body replayed once per restart cycle (eager cycle-top
    work and host synchronisation between replays) gives bit-identical
    iterates and statistics to the eager launches; a second solve reuses the
    cached graph."""
    from ultrasound_solver_amd.operator import HelmholtzOperator
    n = 24
    rho, c, bnd, pml = _scene(n, seed=9, soft=True)
    op = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml, device="cuda:0")
    if pre == "multigrid":
        op.set_multigrid()
    b = _point_rhs(op.n, 3, 19)
    kw = dict(restart=10, rtol=1e-8, precision=precision, chunk=3, maxiter=20000, preconditioner=pre)
    x0, s0 = op.gmres(b, **kw)
    assert all(s["cycles"] > 2 for s in s0), s0  # several replays
    for _ in range(2):
        x1, s1 = op.gmres(b, use_graph=True, **kw)
        assert torch.equal(x1, x0)
        assert s1 == s0


def test_scene_batch_matches_sequential_solves():
    """hz_gmres_batch runs uneven scenes concurrently on worker streams; each
    scene's result and statistics equal its own hz_gmres call bit for bit,
    with and without graph replay.  A failing scene is named in the error."""
    from ultrasound_solver_amd.operator import HelmholtzOperator, gmres_batch
    ops, bs = [], []
    for n, seed in ((17, 1), (23, 2), (29, 3)):
        rho, c, bnd, pml = _scene(n, seed=seed, soft=(seed == 2))
        bnd["+y"] = {"type": "impedance", "impedance": [823.2, 205.8]}
        op = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml, device="cuda:0")
        op.set_multigrid()
        ops.append(op)
        bs.append(_point_rhs(op.n, 2 + seed, 30 + seed))
    kw = dict(rtol=1e-8, precision="c64", chunk=3, maxiter=4000, preconditioner="multigrid")
    seq = [op.gmres(b, **kw) for op, b in zip(ops, bs)]
    for use_graph in (False, True):
        out = gmres_batch(ops, bs, workers=3, use_graph=use_graph, **kw)
        for (x, st), (x0, st0) in zip(out, seq):
            assert torch.equal(x, x0)
            assert st == st0
    rho, c, bnd, pml = _scene(15, seed=4)
    plain = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml, device="cuda:0")  # no hierarchy
    with pytest.raises(RuntimeError, match="scene 1"):
        gmres_batch([ops[0], plain], [bs[0], _point_rhs(plain.n, 1, 5)], **kw)


@pytest.mark.parametrize("name", SCENES)
def test_material_vjp_matches_reference_fixture(fix, name):
    """hz_material_vjp against the gradients the reference's own
    ConservativeHelmholtz.material_vjp wrote for the fixture scenes (every
    face kind, PML, soft rows), to FP64 rounding; two column pairs sum."""
    f, meta = fix
    op, _ = _op(f, meta, name)
    u = f[f"{name}/vjp_u"].ravel(order="F")
    lam = f[f"{name}/vjp_lambda"].ravel(order="F")
    g_rho, g_c = op.material_vjp(u, lam)
    ref_rho = f[f"{name}/grad_rho"].reshape(op.shape, order="F")
    ref_c = f[f"{name}/grad_c"].reshape(op.shape, order="F")
    assert _rel(g_rho.cpu().numpy(), ref_rho) < 1e-12
    assert _rel(g_c.cpu().numpy(), ref_c) < 1e-12
    # Two pairs give the sum of their gradients.
    u2, lam2 = np.stack([u, 2*u], axis=1), np.stack([lam, 1j*lam], axis=1)
    g2_rho, g2_c = op.material_vjp(u2, lam2)
    h_rho, h_c = op.material_vjp(2*u, 1j*lam)
    np.testing.assert_allclose(g2_rho.cpu().numpy(), (g_rho + h_rho).cpu().numpy(), rtol=0,
                               atol=1e-12*np.abs(ref_rho).max())
    np.testing.assert_allclose(g2_c.cpu().numpy(), (g_c + h_c).cpu().numpy(), rtol=0,
                               atol=1e-12*np.abs(ref_c).max())


@pytest.mark.parametrize("guard", [None, "1"])
@pytest.mark.parametrize("pre", ["jacobi", "multigrid"])
def test_two_step_blocks_keep_one_step_iterations(monkeypatch, pre, guard):
    """Even grid width (the paired 16-byte kernels): two-step blocks take
    within 5 % of the one-step iterations per column and solve to the direct
    solution.  guard "1" (HZ_BLOCK_GUARD) sends every block through the
    measured fix-up, which orthogonalises W2 against W1 with their measured
    inner product instead of the predicted coefficient."""
    from ultrasound_solver_amd.operator import HelmholtzOperator
    if guard is not None:
        monkeypatch.setenv("HZ_BLOCK_GUARD", guard)
    n = 22
    rho, c, bnd, pml = _scene(n, seed=8)
    op = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml, device="cuda:0")
    if pre == "multigrid":
        op.set_multigrid()
    a = ohz.OracleOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml).matrix()
    b = _point_rhs(op.n, 4, 31)
    direct = okr.direct_solve(a, b)
    kw = dict(restart=10, rtol=1e-8, precision="c64", chunk=4, maxiter=20000, preconditioner=pre)
    x2, st2 = op.gmres(b, arnoldi_block=2, **kw)
    x1, st1 = op.gmres(b, arnoldi_block=1, **kw)
    it1, it2 = [s["iterations"] for s in st1], [s["iterations"] for s in st2]
    print(f"{pre} guard {guard}: one-step {it1}, two-step {it2}", flush=True)
    for r in range(4):
        assert st2[r]["converged"] == 1, st2[r]
        got = x2[r].cpu().numpy()
        assert np.linalg.norm(got - direct[:, r])/np.linalg.norm(direct[:, r]) < 1e-5
        assert abs(it2[r] - it1[r]) <= max(2, 0.05*it1[r]), (it1, it2)


@pytest.mark.parametrize("pre,adjoint", [("jacobi", False), ("multigrid", False), ("multigrid", True)])
def test_two_step_arnoldi_blocks_solve_to_direct(pre, adjoint):
    """Mixed-form GMRES with two basis vectors per Gram-Schmidt pass reaches
    the same solutions (complex128 LU, field tolerance) as one-step Arnoldi,
    with odd and even restart lengths, and replays from a graph unchanged."""
    from ultrasound_solver_amd.operator import HelmholtzOperator
    n = 21
    rho, c, bnd, pml = _scene(n, seed=6, soft=True)
    bnd["+y"] = {"type": "impedance", "impedance": [823.2, 205.8]}
    op = HelmholtzOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml, device="cuda:0")
    if pre == "multigrid":
        op.set_multigrid()
    a = ohz.OracleOperator(rho, c, 1.07e-3, 2*np.pi*40e3, bnd, pml).matrix()
    mat = okr.adjoint_matrix(a) if adjoint else a
    b = _point_rhs(op.n, 3, 23)
    direct = okr.direct_solve(mat, b)
    for restart in (30, 15):
        kw = dict(restart=restart, rtol=1e-8, precision="c64", chunk=3, maxiter=20000, preconditioner=pre,
                  adjoint=adjoint)
        x, st = op.gmres(b, arnoldi_block=2, **kw)
        x1, st1 = op.gmres(b, arnoldi_block=1, **kw)
        for r in range(3):
            assert st[r]["converged"] == 1, st[r]
            got = x[r].cpu().numpy()
            assert np.linalg.norm(got - direct[:, r])/np.linalg.norm(direct[:, r]) < 1e-5
        xg, stg = op.gmres(b, arnoldi_block=2, use_graph=True, **kw)
        assert torch.equal(xg, x) and stg == st
        print(f"restart {restart}: iterations one-step {[s['iterations'] for s in st1]}, "
              f"two-step {[s['iterations'] for s in st]}", flush=True)
