"""Device-resident Helmholtz operator: a thin Python owner of an hz_operator
handle (``include/hz.h``) with torch tensors as device buffers.

The constructor validates its inputs the way ``ConservativeHelmholtz``
does (src/algorithm/physics/conservative_helmholtz.py:18-103) and builds the
per-axis and face tables on the host; the matrix itself is never formed.
Vectors are torch tensors on a HIP device; k right-hand sides are held as a
(k, N) contiguous tensor (each column of the reference's (N, k) array is a
contiguous row here), N in Fortran node order.
"""
import ctypes

import numpy as np
import torch

from . import _capi, physics

MODES = {"A": _capi.MODE_A, "RAM": _capi.MODE_RAM, "AH": _capi.MODE_AH, "RAMH": _capi.MODE_RAMH,
         "RA": _capi.MODE_RA, "RAH": _capi.MODE_RAH}
PRECISIONS = {"c128": _capi.C128, "c64": _capi.C64}


def _dptr(array):
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _device_tensor(x, dtype, device):
    t = torch.as_tensor(x, device=device)
    if t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def gmres_options(restart=30, maxiter=3000, rtol=1.0e-6, backward_error_tol=1.0e-4, precision="c128",
                  adjoint=False, chunk=1, max_refinements=10, inner_rtol=1.0e-5, preconditioner="jacobi",
                  use_graph=False, arnoldi_block=None, relative_rtol=0.0):
    if arnoldi_block is None:
        arnoldi_block = 2 if precision == "c64" else 1
    return _capi.GmresOpts(restart=int(restart), maxiter=int(maxiter), rtol=float(rtol),
                           backward_error_tol=float(backward_error_tol),
                           precision=PRECISIONS[precision], adjoint=int(bool(adjoint)),
                           chunk=int(chunk), max_refinements=int(max_refinements),
                           inner_rtol=float(inner_rtol), use_graph=int(bool(use_graph)),
                           preconditioner={"jacobi": _capi.PRE_JACOBI,
                                           "multigrid": _capi.PRE_MULTIGRID}[preconditioner],
                           arnoldi_block=int(arnoldi_block), relative_rtol=float(relative_rtol))


def gmres_batch(operators, rhs, workers=4, **options):
    """Solve several scenes concurrently (one hz_gmres_batch call).

    operators: HelmholtzOperator per scene (each at most once); rhs: matching
    right-hand sides ((N,), (N, k) or (k, N) device tensors).  Options as in
    HelmholtzOperator.gmres.  Returns [(x, stats)] in scene order; raises
    RuntimeError naming the scene on an argument or device error."""
    ns = len(operators)
    if ns != len(rhs) or ns == 0:
        raise ValueError("need one right-hand side per operator")
    bs = [op.as_columns(b, torch.complex128) for op, b in zip(operators, rhs)]
    xs = [torch.empty_like(b) for b in bs]
    stats = [(_capi.GmresStats*b.shape[0])() for b in bs]
    opts = gmres_options(**options)
    vp = ctypes.c_void_p
    status = _capi.lib().hz_gmres_batch(
        ns, (vp*ns)(*[op.handle for op in operators]), ctypes.byref(opts),
        (vp*ns)(*[b.data_ptr() for b in bs]), (vp*ns)(*[x.data_ptr() for x in xs]),
        (ctypes.c_int32*ns)(*[b.shape[0] for b in bs]), (ctypes.c_int64*ns)(*[op.n for op in operators]),
        (ctypes.POINTER(_capi.GmresStats)*ns)(*[ctypes.cast(st, ctypes.POINTER(_capi.GmresStats)) for st in stats]),
        int(workers), _stream())
    if status not in (_capi.HZ_OK, _capi.HZ_NOT_CONVERGED):
        _capi.check(status, "hz_gmres_batch")
    return [(x, [s.as_dict() for s in st]) for x, st in zip(xs, stats)]


def launch_counts(reset=False):
    """Solver launches since the last reset: {"graphs", "fused", "single"}
    (graph launches, fused multi-scene kernel launches, single-scene kernel
    launches; launches inside a graph count once, as its graph launch)."""
    out = (ctypes.c_int64*3)()
    _capi.check(_capi.lib().hz_launch_counts(out, int(bool(reset))), "hz_launch_counts")
    return {"graphs": out[0], "fused": out[1], "single": out[2]}


def gmres_fused(operators, rhs, **options):
    """Solve several scenes in lockstep (one hz_gmres_fused call, mixed form).

    Every GMRES cycle runs each of its kernels once for all the scenes still
    iterating: one grid over all their workgroups, each reading its scene's
    arguments from a per-scene table (with use_graph=True each cycle body is one
    graph launch).  Arguments and return value as gmres_batch; the results
    per scene are those of HelmholtzOperator.gmres on the scene alone."""
    ns = len(operators)
    if ns != len(rhs) or ns == 0:
        raise ValueError("need one right-hand side per operator")
    bs = [op.as_columns(b, torch.complex128) for op, b in zip(operators, rhs)]
    xs = [torch.empty_like(b) for b in bs]
    stats = [(_capi.GmresStats*b.shape[0])() for b in bs]
    opts = gmres_options(**options)
    vp = ctypes.c_void_p
    status = _capi.lib().hz_gmres_fused(
        ns, (vp*ns)(*[op.handle for op in operators]), ctypes.byref(opts),
        (vp*ns)(*[b.data_ptr() for b in bs]), (vp*ns)(*[x.data_ptr() for x in xs]),
        (ctypes.c_int32*ns)(*[b.shape[0] for b in bs]), (ctypes.c_int64*ns)(*[op.n for op in operators]),
        (ctypes.POINTER(_capi.GmresStats)*ns)(*[ctypes.cast(st, ctypes.POINTER(_capi.GmresStats)) for st in stats]),
        _stream())
    if status not in (_capi.HZ_OK, _capi.HZ_NOT_CONVERGED):
        _capi.check(status, "hz_gmres_fused")
    return [(x, [s.as_dict() for s in st]) for x, st in zip(xs, stats)]


class HelmholtzOperator:
    """A = div(rho^-1 grad) + omega^2/(rho c^2) with PML and face rows, on the GPU."""

    def __init__(self, rho, c, dx, omega, boundaries, pml=None, device=None):
        rho = np.asarray(rho, dtype=np.float64)
        c = np.asarray(c, dtype=np.float64)
        if rho.ndim != 3 or c.shape != rho.shape or min(rho.shape) < 2:
            raise ValueError("rho and c must have the same 3D grid shape, each dimension >= 2")
        if not (np.isfinite(rho).all() and np.isfinite(c).all() and (rho > 0).all() and (c > 0).all()):
            raise ValueError("density and sound speed must be finite and positive")
        if not (np.isfinite(dx) and dx > 0 and np.isfinite(omega) and omega > 0):
            raise ValueError("dx and omega must be finite and positive")
        if not torch.cuda.is_available():
            raise RuntimeError("HelmholtzOperator needs a HIP device; none is visible")
        self.device = torch.device(device if device is not None else "cuda")
        self.shape = tuple(int(v) for v in rho.shape)
        self.n = int(np.prod(self.shape))
        self.dx, self.omega = float(dx), float(omega)
        self.boundaries = dict(boundaries)
        self.weights, self.stretch = physics.axis_tables(self.shape, self.boundaries, pml)
        self.kinds, self.face_terms = physics.face_tables(rho, c, self.dx, self.omega, self.boundaries,
                                                          self.weights, self.stretch)
        self.fixed = physics.soft_node_mask(self.shape, self.kinds)
        self.pml_mask = physics.pml_node_mask(self.stretch, self.shape)
        self.rho, self.c = rho, c
        self.hierarchy = None
        self.tuning = None
        kappa = 1.0/(rho*c*c)
        keep = []  # host arrays referenced by the descriptor during hz_create

        def hold(a):
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a

        desc = _capi.GridDesc()
        desc.nx, desc.ny, desc.nz = self.shape
        desc.dx, desc.omega = self.dx, self.omega
        desc.rho = _dptr(hold(rho.ravel(order="F")))
        desc.kappa = _dptr(hold(kappa.ravel(order="F")))
        for a in range(3):
            desc.weight[a] = _dptr(hold(self.weights[a].astype(np.float64)))
            desc.stretch[a] = _dptr(hold(self.stretch[a].astype(np.complex128).view(np.float64)))
        for f in range(6):
            desc.face_kind[f] = _capi.FACE_KIND[self.kinds[f]]
            table = self.face_terms[f]
            desc.face_term[f] = (_dptr(hold(table.view(np.float64))) if table is not None
                                 else ctypes.POINTER(ctypes.c_double)())
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _capi.check(_capi.lib().hz_create(ctypes.byref(desc), ctypes.byref(handle)), "hz_create")
        self._h = handle
        self.num_sources = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _capi.lib().hz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("operator is closed")
        return self._h

    def device_bytes(self):
        return int(_capi.lib().hz_device_bytes(self.handle))

    # -- vectors ----------------------------------------------------------
    def as_columns(self, x, dtype=torch.complex128):
        """(N,), (N, k) array-like or (k, N) tensor -> (k, N) contiguous device tensor."""
        if isinstance(x, torch.Tensor) and x.dim() == 2 and x.shape[1] == self.n and x.shape[0] != self.n:
            return _device_tensor(x, dtype, self.device)
        t = torch.as_tensor(x)
        if t.dim() == 1:
            t = t.reshape(1, -1)
        elif t.dim() == 2 and t.shape[0] == self.n:
            t = t.transpose(0, 1)
        else:
            raise ValueError(f"expected ({self.n},) or ({self.n}, k), got {tuple(t.shape)}")
        if t.shape[1] != self.n:
            raise ValueError("vector length does not match the grid")
        return _device_tensor(t, dtype, self.device)

    def apply(self, x, mode="A", precision="c128"):
        """y = op(x) for each column; x is a (k, N) device tensor."""
        dtype = torch.complex128 if precision == "c128" else torch.complex64
        x = self.as_columns(x, dtype)
        y = torch.empty_like(x)
        _capi.check(_capi.lib().hz_apply(self.handle, PRECISIONS[precision], MODES[mode],
                                         ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                         x.shape[0], self.n, _stream()), "hz_apply")
        return y

    def gmres(self, b, restart=30, maxiter=3000, rtol=1.0e-6, backward_error_tol=1.0e-4,
              precision="c128", adjoint=False, chunk=1, max_refinements=10, inner_rtol=1.0e-5,
              preconditioner="jacobi", raise_on_failure=False, use_graph=False, arnoldi_block=None,
              relative_rtol=0.0):
        """Solve A x = b (or A^H x = b) for each column; returns (x, stats).
        arnoldi_block: basis vectors per Gram-Schmidt pass, 2 by default in
        the mixed form (same iterates up to rounding, 43 % less basis traffic
        per cycle), 1 in the plain form (the reference's Arnoldi).
        relative_rtol > 0 also requires ||b - A x|| / ||b|| <= relative_rtol
        (FP64) for "converged", and the mixed form refines until it holds."""
        if arnoldi_block is None:
            arnoldi_block = 2 if precision == "c64" else 1
        b = self.as_columns(b, torch.complex128)
        k = b.shape[0]
        x = torch.empty_like(b)
        opts = gmres_options(restart, maxiter, rtol, backward_error_tol, precision, adjoint, chunk,
                             max_refinements, inner_rtol, preconditioner, use_graph, arnoldi_block, relative_rtol)
        stats = (_capi.GmresStats*k)()
        status = _capi.lib().hz_gmres(self.handle, ctypes.byref(opts), ctypes.c_void_p(b.data_ptr()),
                                      ctypes.c_void_p(x.data_ptr()), k, self.n, stats, _stream())
        if status not in (_capi.HZ_OK, _capi.HZ_NOT_CONVERGED):
            _capi.check(status, "hz_gmres")
        out = [s.as_dict() for s in stats]
        if raise_on_failure and status == _capi.HZ_NOT_CONVERGED:
            bad = next(s for s in out if not s["converged"])
            raise RuntimeError(f"info={bad['info']}, iterations={bad['iterations']}, "
                               f"scaled residual={bad['scaled_residual']:.3e}, "
                               f"backward error={bad['backward_error']:.3e}")
        return x, out

    def autotune(self, nrhs, max_chunk=64, restart=30, preconditioner="multigrid", memory_budget=0):
        """Pick the stencil tile shapes and the mixed-form RHS chunk width by
        timing them on this grid (and multigrid hierarchy).  Returns
        {"chunk": c, "shapes": [...]}; the shapes stay in effect for later
        applies and solves."""
        pre = {"jacobi": _capi.PRE_JACOBI, "multigrid": _capi.PRE_MULTIGRID}[preconditioner]
        nshape = 1 + 3*len(self.hierarchy.levels) if self.hierarchy is not None else 1
        chunk = ctypes.c_int32(0)
        shapes = (ctypes.c_int32*nshape)()
        _capi.check(_capi.lib().hz_autotune(self.handle, int(nrhs), int(max_chunk), int(restart), pre,
                                            int(memory_budget), ctypes.byref(chunk), shapes, nshape, _stream()),
                    "hz_autotune")
        self.tuning = {"chunk": chunk.value, "shapes": list(shapes)}
        return self.tuning

    #: Slab tile shapes by index (z-planes per thread, tile rows, right-hand
    #: sides per pass, non-temporal stores, tile order: 0 rectangular, 1 flat,
    #: 2 rectangular z-fastest, 4 paired: two nodes per lane, even widths), as in the
    #: library's table.
    SLAB_SHAPES = ((4, 4, 1, 0, 0), (2, 4, 1, 0, 0), (4, 8, 1, 0, 0), (2, 8, 1, 0, 0), (4, 2, 1, 0, 0),
                   (2, 4, 2, 0, 0), (4, 4, 1, 1, 0), (2, 4, 2, 1, 0), (4, 4, 1, 1, 1), (2, 4, 1, 1, 1),
                   (2, 4, 2, 1, 1), (4, 4, 1, 1, 2), (2, 4, 2, 1, 2), (4, 8, 1, 1, 2), (2, 4, 1, 1, 4),
                   (2, 8, 1, 1, 4), (4, 4, 1, 1, 4))

    def material_vjp(self, u, lam):
        """dL/drho, dL/dc (grid-shaped, FP64, on the device) for
        dL = -Re(lambda^H dA u), summed over the columns of u and lambda
        ((N,), (N, k) or (k, N)); lambda is zeroed on identity rows."""
        u = self.as_columns(u, torch.complex128)
        lam = self.as_columns(lam, torch.complex128)
        if u.shape != lam.shape:
            raise ValueError("u and lambda must have the same columns")
        g_rho = torch.empty(self.n, dtype=torch.float64, device=self.device)
        g_c = torch.empty_like(g_rho)
        _capi.check(_capi.lib().hz_material_vjp(self.handle, ctypes.c_void_p(u.data_ptr()),
                                                ctypes.c_void_p(lam.data_ptr()), u.shape[0], self.n,
                                                ctypes.c_void_p(g_rho.data_ptr()), ctypes.c_void_p(g_c.data_ptr()),
                                                _stream()), "hz_material_vjp")
        # Fortran node order p = i + nx (j + ny k) -> an (nx, ny, nz) view.
        return tuple(g.view(self.shape[::-1]).permute(2, 1, 0) for g in (g_rho, g_c))

    def release_workspace(self):
        """Free the cached solver workspace (the next solve re-allocates)."""
        _capi.check(_capi.lib().hz_release_workspace(self.handle), "hz_release_workspace")

    def mg_transfer(self, level, x, prolong=False, out=None):
        """The multigrid transfer between `level` and level + 1 (hz_mg_transfer)
        on (k, n) complex64 device vectors: P^T x (fine -> coarse), or
        out + P x (coarse -> fine; out defaults to zeros)."""
        if self.hierarchy is None:
            raise RuntimeError("set_multigrid() first")
        shapes = [lev["shape"] for lev in self.hierarchy.levels]
        n_in = int(np.prod(shapes[level + 1 if prolong else level]))
        n_out = int(np.prod(shapes[level if prolong else level + 1]))
        x = x.to(torch.complex64).contiguous()
        if x.dim() != 2 or x.shape[1] != n_in:
            raise ValueError(f"expected (k, {n_in}) vectors")
        if out is None:
            out = torch.zeros((x.shape[0], n_out), dtype=torch.complex64, device=x.device)
        _capi.check(_capi.lib().hz_mg_transfer(self.handle, int(level), int(bool(prolong)), ctypes.c_void_p(x.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr()), x.shape[0], _stream()), "hz_mg_transfer")
        return out

    def mg_level(self, level):
        """Level `level` of the installed hierarchy as device tensors (hz_mg_level):
        {"rho", "kappa": (n_l,) float32, "faces": {f: (n_b * n_c,) complex64}}."""
        if self.hierarchy is None:
            raise RuntimeError("set_multigrid() first")
        shape = self.hierarchy.levels[level]["shape"]
        n = int(np.prod(shape))
        rho = torch.empty(n, dtype=torch.float32, device=self.device)
        kappa = torch.empty(n, dtype=torch.float32, device=self.device)
        faces = {}
        for f in range(6):
            ax = f//2
            b, c = [a for a in range(3) if a != ax]
            faces[f] = torch.empty(shape[b]*shape[c], dtype=torch.complex64, device=self.device)
        ptrs = (ctypes.c_void_p*6)(*[faces[f].data_ptr() for f in range(6)])
        mask = ctypes.c_int32()
        _capi.check(_capi.lib().hz_mg_level(self.handle, int(level), ctypes.c_void_p(rho.data_ptr()),
                                            ctypes.c_void_p(kappa.data_ptr()), ptrs, ctypes.byref(mask), _stream()),
                    "hz_mg_level")
        return {"rho": rho, "kappa": kappa, "faces": {f: t for f, t in faces.items() if mask.value >> f & 1}}

    def set_tile_shapes(self, shapes):
        """Pin the slab tile shapes (the layout `autotune` reports)."""
        arr = (ctypes.c_int32*len(shapes))(*[int(v) for v in shapes])
        _capi.check(_capi.lib().hz_set_tile_shapes(self.handle, arr, len(shapes)), "hz_set_tile_shapes")

    def set_multigrid(self, shift=1.2, pre=1, post=1, coarse_sweeps=30, weight=0.7, levels=None, cycle=1,
                      post_coarse=5, device_fields=True):
        """Install the shifted-Laplacian multigrid preconditioner (mixed form):
        `post` damped-Jacobi sweeps after the correction on the finest level,
        `post_coarse` on the levels below it (coarse sweeps cost an eighth per
        level); cycle 2 makes it a W-cycle below the finest level.  With
        device_fields the coarse materials and open-face terms are derived on
        the device; otherwise on the host (same definition)."""
        from .multigrid import Hierarchy
        self.hierarchy = Hierarchy(self, levels=levels, device_fields=device_fields).install(
            self, shift, pre, post, coarse_sweeps, weight, cycle, post_coarse)
        return self.hierarchy

    # -- sources ----------------------------------------------------------
    def set_sources(self, node, column, value, num_columns):
        """Source columns B as (node, column, value) triplets (node in 0..N-1)."""
        node = np.ascontiguousarray(node, dtype=np.int64)
        column = np.ascontiguousarray(column, dtype=np.int32)
        value = np.ascontiguousarray(value, dtype=np.float64)
        _capi.check(_capi.lib().hz_set_sources(
            self.handle, int(num_columns), len(node), node.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            column.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            value.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "hz_set_sources")
        self.num_sources = int(num_columns)

    def build_rhs(self, w):
        """b[:, r] = B w[:, r]; w is (k, M) complex (device or host)."""
        w = _device_tensor(torch.as_tensor(w).reshape(-1, self.num_sources), torch.complex128, self.device)
        k = w.shape[0]
        b = torch.empty((k, self.n), dtype=torch.complex128, device=self.device)
        _capi.check(_capi.lib().hz_build_rhs(self.handle, ctypes.c_void_p(w.data_ptr()), self.num_sources,
                                             ctypes.c_void_p(b.data_ptr()), k, self.n, _stream()),
                    "hz_build_rhs")
        return b

    def gather_sources(self, lam):
        """(k, M): B^T lambda for each column of lambda."""
        lam = self.as_columns(lam, torch.complex128)
        k = lam.shape[0]
        out = torch.empty((k, self.num_sources), dtype=torch.complex128, device=self.device)
        _capi.check(_capi.lib().hz_gather_sources(self.handle, ctypes.c_void_p(lam.data_ptr()), self.n,
                                                  ctypes.c_void_p(out.data_ptr()), self.num_sources, k,
                                                  _stream()), "hz_gather_sources")
        return out

    def sample(self, rows, weights, u):
        """Trilinear samples: rows/weights (P, 8); u (k, N) -> (k, P)."""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        if rows.ndim != 2 or rows.shape[1] != 8 or rows.min() < 0 or rows.max() >= self.n:
            raise ValueError("trilinear rows must be (P, 8) node indices inside the grid")
        rows = _device_tensor(rows, torch.int64, self.device)
        weights = _device_tensor(np.ascontiguousarray(weights, dtype=np.float64), torch.float64, self.device)
        u = self.as_columns(u, torch.complex128)
        k, npts = u.shape[0], rows.shape[0]
        out = torch.empty((k, npts), dtype=torch.complex128, device=self.device)
        _capi.check(_capi.lib().hz_sample(self.handle, ctypes.c_void_p(rows.data_ptr()),
                                          ctypes.c_void_p(weights.data_ptr()), npts,
                                          ctypes.c_void_p(u.data_ptr()), self.n,
                                          ctypes.c_void_p(out.data_ptr()), npts, k, _stream()), "hz_sample")
        return out
