"""FP64 restatement of the reference's row-scaled Jacobi GMRES.

Follows ``src/algorithm/solvers/iterative_solver.py``:

* zero RHS short-circuit (:31-35);
* row scale R_ii = 1/max_j |A_ij| (:37-38), scaled operator R A and RHS R b (:39-44);
* Jacobi M = diag(1/(R_ii A_ii)) (:46-56);
* ``cupyx.scipy.sparse.linalg.gmres`` call (:59-69) -- restated below from the
  public CuPy implementation (cupy-cuda12x >= 14, ``requirements.txt:11``):
  right preconditioning, classical Gram-Schmidt (two gemv), full ``restart``
  Arnoldi steps per cycle, Hessenberg least squares once per cycle,
  convergence tested on the true residual at the top of each cycle,
  ``maxiter`` counting inner steps;
* acceptance: info == 0, finite, scaled residual <= rtol, componentwise
  backward error <= tol (:70-97), iterations = cycles * restart (:81).

Also the adjoint route of ``src/algorithm/solvers/helmholtz_solver.py:606-632``
(the same solver run on conj(A)^T) and a complex128 sparse-LU direct oracle.

Test infrastructure only (see ``oracle/__init__.py``).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def row_scale(a):
    row_max = np.asarray(abs(a).max(axis=1).toarray()).ravel()
    return np.where(row_max > 0.0, 1.0/np.where(row_max > 0.0, row_max, 1.0), 1.0)


def jacobi_inverse(a, scale):
    d = scale*a.diagonal()
    safe = np.where(np.abs(d) > 0.0, d, 1.0)
    return np.where(np.abs(d) > 0.0, 1.0/safe, 1.0)


def cupy_style_gmres(matvec, psolve, b, rtol, atol, restart, maxiter, callback=None):
    """Restarted GMRES with CuPy's loop structure; returns (M x, info)."""
    n = b.shape[0]
    b_norm = np.linalg.norm(b)
    if b_norm == 0:
        return b.copy(), 0
    atol = max(float(atol), rtol*float(b_norm)) if atol is not None else rtol*float(b_norm)
    restart = min(restart, n)
    x = np.zeros_like(b)
    basis = np.empty((n, restart), dtype=b.dtype)
    hess = np.zeros((restart + 1, restart), dtype=b.dtype)
    rhs_ls = np.zeros(restart + 1, dtype=b.dtype)
    iters = 0
    while True:
        mx = psolve(x)
        r = b - matvec(mx)
        r_norm = np.linalg.norm(r)
        if callback is not None and iters > 0:
            callback(r_norm/b_norm)
        if r_norm <= atol or iters >= maxiter:
            break
        v = r/r_norm
        basis[:, 0] = v
        rhs_ls[0] = r_norm
        for j in range(restart):
            z = psolve(v)
            u = matvec(z)
            h = basis[:, :j + 1].conj().T @ u
            u = u - basis[:, :j + 1] @ h
            hess[:j + 1, j] = h
            hess[j + 1, j] = np.linalg.norm(u)
            if j + 1 < restart:
                v = u/hess[j + 1, j]
                basis[:, j + 1] = v
        y = np.linalg.lstsq(hess, rhs_ls, rcond=None)[0]
        x = x + basis @ y
        iters += restart
    info = 0
    if iters == maxiter and not (r_norm <= atol):
        info = iters
    return mx, info


class ReferenceGmres:
    """Row-scaled, Jacobi right-preconditioned GMRES with the reference's checks."""

    def __init__(self, restart=30, maxiter=3000, rtol=1.0e-6, backward_error_tol=1.0e-4):
        self.restart = int(restart)
        self.maxiter = int(maxiter)
        self.rtol = float(rtol)
        self.backward_error_tol = float(backward_error_tol)
        self.last_iterations = 0
        self.last_scaled_residual = None
        self.last_backward_error = None
        self.residual_history = []

    def solve(self, a, rhs):
        a = sp.csr_matrix(a)
        rhs = np.asarray(rhs, dtype=np.complex128)
        if float(np.linalg.norm(rhs)) == 0.0:
            self.last_iterations = 0
            self.last_scaled_residual = 0.0
            self.last_backward_error = 0.0
            return np.zeros_like(rhs)
        scale = row_scale(a)
        minv = jacobi_inverse(a, scale)
        b_hat = scale*rhs
        history = []
        x, info = cupy_style_gmres(lambda v: scale*(a @ v), lambda v: minv*v, b_hat,
                                   self.rtol, 0.0, self.restart, self.maxiter, history.append)
        scaled_res = float(np.linalg.norm(scale*(a @ x) - b_hat)/np.linalg.norm(b_hat))
        bwd = backward_error(a, x, rhs)
        iterations = len(history)*self.restart
        self.last_iterations = iterations
        self.last_scaled_residual = scaled_res
        self.last_backward_error = bwd
        self.residual_history = history
        if (info != 0 or not np.isfinite(x).all() or scaled_res > self.rtol
                or bwd > self.backward_error_tol):
            raise RuntimeError(f"info={info}, iterations={iterations}, "
                               f"scaled residual={scaled_res:.3e}, backward error={bwd:.3e}")
        return x


def backward_error(a, x, b):
    """max_i |Ax-b|_i / (|A||x| + |b|)_i (iterative_solver.py:75-79)."""
    res = a @ x - b
    den = abs(a) @ np.abs(x) + np.abs(b)
    return float(np.max(np.abs(res)/np.maximum(den, 1.0e-30)))


def scaled_residual(a, x, b):
    scale = row_scale(a)
    return float(np.linalg.norm(scale*(a @ x - b))/np.linalg.norm(scale*b))


def adjoint_matrix(a):
    """conj(A)^T as CSR (helmholtz_solver.py:613-620)."""
    return sp.csr_matrix(a).conjugate().transpose().tocsr()


def direct_solve(a, b):
    """complex128 sparse LU (scipy SuperLU); b may be (N,) or (N, k)."""
    lu = spla.splu(sp.csc_matrix(a, dtype=np.complex128))
    return lu.solve(np.asarray(b, dtype=np.complex128))
