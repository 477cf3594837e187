"""torch.autograd front end of the phase -> field -> loss path.

``PhaseField`` maps phases phi (B, M) to fields u (B, N) with
u = A^{-1} B_src (a * e^{i phi}) (one batched device GMRES solve), and its
backward is the adjoint solve A^H lambda = c followed by the source gather,
G^H c = a * B_src^T lambda, giving dL/dphi = Re(conj(G^H c) * i e^{i phi}) --
the reference's phase VJP (src/algorithm/training/phase_response_basis.py:208-253)
without forming G, in the PDE-adjoint form of docs/GRADIENT_DERIVATION.html
(:1442-1513).  ``BasisPhaseField`` is the same map through a precomputed
basis G.  torch's gradient of a real loss with respect to a complex tensor
is dL/dRe u + i dL/dIm u, which is exactly the reference's cotangent c
(dL = Re(c^H du)).

The three losses are written with differentiable torch ops so that autograd
produces the reference's cotangents (phase_adjoint_optimizer.py:163-364)
up to the eps guard of |u|.
"""
import numpy as np
import torch


class PhaseEngine:
    """Solver settings shared by the forward and adjoint solves of a scene."""

    def __init__(self, op, amplitudes, precision="c64", preconditioner="jacobi", rtol=1e-6, restart=None,
                 maxiter=20000, chunk=32, backward_error_tol=1e-4):
        """restart: GMRES restart length; by default 10 for the mixed form
        (its inner complex64 GMRES) and 30, the reference's, for FP64."""
        if restart is None:
            restart = 10 if precision == "c64" else 30
        self.op = op
        self.amplitudes = torch.as_tensor(np.asarray(amplitudes, dtype=np.float64), device=op.device)
        self.kw = dict(precision=precision, preconditioner=preconditioner, rtol=rtol, restart=restart,
                       maxiter=maxiter, chunk=chunk, backward_error_tol=backward_error_tol)
        self.last_forward_stats = None
        self.last_adjoint_stats = None

    def _solve(self, b, adjoint):
        x, st = self.op.gmres(b, adjoint=adjoint, **self.kw)
        bad = [s for s in st if not s["converged"]]
        if bad:
            s = bad[0]
            raise RuntimeError(f"{'adjoint' if adjoint else 'forward'} solve failed: info={s['info']}, "
                               f"iterations={s['iterations']}, scaled residual={s['scaled_residual']:.3e}, "
                               f"backward error={s['backward_error']:.3e}")
        return x, st

    def forward(self, phases):
        w = self.amplitudes[None, :]*torch.exp(1j*phases.to(torch.float64))
        u, st = self._solve(self.op.build_rhs(w), adjoint=False)
        self.last_forward_stats = st
        return u

    def response_adjoint(self, cot):
        """G^H c for each row of c (B, N): a * B^T lambda with A^H lambda = c."""
        lam, st = self._solve(cot.to(torch.complex128).contiguous(), adjoint=True)
        self.last_adjoint_stats = st
        return self.amplitudes[None, :]*self.op.gather_sources(lam)


class PhaseField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phases, engine):
        ctx.engine = engine
        ctx.save_for_backward(phases)
        return engine.forward(phases)

    @staticmethod
    def backward(ctx, grad_u):
        (phases,) = ctx.saved_tensors
        resp = ctx.engine.response_adjoint(grad_u)
        w = torch.exp(1j*phases.to(torch.float64))
        grad = torch.real(torch.conj(resp)*(1j*w))
        return grad.to(phases.dtype), None


class BasisPhaseField(torch.autograd.Function):
    """u = G e^{i phi} for (B, M) phases on a device basis G (N, M), through
    the library's one-pass basis kernels; backward is G^H c per row."""

    @staticmethod
    def forward(ctx, phases, basis):
        from . import basis as dev_basis
        ctx.basis = basis
        ctx.save_for_backward(phases)
        w = torch.exp(1j*phases.to(torch.float64))
        return torch.stack([dev_basis.field(basis, w[b]) for b in range(w.shape[0])])

    @staticmethod
    def backward(ctx, grad_u):
        from . import basis as dev_basis
        (phases,) = ctx.saved_tensors
        w = torch.exp(1j*phases.to(torch.float64))
        resp = torch.stack([dev_basis.adjoint(ctx.basis, grad_u[b].contiguous()) for b in range(w.shape[0])])
        grad = dev_basis.phase_gradient(resp, w)
        return grad.to(phases.dtype), None


def phase_field(phases, engine):
    """(M,) or (B, M) phases -> (N,) or (B, N) complex128 fields (Fortran node order)."""
    single = phases.dim() == 1
    u = PhaseField.apply(phases.reshape(1, -1) if single else phases, engine)
    return u[0] if single else u


def grid_view(u, shape):
    """(..., N) Fortran-ordered vector -> (..., nx, ny, nz) view."""
    nx, ny, nz = shape
    v = u.reshape(*u.shape[:-1], nz, ny, nx)
    return v.permute(*range(v.dim() - 3), v.dim() - 1, v.dim() - 2, v.dim() - 3)


# --- differentiable losses (per field; fields are (nx, ny, nz) views) ----
def field_match_loss(u, target, weight):
    amp = u.abs()
    return 0.5*torch.sum((weight*(amp - target))**2)


def _focal(u, indices, weight, focal_field=None):
    if focal_field is None:
        idx = tuple(torch.as_tensor(a, device=u.device) for a in np.asarray(indices, dtype=np.int64).T)
        values = u[idx]
    else:
        values = focal_field
    return -weight*values.abs().mean()


def focal_pressure_loss(u, indices, target_weight, background_weight, focal_field=None):
    loss = _focal(u, indices, target_weight, focal_field)
    if background_weight > 0:
        loss = loss + 0.5*background_weight*torch.mean(u.abs()**2)
    return loss


def focal_contrast_loss(u, indices, target_weight, sidelobe_weight, temperature, mask, focal_field=None):
    loss = _focal(u, indices, target_weight, focal_field)
    bg = u.abs()[mask]
    smooth = temperature*(torch.logsumexp(bg/temperature, 0) - np.log(bg.numel()))
    return loss + sidelobe_weight*smooth




def focal_contrast_ratio_loss(u, ball, rest):
    """mean |u|^2 over the ball / mean |u|^2 over the rest (boolean masks)."""
    p2 = u.real**2 + u.imag**2
    return p2[ball].mean()/p2[rest].mean()
