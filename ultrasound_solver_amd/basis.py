"""Device response-basis operations on a phase response basis G held as an
(N, M) complex128 torch tensor (rows = nodes in Fortran order), through the
library's one-pass kernels (phase_response_basis.py:177-253,
phase_adjoint_optimizer.py:163-364)."""
import ctypes

import torch

from . import _capi
from .operator import _stream

EPS = 1.0e-12  # amplitude regulariser of the reference's cotangents


def _check(G):
    if not isinstance(G, torch.Tensor) or G.dtype != torch.complex128 or G.dim() != 2 or not G.is_cuda:
        raise ValueError("G must be a (N, M) complex128 device tensor")
    if G.stride(1) != 1:
        raise ValueError("G rows must be contiguous")
    return G.shape[0], G.shape[1], G.stride(0)


def _vec(x, n, like):
    x = torch.as_tensor(x, device=like.device).to(torch.complex128).contiguous().reshape(-1)
    if x.numel() != n:
        raise ValueError(f"expected {n} entries, got {x.numel()}")
    return x


def field(G, w):
    """u = G w."""
    n, m, ld = _check(G)
    w = _vec(w, m, G)
    u = torch.empty(n, dtype=torch.complex128, device=G.device)
    _capi.check(_capi.lib().hz_basis_field(ctypes.c_void_p(G.data_ptr()), ld, n, m, ctypes.c_void_p(w.data_ptr()),
                                           ctypes.c_void_p(u.data_ptr()), _stream()), "hz_basis_field")
    return u


def adjoint(G, c):
    """g = G^H c."""
    n, m, ld = _check(G)
    c = _vec(c, n, G)
    g = torch.empty(m, dtype=torch.complex128, device=G.device)
    _capi.check(_capi.lib().hz_basis_adjoint(ctypes.c_void_p(G.data_ptr()), ld, n, m, ctypes.c_void_p(c.data_ptr()),
                                             ctypes.c_void_p(g.data_ptr()), _stream()), "hz_basis_adjoint")
    return g


def phase_gradient(g, w):
    """dL/dphi = Re(conj(G^H c) * i w) from g = G^H c and w = a e^{i phi}."""
    return torch.real(torch.conj(g)*(1j*w))


def field_match(G, phases, target, weight, eps=EPS, return_field=False):
    """One pass over G: (loss, dL/dphi[, u]) of the reference's field_match
    loss 0.5 |W (|u| - a)|^2 with u = G e^{i phi} (target a, weight W given
    per node in Fortran order)."""
    n, m, ld = _check(G)
    ph = torch.as_tensor(phases, device=G.device, dtype=torch.float64).reshape(-1)
    if ph.numel() != m:
        raise ValueError(f"expected {m} phases")
    w = torch.exp(1j*ph).contiguous()
    tgt = torch.as_tensor(target, device=G.device, dtype=torch.float64).contiguous().reshape(-1)
    wt = torch.as_tensor(weight, device=G.device, dtype=torch.float64).contiguous().reshape(-1)
    if tgt.numel() != n or wt.numel() != n:
        raise ValueError("target and weight need one entry per basis row")
    loss = torch.empty(1, dtype=torch.float64, device=G.device)
    g = torch.empty(m, dtype=torch.complex128, device=G.device)
    u = torch.empty(n, dtype=torch.complex128, device=G.device) if return_field else None
    _capi.check(_capi.lib().hz_basis_field_match(
        ctypes.c_void_p(G.data_ptr()), ld, n, m, ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(tgt.data_ptr()),
        ctypes.c_void_p(wt.data_ptr()), float(eps), ctypes.c_void_p(loss.data_ptr()), ctypes.c_void_p(g.data_ptr()),
        ctypes.c_void_p(u.data_ptr() if u is not None else 0), _stream()), "hz_basis_field_match")
    out = (loss[0], phase_gradient(g, w))
    return out + (u,) if return_field else out


FUSED_MAX_EMITTERS = 512
LOSS_KINDS = {"field_match": _capi.LOSS_FIELD_MATCH, "energy": _capi.LOSS_ENERGY,
              "softmax": _capi.LOSS_SOFTMAX, "ratio": _capi.LOSS_RATIO}


PHASE_MAX_TARGETS = 256


def phase_prologue(ph, response=None):
    """(w = e^{i ph}, focal = response @ w or None) on the device in one
    launch (hz_phase_prologue): ph (M,) float64, response (P, M) complex128."""
    ph = ph.to(torch.float64).contiguous()
    m = ph.numel()
    if not 1 <= m <= FUSED_MAX_EMITTERS:
        raise ValueError(f"need 1 <= M <= {FUSED_MAX_EMITTERS} emitters")
    w = torch.empty(m, dtype=torch.complex128, device=ph.device)
    p, r_ptr, focal = 0, 0, None
    if response is not None:
        response = response.contiguous()
        p = response.shape[0]
        if response.shape != (p, m) or response.dtype != torch.complex128 or not 1 <= p <= PHASE_MAX_TARGETS:
            raise ValueError(f"response must be (P, M) complex128 with 1 <= P <= {PHASE_MAX_TARGETS}")
        focal = torch.empty(p, dtype=torch.complex128, device=ph.device)
        r_ptr = response.data_ptr()
    _capi.check(_capi.lib().hz_phase_prologue(ctypes.c_void_p(ph.data_ptr()), m, ctypes.c_void_p(r_ptr), p,
                                              ctypes.c_void_p(w.data_ptr()),
                                              ctypes.c_void_p(focal.data_ptr() if focal is not None else 0),
                                              _stream()), "hz_phase_prologue")
    return w, focal


def phase_epilogue(w, g, loss, response=None, focal=None, target_weight=0.0, eps=EPS, gradient=True):
    """(M + 1,) float64 device tensor [loss + focal term, phase gradient]
    (hz_phase_epilogue): the focal term -target_weight mean |focal| and the
    gradient Re(conj(g + R^H c) i w) with the focal cotangent c; g = G^H c
    of the grid loss (None without gradient), loss its 0-d value."""
    m = w.numel()
    out = torch.empty(m + 1, dtype=torch.float64, device=w.device)
    p = 0 if focal is None else focal.numel()
    _capi.check(_capi.lib().hz_phase_epilogue(
        ctypes.c_void_p(w.data_ptr()), m, ctypes.c_void_p(g.data_ptr() if g is not None else 0),
        ctypes.c_void_p(response.data_ptr() if p else 0), p, ctypes.c_void_p(focal.data_ptr() if p else 0),
        float(target_weight), float(eps), ctypes.c_void_p(loss.data_ptr()), int(bool(gradient)),
        ctypes.c_void_p(out.data_ptr()), _stream()), "hz_phase_epilogue")
    return out


def fused_loss(G, w, kind, target=None, weight=None, mask=None, eps=EPS, scale=0.0, temperature=1.0,
               count_a=0, count_b=0, return_field=False, static=False):
    """One pass over G for a loss of the field u = G w (hz_basis_loss):
    returns (loss as a 0-d device tensor, g = G^H c (M,) device[, u]).

    kind: "field_match" (target, weight per row), "energy" (0.5 scale sum
    |u|^2), "softmax" (scale * smooth-max of |u| over rows with mask != 0 at
    `temperature`; count_a masked rows) or "ratio" (mean |u|^2 over mask == 1
    divided by mean over mask == 2; count_a, count_b rows).  Per-row arrays
    are device tensors in the row order of G; mask is uint8.  static=True
    calls hz_basis_loss_static, which allocates nothing (its partial sums sit
    in a static device array shared by every call on the GPU), so that the
    call can be captured into a graph."""
    n, m, ld = _check(G)
    if m > FUSED_MAX_EMITTERS:
        raise ValueError(f"the fused loss kernels take at most {FUSED_MAX_EMITTERS} emitters")
    w = _vec(w, m, G)
    desc = _capi.BasisLossDesc()
    desc.kind = LOSS_KINDS[kind]
    keep = []
    for name, arr, dtype in (("target", target, torch.float64), ("weight", weight, torch.float64),
                             ("mask", mask, torch.uint8)):
        if arr is None:
            continue
        a = torch.as_tensor(arr, device=G.device).to(dtype).contiguous().reshape(-1)
        if a.numel() != n:
            raise ValueError(f"{name} needs one entry per basis row")
        keep.append(a)
        setattr(desc, name, a.data_ptr())
    desc.eps, desc.scale, desc.temperature = float(eps), float(scale), float(temperature)
    desc.count_a, desc.count_b = int(count_a), int(count_b)
    loss = torch.empty(1, dtype=torch.float64, device=G.device)
    g = torch.empty(m, dtype=torch.complex128, device=G.device)
    u = torch.empty(n, dtype=torch.complex128, device=G.device) if return_field else None
    name = "hz_basis_loss_static" if static else "hz_basis_loss"
    _capi.check(getattr(_capi.lib(), name)(ctypes.c_void_p(G.data_ptr()), ld, n, m, ctypes.c_void_p(w.data_ptr()),
                                           ctypes.byref(desc), ctypes.c_void_p(loss.data_ptr()),
                                           ctypes.c_void_p(g.data_ptr()),
                                           ctypes.c_void_p(u.data_ptr() if u is not None else 0), _stream()), name)
    out = (loss[0], g)
    return out + (u,) if return_field else out
