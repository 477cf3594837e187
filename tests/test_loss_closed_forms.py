"""Closed forms that pin the focal_pressure loss and the focal gradient.

focal_pressure is defined in docs/GRADIENT_DERIVATION.html:914-930 as

    L = -alpha/K sum_k |u_{t_k}| + beta/(2N) sum_j |u_j|^2,
    c_{t_k} += -alpha/K u_{t_k}/(|u_{t_k}| + eps),   c_j += beta/N u_j,

(code: src/algorithm/training/phase_adjoint_optimizer.py:244-258).  A constant
field and a field with one nonzero node fix the 1/2, the mean-vs-sum of both
terms and the sign; each check runs against the oracle, the host and torch
evaluations of the product loss class and the differentiable torch loss.
Cotangents agree to 1e-11 relative: autograd differentiates |u| exactly,
without the eps = 1e-12 guard, a relative difference of eps/|u|.

The focal gradient (src/algorithm/baselines/common.py:189-195) is the phase
derivative of -alpha |f| for one exact sample f = r1 e^{i phi1} + r2 e^{i phi2}.
Trigonometry gives |f|^2 = |r1|^2 + |r2|^2 + 2|r1||r2| cos(D) with
D = phi1 + arg r1 - phi2 - arg r2, so dL/dphi1 = alpha |r1||r2| sin(D)/|f| and
dL/dphi2 = -dL/dphi1 (to 1e-10 relative: the eps guard again).
"""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import losses as olo
from ultrasound_solver_amd import autograd as hza
from ultrasound_solver_amd import training

SHAPE = (5, 4, 3)
N = int(np.prod(SHAPE))
DX = 1.0e-3


def _cfg(alpha, beta, indices):
    """The config fields FocalPressureLoss reads; targets on grid nodes."""
    pts = [np.asarray(i, float)*DX for i in indices]
    return SimpleNamespace(targets=pts, domain=SimpleNamespace(origin_m=(0.0, 0.0, 0.0), dx=DX),
                           training=SimpleNamespace(target_weight=alpha, background_weight=beta))


def _all_forms(u, indices, alpha, beta):
    """(loss, cotangent) from the oracle, the product loss (host, torch) and autograd."""
    out = {"oracle": olo.focal_pressure(u, indices, alpha, beta)}
    lm = training.FocalPressureLoss(_cfg(alpha, beta, indices),
                                    training.AmplitudeTarget(np.zeros(SHAPE), np.ones(SHAPE)))
    assert lm.target_indices == [tuple(i) for i in indices]
    out["host"] = lm.evaluate(u)
    val, cot = lm.evaluate_device(torch.as_tensor(u))
    out["torch"] = (float(val), cot.numpy())
    ut = torch.tensor(u, requires_grad=True)
    val = hza.focal_pressure_loss(ut, indices, alpha, beta)
    val.backward()
    out["autograd"] = (float(val.detach()), ut.grad.numpy())
    return out


@pytest.mark.parametrize("z0", [2.0 - 1.5j, 0.3j])
def test_focal_pressure_constant_field(z0):
    alpha, beta = 1.7, 0.6
    indices = [(1, 2, 0), (4, 0, 2)]
    u = np.full(SHAPE, z0, dtype=np.complex128)
    a = abs(z0)
    want_loss = -alpha*a + 0.5*beta*a*a
    want = np.full(SHAPE, beta/N*z0)
    for i in indices:
        want[i] += -alpha/2*z0/(a + olo.EPS)
    for name, (loss, cot) in _all_forms(u, indices, alpha, beta).items():
        assert loss == pytest.approx(want_loss, rel=1e-14), name
        np.testing.assert_allclose(cot, want, rtol=1e-11, atol=1e-16, err_msg=name)


def test_focal_pressure_single_node():
    alpha, beta = 0.9, 2.5
    t = (3, 1, 2)
    v = -1.2 + 0.7j
    u = np.zeros(SHAPE, dtype=np.complex128)
    u[t] = v
    want_loss = -alpha*abs(v) + beta/(2*N)*abs(v)**2
    want = np.zeros(SHAPE, dtype=np.complex128)
    want[t] = -alpha*v/(abs(v) + olo.EPS) + beta/N*v
    for name, (loss, cot) in _all_forms(u, [t], alpha, beta).items():
        assert loss == pytest.approx(want_loss, rel=1e-14), name
        np.testing.assert_allclose(cot, want, rtol=1e-11, atol=1e-16, err_msg=name)


def test_focal_pressure_repeated_target_accumulates():
    """Two copies of one target index: K = 2, both cotangent entries land on one node."""
    alpha, beta = 1.0, 0.0
    t = (2, 2, 1)
    v = 3.0 + 4.0j
    u = np.zeros(SHAPE, dtype=np.complex128)
    u[t] = v
    for name, (loss, cot) in _all_forms(u, [t, t], alpha, beta).items():
        assert loss == pytest.approx(-5.0, rel=1e-15), name
        assert cot[t] == pytest.approx(-(0.6 + 0.8j), rel=1e-12), name
        assert np.count_nonzero(cot) == 1, name


def _two_emitter_case(seed):
    rng = np.random.default_rng(seed)
    r = (rng.standard_normal(2) + 1j*rng.standard_normal(2)).reshape(1, 2)
    ph = rng.uniform(0, 2*np.pi, 2)
    alpha = 1.3
    f = r @ np.exp(1j*ph)
    d = ph[0] + np.angle(r[0, 0]) - ph[1] - np.angle(r[0, 1])
    g1 = alpha*abs(r[0, 0])*abs(r[0, 1])*np.sin(d)/abs(f[0])
    return r, ph, alpha, f, np.array([g1, -g1])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_focal_gradient_two_emitter_closed_form(seed):
    r, ph, alpha, f, want = _two_emitter_case(seed)
    np.testing.assert_allclose(olo.focal_gradient(r, f, ph, alpha), want, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(training.focal_gradient(r, f, ph, alpha), want, rtol=1e-10, atol=1e-14)
    got = training.focal_gradient(torch.as_tensor(r), torch.as_tensor(f), ph, alpha).numpy()
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-14)


def test_focal_gradient_averages_over_samples():
    """Two samples that are each a single emitter: |f_p| does not depend on the
    phases, so the gradient vanishes; P copies of the two-emitter sample give
    the two-emitter gradient (the 1/P mean)."""
    r = np.array([[1.0 + 2.0j, 0.0], [0.0, -0.5j]])
    ph = np.array([0.4, 2.0])
    for fn in (olo.focal_gradient, training.focal_gradient):
        np.testing.assert_allclose(fn(r, r @ np.exp(1j*ph), ph, 2.0), 0.0, atol=1e-15)
    r2, ph2, alpha, f, want = _two_emitter_case(5)
    rr = np.repeat(r2, 3, axis=0)
    for fn in (olo.focal_gradient, training.focal_gradient):
        np.testing.assert_allclose(fn(rr, rr @ np.exp(1j*ph2), ph2, alpha), want, rtol=1e-10)


def _ratio_case():
    ball = np.zeros(SHAPE, dtype=bool)
    ball[2, 1:3, 2] = True
    rest = np.zeros(SHAPE, dtype=bool)
    rest[:, :, 1:] = True
    rest &= ~ball
    return ball, rest


def test_focal_contrast_ratio_piecewise_constant():
    """u = a on the ball, b on the rest (and anything elsewhere): L = |a|^2/|b|^2."""
    ball, rest = _ratio_case()
    a, b = 1.5 - 2.0j, 0.25 + 0.5j
    u = np.full(SHAPE, 7.0 + 1j)
    u[ball], u[rest] = a, b
    want = abs(a)**2/abs(b)**2
    loss, cot = olo.focal_contrast_ratio(u, ball, rest)
    assert loss == pytest.approx(want, rel=1e-14)
    assert np.all(cot[~(ball | rest)] == 0)
    ut = torch.tensor(u, requires_grad=True)
    val = hza.focal_contrast_ratio_loss(ut, torch.as_tensor(ball), torch.as_tensor(rest))
    val.backward()
    assert float(val.detach()) == pytest.approx(want, rel=1e-14)
    np.testing.assert_allclose(ut.grad.numpy(), cot, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("seed", [0, 1])
def test_focal_contrast_ratio_scale_invariance_and_differences(seed):
    """L(s u) = L(u): the cotangent is orthogonal to u (Re(c^H u) = 0), and
    central differences of L along random directions match Re(c^H du)."""
    ball, rest = _ratio_case()
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(SHAPE) + 1j*rng.standard_normal(SHAPE)
    loss, cot = olo.focal_contrast_ratio(u, ball, rest)
    assert abs(np.real(np.vdot(cot, u))) < 1e-13*np.linalg.norm(cot)*np.linalg.norm(u)
    assert olo.focal_contrast_ratio(3.7j*u, ball, rest, False)[0] == pytest.approx(loss, rel=1e-14)
    for _ in range(3):
        du = rng.standard_normal(SHAPE) + 1j*rng.standard_normal(SHAPE)
        h = 1e-6
        fd = (olo.focal_contrast_ratio(u + h*du, ball, rest, False)[0]
              - olo.focal_contrast_ratio(u - h*du, ball, rest, False)[0])/(2*h)
        assert np.real(np.vdot(cot, du)) == pytest.approx(fd, rel=1e-7)
