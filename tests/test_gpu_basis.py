"""One-pass response-basis kernels (hz_basis_*) against the oracle's basis
field, phase VJP and field_match loss (phase_response_basis.py:177-253,
phase_adjoint_optimizer.py:163-200) on seeded random bases, to FP64
rounding (1e-12 relative)."""
import numpy as np
import pytest

from oracle import losses as olo

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _basis(n, m, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, m)) + 1j*rng.standard_normal((n, m))
    return g, torch.as_tensor(g, device="cuda:0")


def _rel(a, b):
    return np.max(np.abs(a - b))/np.max(np.abs(b))


@pytest.mark.parametrize("n,m", [(5000, 16), (4099, 300), (1000, 2048)])
def test_field_and_adjoint_match_numpy(n, m):
    from ultrasound_solver_amd import basis
    g, gd = _basis(n, m, 1)
    rng = np.random.default_rng(2)
    ph = rng.uniform(0, 2*np.pi, m)
    u = basis.field(gd, torch.as_tensor(np.exp(1j*ph), device="cuda:0")).cpu().numpy()
    assert _rel(u, g @ np.exp(1j*ph)) < 1e-12
    c = rng.standard_normal(n) + 1j*rng.standard_normal(n)
    resp = basis.adjoint(gd, torch.as_tensor(c, device="cuda:0")).cpu().numpy()
    assert _rel(resp, g.conj().T @ c) < 1e-12
    grad = basis.phase_gradient(torch.as_tensor(resp, device="cuda:0"),
                                torch.as_tensor(np.exp(1j*ph), device="cuda:0")).cpu().numpy()
    assert _rel(grad, olo.phase_vjp(g, c, ph)) < 1e-12


@pytest.mark.parametrize("n,m", [(6000, 64), (3001, 512)])
def test_fused_field_match_matches_oracle(n, m):
    from ultrasound_solver_amd import basis
    g, gd = _basis(n, m, 3)
    rng = np.random.default_rng(4)
    ph = rng.uniform(0, 2*np.pi, m)
    u_ref = g @ np.exp(1j*ph)
    target = np.abs(u_ref)*rng.uniform(0.5, 1.5, n)
    weight = rng.uniform(0.0, 2.0, n)
    loss_ref, cot = olo.field_match(u_ref, target, weight)
    grad_ref = olo.phase_vjp(g, cot, ph)
    loss, grad, u = basis.field_match(gd, ph, target, weight, return_field=True)
    assert float(loss) == pytest.approx(loss_ref, rel=1e-12)
    assert _rel(grad.cpu().numpy(), grad_ref) < 1e-12
    assert _rel(u.cpu().numpy(), u_ref) < 1e-12
    loss2, grad2 = basis.field_match(gd, ph, target, weight)
    assert torch.equal(grad2, grad) and float(loss2) == float(loss)


def test_fused_kernel_rejects_wide_rows():
    from ultrasound_solver_amd import basis
    _, gd = _basis(100, 600, 5)
    with pytest.raises(RuntimeError, match="512"):
        basis.field_match(gd, np.zeros(600), np.ones(100), np.ones(100))


def test_basis_autograd_vjp_matches_oracle_and_fd():
    """BasisPhaseField (device kernels) under torch autograd: the phase
    gradient of field_match equals the oracle basis VJP and central
    differences."""
    from ultrasound_solver_amd import autograd as hza
    shape = (6, 5, 4)
    n, m = int(np.prod(shape)), 7
    rng = np.random.default_rng(3)
    g = rng.standard_normal((n, m)) + 1j*rng.standard_normal((n, m))
    target = np.abs(rng.standard_normal(shape))
    weight = rng.uniform(0.1, 1, shape)
    ph = torch.tensor(rng.uniform(0, 2*np.pi, m), requires_grad=True, device="cuda:0")
    u = hza.BasisPhaseField.apply(ph.reshape(1, -1), torch.as_tensor(g, device="cuda:0"))[0]
    val = hza.field_match_loss(hza.grid_view(u, shape), torch.as_tensor(target, device="cuda:0"),
                               torch.as_tensor(weight, device="cuda:0"))
    val.backward()
    grad = ph.grad.cpu().numpy()
    p0 = ph.detach().cpu().numpy()
    u_np = olo.basis_field(g, p0, shape)
    _, cot = olo.field_match(u_np, target, weight)
    ref = olo.phase_vjp(g, cot, p0)
    np.testing.assert_allclose(grad, ref, rtol=1e-9, atol=1e-12)
    fd = olo.central_difference(
        lambda p: olo.field_match(olo.basis_field(g, p, shape), target, weight, False)[0], p0, 1e-6, count=m)
    assert max(olo.scaled_gradient_error(a, b) for a, b in zip(grad, fd)) < 1e-6
