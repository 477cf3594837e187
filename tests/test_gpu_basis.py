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


def _fused_case(kind, n, m, seed, temperature=2.0):
    """(g, ph, u, kwargs for basis.fused_loss, oracle loss, oracle G^H c, the
    bound |G|^T |c| that scales the rounding of each entry of G^H c)."""
    g, gd = _basis(n, m, seed)
    rng = np.random.default_rng(seed + 1)
    ph = rng.uniform(0, 2*np.pi, m)
    u = g @ np.exp(1j*ph)
    origin = [(0,)]  # a focal index with weight 0: only the background term remains
    if kind == "energy":
        beta = 0.7
        ref_loss, cot = olo.focal_pressure(u, origin, 0.0, beta)
        kw = dict(scale=beta/n)
    elif kind == "softmax":
        beta = 1.3
        mask = rng.random(n) < 0.7
        mask[rng.integers(0, n)] = True
        ref_loss, cot = olo.focal_contrast(u, origin, 0.0, beta, temperature, mask)
        kw = dict(mask=torch.as_tensor(mask.astype(np.uint8), device="cuda:0"), scale=beta,
                  temperature=temperature, count_a=int(mask.sum()))
    else:
        cls = rng.integers(0, 3, n)
        cls[0], cls[-1] = 1, 2
        ref_loss, cot = olo.focal_contrast_ratio(u, cls == 1, cls == 2)
        kw = dict(mask=torch.as_tensor(cls.astype(np.uint8), device="cuda:0"),
                  count_a=int((cls == 1).sum()), count_b=int((cls == 2).sum()))
    return gd, ph, u, kw, ref_loss, g.conj().T @ cot, np.abs(g).T @ np.abs(cot)


def _close(got, want, bound, tol):
    """Entries of G^H c within tol of the magnitudes summed into them (the
    ratio loss with one emitter has G^H c = 0 exactly: scale invariance)."""
    return np.max(np.abs(got - want)) <= tol*np.max(bound)


@pytest.mark.parametrize("n,m", [(7001, 1), (6000, 144), (3001, 512), (5, 33), (299, 40), (256, 7)])
@pytest.mark.parametrize("kind", ["energy", "softmax", "ratio"])
def test_fused_losses_match_oracle(kind, n, m):
    """hz_basis_loss: the focal_pressure background, the focal_contrast
    smooth-max and the contrast ratio, loss and G^H c to 1e-12 of the oracle,
    from one row up to the widest row (512 emitters), fewer rows than waves,
    75 workgroups (folded into 64 slices of 1 or 2) and exactly 64 (combined
    without a fold)."""
    from ultrasound_solver_amd import basis
    gd, ph, u_ref, kw, ref_loss, ref_g, bound = _fused_case(kind, n, m, 11)
    w = torch.as_tensor(np.exp(1j*ph), device="cuda:0")
    loss, g, u = basis.fused_loss(gd, w, kind, return_field=True, **kw)
    assert float(loss) == pytest.approx(ref_loss, rel=1e-12)
    assert _close(g.cpu().numpy(), ref_g, bound, 1e-12)
    assert _rel(u.cpu().numpy(), u_ref) < 1e-12
    loss2, g2 = basis.fused_loss(gd, w, kind, **kw)
    assert torch.equal(g2, g) and float(loss2) == float(loss)


@pytest.mark.parametrize("temperature", [0.05, 40.0])
def test_fused_softmax_running_maximum(temperature):
    """A sharp and a flat softmax (T = 0.05 underflows most weights to 0)."""
    from ultrasound_solver_amd import basis
    gd, ph, _, kw, ref_loss, ref_g, bound = _fused_case("softmax", 20011, 96, 13, temperature)
    loss, g = basis.fused_loss(gd, torch.as_tensor(np.exp(1j*ph), device="cuda:0"), "softmax", **kw)
    assert float(loss) == pytest.approx(ref_loss, rel=1e-12)
    assert _close(g.cpu().numpy(), ref_g, bound, 1e-12)


def test_fused_softmax_single_masked_row():
    """One background row r: the smooth max is |u_r| exactly and c_r = beta u_r/(|u_r| + eps)."""
    from ultrasound_solver_amd import basis
    g, gd = _basis(4097, 20, 17)
    ph = np.random.default_rng(18).uniform(0, 2*np.pi, 20)
    u = g @ np.exp(1j*ph)
    r = 3001
    mask = np.zeros(4097, dtype=np.uint8)
    mask[r] = 1
    loss, gg = basis.fused_loss(gd, torch.as_tensor(np.exp(1j*ph), device="cuda:0"), "softmax",
                                mask=torch.as_tensor(mask, device="cuda:0"), scale=0.5, temperature=3.0, count_a=1)
    assert float(loss) == pytest.approx(0.5*abs(u[r]), rel=1e-13)
    want = g[r].conj()*(0.5*u[r]/(abs(u[r]) + olo.EPS))
    assert _rel(gg.cpu().numpy(), want) < 1e-12


def test_fused_loss_rejects_bad_descriptors():
    from ultrasound_solver_amd import basis
    _, gd = _basis(100, 8, 19)
    w = torch.ones(8, dtype=torch.complex128, device="cuda:0")
    with pytest.raises(RuntimeError, match="mask"):
        basis.fused_loss(gd, w, "softmax", temperature=1.0, count_a=3)
    mask = torch.ones(100, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(RuntimeError, match="temperature"):
        basis.fused_loss(gd, w, "softmax", mask=mask, temperature=0.0, count_a=100)
    with pytest.raises(RuntimeError, match="count"):
        basis.fused_loss(gd, w, "ratio", mask=mask, count_a=60, count_b=60)
