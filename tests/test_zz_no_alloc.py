"""hz_basis_loss_static (no allocation; partials in a static device array)
against hz_basis_loss through the C ABI: the same kernels and launch shape,
so loss, G^H c and u must agree bit for bit.  Two calls in a row, the
second on a smaller basis, then every loss kind; the static array keeps the
previous call's contents, which a correct call never reads."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _call(entry, G, w, kind, mask=None, target=None, weight=None, **kw):
    from ultrasound_solver_amd import _capi
    n, m = G.shape
    desc = _capi.BasisLossDesc()
    desc.kind = kind
    desc.mask = mask.data_ptr() if mask is not None else None
    desc.target = target.data_ptr() if target is not None else None
    desc.weight = weight.data_ptr() if weight is not None else None
    desc.eps = 1e-12
    desc.scale = kw.get("scale", 0.0)
    desc.temperature = kw.get("temperature", 1.0)
    desc.count_a = kw.get("count_a", 0)
    desc.count_b = kw.get("count_b", 0)
    loss = torch.full((1,), float("nan"), dtype=torch.float64, device=G.device)
    g = torch.full((m,), float("nan"), dtype=torch.complex128, device=G.device)
    u = torch.full((n,), float("nan"), dtype=torch.complex128, device=G.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    status = getattr(_capi.lib(), entry)(ctypes.c_void_p(G.data_ptr()), m, n, m, ctypes.c_void_p(w.data_ptr()),
                                         ctypes.byref(desc), ctypes.c_void_p(loss.data_ptr()),
                                         ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(u.data_ptr()), stream)
    _capi.check(status, entry)
    torch.cuda.synchronize()
    return loss, g, u


def _case(n, m, seed):
    rng = np.random.default_rng(seed)
    G = torch.as_tensor(rng.standard_normal((n, m)) + 1j*rng.standard_normal((n, m)), device="cuda:0")
    w = torch.as_tensor(np.exp(1j*rng.uniform(0, 2*np.pi, m)), device="cuda:0")
    return rng, G, w


def _same(kind, G, w, **kw):
    a = _call("hz_basis_loss", G, w, kind, **kw)
    b = _call("hz_basis_loss_static", G, w, kind, **kw)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    assert torch.isfinite(torch.view_as_real(b[1])).all() and torch.isfinite(b[0]).all()
    return b


def test_static_scratch_matches_allocating_entry_on_two_calls():
    from ultrasound_solver_amd import _capi
    print("no-alloc: softmax, 200003 x 256, then energy, 3001 x 64", flush=True)
    rng, G, w = _case(200003, 256, 1)
    mask = torch.as_tensor((rng.random(200003) < 0.6).astype(np.uint8), device="cuda:0")
    first = _same(_capi.LOSS_SOFTMAX, G, w, mask=mask, scale=1.3, temperature=4.0,
                  count_a=int(mask.sum()))
    _, G2, w2 = _case(3001, 64, 2)
    second = _same(_capi.LOSS_ENERGY, G2, w2, scale=0.5/3001)
    u_ref = (G2 @ w2).cpu().numpy()
    np.testing.assert_allclose(second[2].cpu().numpy(), u_ref, rtol=1e-12, atol=1e-12*np.abs(u_ref).max())
    assert float(second[0]) == pytest.approx(0.25/3001*np.sum(np.abs(u_ref)**2), rel=1e-12)
    assert float(first[0]) != float(second[0])


def test_static_scratch_every_kind():
    from ultrasound_solver_amd import _capi
    print("no-alloc: every loss kind at 20011 x 512 and 97 x 33", flush=True)
    for n, m, seed in ((20011, 512, 3), (97, 33, 4)):
        rng, G, w = _case(n, m, seed)
        target = torch.as_tensor(rng.uniform(0, 20, n), device="cuda:0")
        weight = torch.as_tensor(rng.uniform(0, 1, n), device="cuda:0")
        _same(_capi.LOSS_FIELD_MATCH, G, w, target=target, weight=weight)
        _same(_capi.LOSS_ENERGY, G, w, scale=1.0/n)
        mask = torch.as_tensor((rng.random(n) < 0.5).astype(np.uint8), device="cuda:0")
        mask[0] = 1
        _same(_capi.LOSS_SOFTMAX, G, w, mask=mask, scale=0.7, temperature=2.0, count_a=int(mask.sum()))
        cls = rng.integers(0, 3, n).astype(np.uint8)
        cls[0], cls[1] = 1, 2
        _same(_capi.LOSS_RATIO, G, w, mask=torch.as_tensor(cls, device="cuda:0"),
              count_a=int((cls == 1).sum()), count_b=int((cls == 2).sum()))


def test_static_scratch_rejects_wide_rows():
    from ultrasound_solver_amd import _capi
    _, G, w = _case(64, 513, 5)
    with pytest.raises(_capi.HzError, match="512"):
        _call("hz_basis_loss_static", G, w, _capi.LOSS_ENERGY, scale=1.0)
