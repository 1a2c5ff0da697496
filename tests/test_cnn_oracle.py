"""Pinning of the FP64 numpy CNN oracle (no GPU): its forward against
torch's conv2d/linear in float64, its hand-derived backward against torch
autograd of the same loss, and its loss gradient against central finite
differences."""
import numpy as np
import pytest

import cnn_oracle as co

C = 3


def _torch():
    """torch is imported inside the tests that use it, never at collection:
    its wheel bundles its own HIP runtime under the same soname as the ROCm
    install's, and a GPU test session that collected this module would
    otherwise bind the expectation library to that runtime."""
    return pytest.importorskip("torch")


def _rand_problem(B=3, seed=0):
    rng = np.random.default_rng(seed)
    shapes = co.layout(C)
    arrs = []
    for s in shapes:
        fan = int(np.prod(s[1:])) if len(s) > 1 else 1
        arrs.append(rng.normal(0, 1.0 / np.sqrt(fan), s) if len(s) > 1 else rng.normal(0, 0.1, s))
    p = co.pack(arrs)
    x = rng.normal(size=(B, 1, 64, 64))
    q = rng.normal(size=(B, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.normal(0, 2, size=(B, 2))
    return p, x, q, t


def _torch_loss(p, x, q, t):
    torch = _torch()
    F = torch.nn.functional
    W = [torch.tensor(a, requires_grad=True) for a in co.unpack(p, C)]
    h = torch.tensor(x)
    for l, (ci, hi, c_o, ho, ks, st, pd) in enumerate(co.CONV):
        h = F.relu(F.conv2d(h, W[2 * l], W[2 * l + 1], stride=st, padding=pd))
    f = h.reshape(h.shape[0], -1)
    z = F.relu(f @ W[8].T + W[9])
    out = z @ W[10].T + W[11]
    B = out.shape[0]
    raw = out[:, :6 * C].reshape(B, C, 6)
    m = raw[:, :, :4] / (raw[:, :, :4].norm(dim=2, keepdim=True) + 1e-12)
    qq = torch.tensor(q)
    d = (m * qq[:, None, :]).sum(2)
    k = torch.exp(-(1.0 + 6.0 * torch.sigmoid(raw[:, :, 4])))
    s = d * d + (1 - d * d) / k
    logA = -np.log(2 * np.pi ** 2) - 1.5 * torch.log(k) - 2 * torch.log(s)
    logpi = torch.log_softmax(raw[:, :, 5], dim=1)
    lrot = -torch.logsumexp(logpi + logA, dim=1)
    sig = torch.clamp(torch.exp(out[:, 6 * C + 2]), 0.2, 20.0)
    tt = torch.tensor(t)
    r2 = ((tt - out[:, 6 * C:6 * C + 2]) ** 2).sum(1)
    lt = np.log(2 * np.pi) + 2 * torch.log(sig) + r2 / (2 * sig * sig)
    loss = (lrot + lt).mean()
    loss.backward()
    return loss.item(), np.concatenate([w.grad.numpy().ravel() for w in W]), out.detach().numpy()


def test_forward_matches_torch():
    p, x, q, t = _rand_problem()
    out, _ = co.forward(p, x, C)
    _, _, tout = _torch_loss(p, x, q, t)
    np.testing.assert_allclose(out, tout, rtol=1e-10, atol=1e-10)


def test_backward_matches_torch_autograd():
    p, x, q, t = _rand_problem(seed=1)
    loss, g, _ = co.loss_grad(p, x, q, t, C)
    tl, tg, _ = _torch_loss(p, x, q, t)
    assert loss == pytest.approx(tl, rel=1e-12)
    np.testing.assert_allclose(g, tg, rtol=1e-8, atol=1e-10 * np.abs(tg).max())


def test_loss_gradient_finite_differences():
    p, x, q, t = _rand_problem(B=2, seed=2)
    loss, g, _ = co.loss_grad(p, x, q, t, C)
    rng = np.random.default_rng(3)
    # one parameter of every tensor plus a few random ones
    offs = np.cumsum([0] + [int(np.prod(s)) for s in co.layout(C)])
    idx = list(offs[:-1]) + list(rng.integers(0, p.size, 8))
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (co.loss_grad(pp, x, q, t, C)[0] - co.loss_grad(pm, x, q, t, C)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8)
