"""Pins for the NumPy U-Net oracle: PyTorch CPU autograd (an independent
implementation of the same network) and central finite differences."""
import numpy as np
import pytest

import unet_ref as U


def _case(seed, B=2, H=8, W=12):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    x = np.zeros((B, H, W, 16))
    x[..., :6] = rng.standard_normal((B, H, W, 6))
    t = rng.standard_normal((B, H, W)) * 0.1
    return p, x, t


def test_gradients_match_torch_autograd():
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    p, x, t = _case(0)
    L, g = U.loss_and_grads(p, x, t)
    ws = [(torch.tensor(w, requires_grad=True), torch.tensor(b, requires_grad=True)) for w, b in U.unflatten(p)]

    def lay(i, xx):
        w, b = ws[i]
        k = w.reshape(3, 3, w.shape[1], w.shape[2]).permute(2, 3, 0, 1)
        z = F.conv2d(xx.permute(0, 3, 1, 2), k, padding=1).permute(0, 2, 3, 1) + b
        return torch.relu(z) if U.LAYERS[i][4] else z

    pool = lambda a: F.avg_pool2d(a.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)
    up = lambda a: a.repeat_interleave(2, 1).repeat_interleave(2, 2)
    xt = torch.tensor(x)
    e1 = lay(1, lay(0, xt))
    e2 = lay(3, lay(2, pool(e1)))
    m = lay(5, lay(4, pool(e2)))
    d2 = lay(7, lay(6, torch.cat([up(m), e2], -1)))
    d1 = lay(9, lay(8, torch.cat([up(d2), e1], -1)))
    out = lay(10, d1)[..., 0]
    loss = ((out - torch.tensor(t)) ** 2).mean()
    loss.backward()
    tg = U.flatten_grads([(w.grad.numpy(), b.grad.numpy()) for w, b in ws])
    assert abs(loss.item() - L) < 1e-12 * L
    assert np.max(np.abs(tg - g)) < 1e-12 * np.max(np.abs(tg))


def test_gradients_match_finite_differences():
    p, x, t = _case(1)
    _, g = U.loss_and_grads(p, x, t)
    rng = np.random.default_rng(7)
    idx = rng.choice(len(p), 25, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (U.loss_and_grads(pp, x, t)[0] - U.loss_and_grads(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
