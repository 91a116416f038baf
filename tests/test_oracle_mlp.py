"""Pins of the surrogate's oracle (CPU): the hand-written backpropagation
against central differences of the loss and against torch.autograd, the
BF16 rounding against torch's cast."""
import numpy as np

from oracle import mlp


def _case(dim=3, width=8, hidden=2, n=5, seed=0):
    lay, tot = mlp.layout(dim, width, hidden)
    rng = np.random.default_rng(seed)
    p = rng.uniform(-0.5, 0.5, tot)
    x = rng.uniform(-1, 1, (n, dim))
    t = rng.uniform(-0.3, 0.3, n)
    sh = np.array([0.1, -0.2, 0.05])[:dim]
    sc = np.full(dim, 1.3)
    return p, x, t, sh, sc


def test_gradient_against_central_differences():
    for dim, lam, trunc in ((2, 0.0, 0.0), (3, 0.1, 0.0), (3, 0.1, 0.2)):
        p, x, t, sh, sc = _case(dim=dim)
        if trunc:
            t = np.clip(t, -trunc, trunc)  # some targets sit at the truncation
        L, g = mlp.loss_and_grad(p, dim, 8, 2, x, t, sh, sc, lam=lam, eik_h=1e-2, trunc=trunc)
        fd = np.empty_like(p)
        for i in range(len(p)):
            e = np.zeros_like(p)
            e[i] = 1e-6
            fd[i] = (mlp.loss(p + e, dim, 8, 2, x, t, sh, sc, lam, 1e-2, trunc)
                     - mlp.loss(p - e, dim, 8, 2, x, t, sh, sc, lam, 1e-2, trunc)) / 2e-6
        assert np.abs(fd - g).max() <= 1e-7 * max(1.0, np.abs(g).max())


def test_gradient_against_torch_autograd():
    import torch
    dim, W, H, n, h = 3, 8, 2, 5, 1e-2
    p, x, t, sh, sc = _case(dim=dim, width=W, hidden=H, n=n)
    L, g = mlp.loss_and_grad(p, dim, W, H, x, t, sh, sc, lam=0.1, eik_h=h)
    lay, _ = mlp.layout(dim, W, H)
    pt = torch.tensor(p, requires_grad=True)

    def seg(k):
        o, m = lay[k]
        return pt[o:o + m]

    X = torch.tensor(mlp.make_inputs(x, sh, sc, h))
    z = X @ seg("w_in").view(W, dim).T + seg("b_in")
    a = z * torch.sigmoid(z)
    for l in range(H):
        z = a @ seg("w_hid").view(H, W, W)[l].T + seg("b_hid").view(H, W)[l]
        a = z * torch.sigmoid(z)
    y = a @ seg("w_out") + seg("b_out")[0]
    loss = ((y[:n] - torch.tensor(t)) ** 2).mean()
    gg = torch.stack([(y[(2 * k + 1) * n:(2 * k + 2) * n] - y[(2 * k + 2) * n:(2 * k + 3) * n])
                      / (2 * h) for k in range(dim)], 1)
    loss = loss + 0.1 * ((gg.norm(dim=1) - 1) ** 2).mean()
    loss.backward()
    assert abs(float(loss.detach()) - L) < 1e-12
    assert np.abs(pt.grad.numpy() - g).max() < 1e-12


def test_bf16_round_matches_torch():
    import torch
    a = np.random.default_rng(1).standard_normal(4096).astype(np.float32) * 10.0
    a[:4] = [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -0.0, 65504.0]  # ties to even
    ref = torch.tensor(a).to(torch.bfloat16).float().numpy()
    assert np.array_equal(mlp.bf16_round(a), ref)
