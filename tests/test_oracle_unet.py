"""Pin the U-Net oracle (oracle/unet_ref.py): the flat weight layout against a
brute-force convolution from its definition, and the gradient against central
finite differences along random directions (FP64).  CPU only."""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U


def brute_conv(W, x):
    """out[co,i,j,k] = sum_tap sum_ci W[co,tap,ci] x[ci, i+di-1, j+dj-1, k+dk-1], zero outside."""
    co_n, _, ci_n = W.shape
    n = x.shape[1:]
    xp = np.zeros((ci_n,) + tuple(s + 2 for s in n))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    out = np.zeros((co_n,) + n)
    for tap in range(27):
        di, dj, dk = tap // 9, (tap // 3) % 3, tap % 3
        sl = xp[:, di:di + n[0], dj:dj + n[1], dk:dk + n[2]]
        out += np.einsum("oc,cijk->oijk", W[:, tap, :], sl)
    return out


def test_flat_layout_is_the_defined_convolution():
    rng = np.random.default_rng(0)
    flat = U.init_params(seed=3)
    P = U.split_params(flat)
    # layer d1a (96 -> 32): its slice of the flat vector, read as W[co][tap][ci]
    off = 0
    for name, c0, c1, co in U.LAYERS:
        if name == "d1a":
            break
        off += co * 27 * (c0 + c1) + co
    W = flat[off:off + 32 * 27 * 96].numpy().reshape(32, 27, 96)
    x = rng.normal(size=(96, 4, 5, 6))
    got = torch.nn.functional.conv3d(torch.tensor(x)[None], P["d1a.w"], padding=1)[0].numpy()
    np.testing.assert_allclose(got, brute_conv(W, x), rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    torch.manual_seed(0)
    rng = np.random.default_rng(1)
    n = (8, 8, 8)
    flat = U.init_params(seed=2) + 0.01 * torch.tensor(rng.normal(size=U.nparams()))
    b = torch.tensor(rng.normal(size=n))
    target = torch.tensor(rng.normal(size=n))
    L0, g = U.gradients(flat, b, target)
    for t in range(4):
        v = torch.tensor(rng.normal(size=U.nparams()))
        v /= torch.linalg.norm(v)   # a unit step: few ReLU kinks within +-eps
        eps = 1e-7
        lp = float(_loss(flat + eps * v, b, target))
        lm = float(_loss(flat - eps * v, b, target))
        fd = (lp - lm) / (2 * eps)
        an = float(g @ v)
        assert abs(fd - an) <= 1e-5 * max(1e-3, abs(an)), (fd, an)


def _loss(flat, b, target):
    p, scale = U.forward(flat, b)
    return U.loss(p, target, scale)


def test_zero_network_gives_the_head_bias():
    flat = torch.zeros(U.nparams(), dtype=torch.float64)
    flat[-1] = 0.5   # head bias
    b = torch.tensor(np.random.default_rng(3).normal(size=(8, 8, 8)))
    p, scale = U.forward(flat, b)
    np.testing.assert_allclose(p.numpy(), 0.5 * float(scale))
    np.testing.assert_allclose(float(scale), float(torch.sqrt(torch.mean(b ** 2))))


def test_energy_loss_operator_and_eigenvector():
    """The energy loss's operator is the projection's 7-point operator with
    walls all round: it equals the oracle's restated operator on a random
    field, and on a discrete cosine mode cos(pi k (i + 1/2) / n) per axis, an
    eigenvector with eigenvalue sum(2 - 2 cos(pi k_d / n_d)), the loss is half
    that eigenvalue times |v|^2 / N.  A missing wall term or a dropped 1/2
    fails."""
    import torch
    from oracle import np_ref
    from oracle import unet_ref as U
    n = (6, 5, 7)
    rng = np.random.default_rng(3)
    e = rng.normal(size=n)
    ns, dg = np_ref._nsum_diag(e)
    np.testing.assert_allclose(U.apply_A(torch.tensor(e)).numpy(), dg * e - ns, atol=1e-12)
    k = (2, 1, 3)
    idx = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    v = np.ones(n)
    lam = 0.0
    for d in range(3):
        v = v * np.cos(np.pi * k[d] * (idx[d] + 0.5) / n[d])
        lam += 2 - 2 * np.cos(np.pi * k[d] / n[d])
    got = float(U.energy_loss(torch.tensor(v), torch.zeros(n, dtype=torch.float64), torch.tensor(1.0)))
    assert got == pytest.approx(0.5 * lam * (v ** 2).sum() / v.size, rel=1e-12)
