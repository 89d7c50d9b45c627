"""GPU parity of the learned warm start (sim/learned.py) against the PyTorch-CPU
oracle (oracle/unet_ref.py).

BF16 storage: the oracle rounds to bfloat16 at the same points (weights,
input features, activations), so the remaining difference is FP32
accumulation order plus the occasional one-ulp BF16 flip it causes:
convolutions to 1e-2 of scale (max), the whole network to 2e-2 (relative L2).
Gradients: the GPU rounds the back-propagated gradients to BF16 too, and is
compared with FP32 autograd of the unrounded network: 5e-2 relative L2 per
layer.  The warm start keeps the solver's answer: ||b - A p|| <= 1e-6 ||b||
whether or not p0 is accepted.
"""
import ctypes

import numpy as np
import pytest
import torch

from helpers import smooth_velocity

pytestmark = pytest.mark.gpu

N3 = (8, 16, 64)


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def max_scale(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def cl(x):
    """(C, n0, n1, n2) -> channels-last (N, C) bf16 on the GPU."""
    C = x.shape[0]
    return x.reshape(C, -1).t().contiguous().to(torch.bfloat16).cuda()


def from_cl(y, C, n):
    return y.float().cpu().t().reshape((C,) + tuple(n))


@pytest.mark.parametrize("c0,c1,co", [(8, 0, 16), (16, 0, 32), (32, 64, 32), (16, 32, 16), (64, 0, 64)])
def test_conv_forward_and_gradients(dev, c0, c1, co):
    from edge_hip import _lib
    from edge_hip.field import stream_handle
    rng = np.random.default_rng(c0 + c1 + co)
    n = N3
    nh = tuple(x // 2 for x in n)
    x0 = bf(torch.tensor(rng.normal(size=(c0,) + n), dtype=torch.float32))
    x1 = bf(torch.tensor(rng.normal(size=(c1,) + nh), dtype=torch.float32)) if c1 else None
    cin = c0 + c1
    W = bf(torch.tensor(rng.normal(size=(co, 27, cin)) / np.sqrt(27 * cin), dtype=torch.float32))
    bias = torch.tensor(rng.normal(size=co), dtype=torch.float32)
    # oracle: torch conv3d on the same BF16 values
    Wt = W.permute(0, 2, 1).reshape(co, cin, 3, 3, 3)
    xin = x0 if x1 is None else torch.cat([x0, torch.nn.functional.interpolate(x1[None], scale_factor=2,
                                                                                   mode="nearest")[0]], 0)
    xin = xin.clone().requires_grad_(True)
    Wt_ = Wt.clone().requires_grad_(True)
    pre = torch.nn.functional.conv3d(xin[None], Wt_, bias, padding=1)[0]
    ref = torch.relu(pre)
    s = stream_handle()
    gx0, gx1 = cl(x0), (cl(x1) if x1 is not None else None)
    gw = W.reshape(-1).to(torch.bfloat16).cuda()
    gb = bias.cuda()
    y = torch.empty((int(np.prod(n)), co), dtype=torch.bfloat16, device="cuda")
    _lib.call("edge_conv3d_bf16_fwd", gx0.data_ptr(), c0, gx1.data_ptr() if c1 else None, c1, gw.data_ptr(),
              gb.data_ptr(), y.data_ptr(), co, 1, *n, s)
    got = from_cl(y, co, n)
    assert max_scale(got, ref.detach()) < 1e-2
    # gradients: upstream gradient G on the output (through ReLU)
    G = torch.tensor(rng.normal(size=ref.shape), dtype=torch.float32)
    (ref * G).sum().backward()
    gpre = bf(G * (ref.detach() > 0))
    # GPU: data gradient with flipped/transposed weights, weight gradient
    wf = torch.empty_like(gw)
    wt = torch.empty_like(gw)
    wfl = W.reshape(-1).cuda()
    _lib.call("edge_unet_pack_weights", wfl.data_ptr(), co, cin, wf.data_ptr(), wt.data_ptr(), s)
    assert torch.equal(wf.cpu(), gw.cpu())
    gp = cl(gpre)
    g0 = torch.zeros((int(np.prod(n)), c0), dtype=torch.float32, device="cuda")
    g1 = torch.zeros((int(np.prod(nh)), c1), dtype=torch.float32, device="cuda") if c1 else None
    if co % 16 == 0 and cin % 16 == 0:
        _lib.call("edge_conv3d_bf16_bwd_data", gp.data_ptr(), co, wt.data_ptr(), cin, c0, g0.data_ptr(),
                  g1.data_ptr() if c1 else None, *n, s)
        gx_ref = xin.grad
        got0 = g0.cpu().t().reshape((c0,) + n)
        assert rel_l2(got0, gx_ref[:c0]) < 1e-2
        if c1:
            # the upsampling's transpose: sum over each parent's 8 children
            up_ref = gx_ref[c0:].reshape(c1, nh[0], 2, nh[1], 2, nh[2], 2).sum(dim=(2, 4, 6))
            assert rel_l2(g1.cpu().t().reshape((c1,) + nh), up_ref) < 1e-2
    dw = torch.zeros(co * 27 * cin, dtype=torch.float32, device="cuda")
    _lib.call("edge_conv3d_bf16_bwd_weight", gp.data_ptr(), co, gx0.data_ptr(), c0, gx1.data_ptr() if c1 else None,
              c1, dw.data_ptr(), *n, s)
    dw_ref = Wt_.grad.reshape(co, cin, 27).permute(0, 2, 1).reshape(-1)
    assert rel_l2(dw.cpu(), dw_ref) < 1e-2


def test_network_forward_and_backward_match_oracle(dev):
    from sim.learned import UNet
    from oracle import unet_ref as U
    n = (16, 16, 64)
    net = UNet(seed=5)
    # non-trivial biases too
    rng = np.random.default_rng(7)
    with torch.no_grad():
        net.params += 0.02 * torch.tensor(rng.normal(size=net.nparams), dtype=torch.float32, device="cuda")
    net._packed = False
    u = smooth_velocity(*n, 1.0 / 16, seed=4)
    b = (u[0][1:] - u[0][:-1]) + (u[1][:, 1:] - u[1][:, :-1]) + (u[2][:, :, 1:] - u[2][:, :, :-1])
    b = -(b - b.mean())
    tb = torch.tensor(b, dtype=torch.float32)
    p = net.forward(tb.cuda())
    flat = net.params.detach().cpu().to(torch.float32)
    with torch.no_grad():
        pref, scale = U.forward(flat, tb, bf16=True)
    assert rel_l2(p.cpu().reshape(n), pref) < 2e-2
    # backward against FP32 autograd of the unrounded network
    target = torch.tensor(rng.normal(size=n), dtype=torch.float32) * float(scale)
    loss = net.backward(target.cuda().reshape(-1))
    Lref, gref = U.gradients(flat.double(), tb.double(), target.double())
    assert loss == pytest.approx(Lref, rel=5e-2)
    g = net.grads.cpu().double()
    for key, (off, cnt) in net.offsets.items():
        if key.startswith("e0a.w"):
            continue  # compared below without its padding channels (zero gradient on both sides)
        if key == "head.b":
            # the loss ignores constant offsets: zero up to rounding
            assert abs(float(g[off])) < 1e-6 * float(g.abs().max())
            continue
        assert rel_l2(g[off:off + cnt], gref[off:off + cnt]) < 5e-2, key
    off, cnt = net.offsets["e0a.w"]
    a = g[off:off + cnt].reshape(16, 27, 8)[:, :, :2]
    r = gref[off:off + cnt].reshape(16, 27, 8)[:, :, :2]
    assert rel_l2(a, r) < 5e-2
    assert float(g[off:off + cnt].reshape(16, 27, 8)[:, :, 2:].abs().max()) == 0.0


def test_energy_loss_gradient_matches_oracle(dev):
    """The energy-norm loss (1/2 e'Ae / (N rms^2), e = p - p*; its dL/dp is A e
    through the solver's operator) and the network's gradient under it, against
    FP32 autograd of the unrounded CPU network with the restated operator."""
    from sim.learned import UNet
    from sim.mgpcg import MGPCG_3
    from oracle import unet_ref as U
    n = (16, 16, 64)
    net = UNet(seed=6)
    rng = np.random.default_rng(17)
    u = smooth_velocity(*n, 1.0 / 16, seed=8)
    b = (u[0][1:] - u[0][:-1]) + (u[1][:, 1:] - u[1][:, :-1]) + (u[2][:, :, 1:] - u[2][:, :, :-1])
    b = -(b - b.mean())
    tb = torch.tensor(b, dtype=torch.float32)
    p0 = net.forward(tb.cuda()).clone()
    flat = net.params.detach().cpu().to(torch.float32)
    with torch.no_grad():
        _, scale = U.forward(flat, tb, bf16=True)
    # a smooth target, as a converged pressure is (a rough one makes dL/dp = A e
    # white noise, whose products with the activations cancel over the voxels
    # far below the BF16 rounding of the backward chain)
    idx = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    smooth = sum(rng.normal() * np.cos(np.pi * k * (idx[d] + 0.5) / n[d]) for d in range(3) for k in (1, 2))
    target = torch.tensor(smooth, dtype=torch.float32) * float(scale)
    solver = MGPCG_3(np.full((3, 2), 2), list(n), levels=2)
    loss = net.backward(target.cuda().reshape(-1), loss="energy", apply_A=lambda x, out: solver.get_Ap(x, out))
    # dL/dp itself: A (p0 - target) / (N rms^2) with the restated operator, FP64
    e = p0.cpu().double().reshape(n) - target.double()
    want = U.apply_A(e) / (e.numel() * float(net._scale.item()) ** 2)
    got = net._acts["g_p"].cpu().double().reshape(n)
    assert float((got - want).abs().max() / want.abs().max()) < 1e-5
    Lref, gref = U.gradients(flat.double(), tb.double(), target.double(), kind="energy")
    print(f"energy loss GPU {loss:.5g}, oracle {Lref:.5g}")
    assert loss == pytest.approx(Lref, rel=5e-2)
    g = net.grads.cpu().double()
    errs = {}
    for key, (off, cnt) in net.offsets.items():
        if key.startswith("e0a.w") or key == "head.b":
            continue   # padding channels; the constant offset has zero gradient (A 1 = 0)
        errs[key] = rel_l2(g[off:off + cnt], gref[off:off + cnt])
    print("energy-loss gradient, rel. L2 vs FP32 autograd:", {k: round(v, 4) for k, v in errs.items()})
    for key, err in errs.items():
        assert err < 1e-1, key
    off, cnt = net.offsets["head.b"]
    assert abs(float(g[off])) < 1e-5 * float(g.abs().max())


def test_training_reduces_the_loss(dev):
    from sim.learned import UNet
    n = (16, 16, 64)
    rng = np.random.default_rng(8)
    net = UNet(seed=1)
    b = torch.tensor(rng.normal(size=n), dtype=torch.float32, device="cuda")
    # a target the network can represent: another network's output
    target = UNet(seed=9).forward(b).clone()
    losses = []
    for _ in range(30):
        net.forward(b)
        losses.append(net.backward(target.reshape(-1)))
        net.adam_step(lr=3e-3)
    assert losses[-1] < 0.7 * losses[0]


def test_warm_start_keeps_the_solution(dev):
    from edge_hip import Field
    from sim.learned import UNet
    from sim.mgpcg import MGPCG_3
    n = (32, 16, 64)
    u = smooth_velocity(*n, 1.0 / 16, seed=11)

    def fields():
        out = []
        for c in u:
            f = Field(None, float, c.shape)
            f.from_numpy(c.astype(np.float32))
            out.append(f)
        return out

    bt = np.full((3, 2), 2)
    cold = MGPCG_3(bt, list(n), levels=3, mode="mixed", stop="relative", rel_tol=1e-6)
    fc = fields()
    cold.Poisson(*fc)
    warm = MGPCG_3(bt, list(n), levels=3, mode="mixed", stop="relative", rel_tol=1e-6, warm_start=UNet(seed=2))
    fw = fields()
    warm.Poisson(*fw)
    assert warm.stats["converged"] == 1.0
    # the same projected velocity up to the solver tolerance, whichever start was kept
    for a, b_ in zip(fc, fw):
        assert max_scale(b_.to_numpy(), a.to_numpy()) < 1e-4
    # x0 was checked: used iff ||b - A x0||^2 <= ||b||^2
    assert warm.stats["x0_used"] == (1.0 if warm.stats["r02"] <= warm.stats["b2"] else 0.0)
