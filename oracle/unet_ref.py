"""CPU restatement of the learned warm start's U-Net with PyTorch library ops.

TEST INFRASTRUCTURE ONLY: the checker for the GPU network of
edge-flow-maps_amd/sim/learned.py (which has no counterpart in the reference,
so this restatement is the definition it is held to).  torch.nn.functional
conv3d (cross-correlation, zero padding 1), avg_pool3d(2), nearest
interpolation x2 and autograd supply every operation.

The flat parameter vector has the GPU layout: per layer W[co][tap][ci]
(tap = (di*3 + dj)*3 + dk) then bias[co]; then the head's w[16] and bias[1].
`bf16=True` rounds to bfloat16 where the GPU path stores BF16 (weights, the
input features, every activation).
"""
import numpy as np
import torch
import torch.nn.functional as Fn

LAYERS = [
    ("e0a", 8, 0, 16), ("e0b", 16, 0, 16), ("e1a", 16, 0, 32), ("e1b", 32, 0, 32), ("ba", 32, 0, 64),
    ("bb", 64, 0, 64), ("d1a", 32, 64, 32), ("d1b", 32, 0, 32), ("d0a", 16, 32, 16), ("d0b", 16, 0, 16),
]
HEAD_C = 16


def split_params(flat):
    """dict name -> tensor views (conv weights as torch's [co][ci][3][3][3])."""
    out, off = {}, 0
    for name, c0, c1, co in LAYERS:
        cin = c0 + c1
        n = co * 27 * cin
        w = flat[off:off + n].reshape(co, 27, cin)
        out[name + ".w"] = w.permute(0, 2, 1).reshape(co, cin, 3, 3, 3)
        off += n
        out[name + ".b"] = flat[off:off + co]
        off += co
    out["head.w"] = flat[off:off + HEAD_C]
    out["head.b"] = flat[off + HEAD_C:off + HEAD_C + 1]
    assert off + HEAD_C + 1 == flat.numel()
    return out


def nparams():
    return sum(co * 27 * (c0 + c1) + co for _, c0, c1, co in LAYERS) + HEAD_C + 1


def forward(flat, b, mask=None, bf16=False):
    """p (shape of b) for the right-hand side b (3-D tensor); mask: 1 = solid."""
    dt = flat.dtype

    def rnd(t):
        return t.to(torch.bfloat16).to(dt) if bf16 else t

    P = split_params(flat)
    scale = torch.sqrt(torch.mean(b.detach().double() ** 2)).to(dt)
    if float(scale) == 0.0:
        scale = torch.ones((), dtype=dt)
    x = torch.zeros((1, 8) + tuple(b.shape), dtype=dt)
    x[0, 0] = b / scale
    x[0, 1] = 1.0 if mask is None else (mask <= 0).to(dt)
    x = rnd(x)

    def conv(name, t):
        return rnd(torch.relu(Fn.conv3d(t, rnd(P[name + ".w"]), P[name + ".b"], padding=1)))

    def up(t):
        return Fn.interpolate(t, scale_factor=2, mode="nearest")

    e0a = conv("e0a", x)
    s0 = conv("e0b", e0a)
    p0 = rnd(Fn.avg_pool3d(s0, 2))
    e1a = conv("e1a", p0)
    s1 = conv("e1b", e1a)
    p1 = rnd(Fn.avg_pool3d(s1, 2))
    ba = conv("ba", p1)
    bb = conv("bb", ba)
    d1a = conv("d1a", torch.cat([s1, up(bb)], dim=1))
    d1b = conv("d1b", d1a)
    d0a = conv("d0a", torch.cat([s0, up(d1b)], dim=1))
    d0b = conv("d0b", d0a)
    h = torch.einsum("c,ncijk->nijk", P["head.w"], d0b)[0] + P["head.b"][0]
    return scale * h, scale


def loss(p, target, scale):
    """Mean-free squared error in units of the input's RMS."""
    q = (p - target) / scale
    q = q - q.mean()
    return torch.mean(q * q)


def apply_A(e):
    """The projection's operator with walls all round (sim/mgpcg.py:87-117):
    (6 - #box faces) e - sum of the in-box neighbours, by tensor slicing."""
    out = 6.0 * e
    for a in range(3):
        n = e.shape[a]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a], hi[a] = slice(0, n - 1), slice(1, n)
        out = out.clone()
        out[tuple(lo)] -= e[tuple(hi)]
        out[tuple(hi)] -= e[tuple(lo)]
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[a], last[a] = slice(0, 1), slice(n - 1, n)
        out[tuple(first)] -= e[tuple(first)]
        out[tuple(last)] -= e[tuple(last)]
    return out


def energy_loss(p, target, scale):
    """The error's energy norm 1/2 e'Ae / (N scale^2), e = p - target."""
    e = p - target
    return 0.5 * torch.sum(e * apply_A(e)) / (e.numel() * scale * scale)


def gradients(flat, b, target, mask=None, kind="mse"):
    """(loss, dloss/dflat) by autograd in flat's dtype (no BF16 rounding);
    kind "mse" (loss) or "energy" (energy_loss)."""
    f = flat.clone().requires_grad_(True)
    p, scale = forward(f, b, mask)
    L = (loss if kind == "mse" else energy_loss)(p, target, scale.detach())
    L.backward()
    return float(L), f.grad.detach()


def init_params(seed=0, dtype=torch.float64):
    """The GPU network's initialisation (He normal, zero biases, first layer's padding channels zero)."""
    rng = np.random.default_rng(seed)
    out = []
    for name, c0, c1, co in LAYERS:
        cin = c0 + c1
        w = rng.normal(0.0, np.sqrt(2.0 / (27 * cin)), size=(co, 27, cin))
        if name == "e0a":
            w[:, :, 2:] = 0.0
        out += [w.ravel(), np.zeros(co)]
    out += [rng.normal(0.0, np.sqrt(1.0 / HEAD_C), size=HEAD_C), np.zeros(1)]
    return torch.tensor(np.concatenate(out), dtype=dtype)
