"""Learned warm start for the pressure projection (no counterpart in the reference).

A 3-level 3D U-Net maps the projection's right-hand side b (scaled by its
RMS) and the fluid mask to a pressure guess p0; MGPCG then starts from p0 and
keeps it only if ||b - A p0|| <= ||b||, so the converged answer and the stop
rule are those of the zero start.  Every layer is a libedge_hip.so kernel
(BF16 MFMA convolutions with FP32 accumulation, channels-last activations);
torch only holds the device memory.

    encoder  e0a 8->16, e0b 16->16 (s0) | pool | e1a 16->32, e1b 32->32 (s1) | pool
    bottom   ba 32->64, bb 64->64
    decoder  d1a [s1, up(bb)] 96->32, d1b 32->32 | d0a [s0, up(d1b)] 48->16, d0b 16->16
    head     1x1 16->1 (FP32), times the input's RMS

All 3x3x3 layers use ReLU.  Input channel 0 is b / rms(b), channel 1 the fluid
mask, channels 2..7 zero (the first layer's input is padded to 8 channels).
Training (`train`) uses online (b, p) pairs from the step loop itself, with
Adam on FP32 master weights; the loss is the energy norm of the guess's error
(1/2 e'Ae, e = p0 - p, whose gradient is the guess's negative residual), or
the mean-free squared error of p0 against the converged pressure.
"""
import argparse
import math
import os
import sys
import time

import numpy as np
import torch

from edge_hip import _lib
from edge_hip.field import as_ptr, stream_handle

LAYERS = [  # name, input channels (C0, C1: same resolution, half resolution), output channels, level
    ("e0a", 8, 0, 16, 0),
    ("e0b", 16, 0, 16, 0),
    ("e1a", 16, 0, 32, 1),
    ("e1b", 32, 0, 32, 1),
    ("ba", 32, 0, 64, 2),
    ("bb", 64, 0, 64, 2),
    ("d1a", 32, 64, 32, 1),
    ("d1b", 32, 0, 32, 1),
    ("d0a", 16, 32, 16, 0),
    ("d0b", 16, 0, 16, 0),
]
HEAD_C = 16
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "unet_warmstart.npz")


def _ptr(t):
    return None if t is None else t.data_ptr()


class UNet:
    """The network, its parameters (one flat FP32 vector) and, for training, Adam state."""

    def __init__(self, seed=0, device="cuda"):
        self.device = torch.device(device)
        sizes = []
        for name, c0, c1, co, _ in LAYERS:
            sizes += [(name + ".w", co * 27 * (c0 + c1)), (name + ".b", co)]
        sizes += [("head.w", HEAD_C), ("head.b", 1)]
        self.offsets = {}
        off = 0
        for k, n in sizes:
            self.offsets[k] = (off, n)
            off += n
        self.nparams = off
        self.params = torch.zeros(off, dtype=torch.float32, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.m1 = self.m2 = None
        self.t = 0
        self._init(seed)
        self._wf = {}
        self._wt = {}
        for name, c0, c1, co, _ in LAYERS:
            n = co * 27 * (c0 + c1)
            self._wf[name] = torch.empty(n, dtype=torch.bfloat16, device=self.device)
            self._wt[name] = torch.empty(n, dtype=torch.bfloat16, device=self.device)
        self._packed = False
        self._work = torch.zeros(4, dtype=torch.float64, device=self.device)
        self._scale = torch.ones(1, dtype=torch.float32, device=self.device)
        self._acts = {}

    # -- parameters -------------------------------------------------------------------
    def view(self, key, grad=False):
        off, n = self.offsets[key]
        return (self.grads if grad else self.params)[off:off + n]

    def _init(self, seed):
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparams, np.float32)
        for name, c0, c1, co, _ in LAYERS:
            cin = c0 + c1
            w = rng.normal(0.0, math.sqrt(2.0 / (27 * cin)), size=(co, 27, cin)).astype(np.float32)
            if name == "e0a":
                w[:, :, 2:] = 0.0   # padding channels
            off, n = self.offsets[name + ".w"]
            p[off:off + n] = w.ravel()
        off, n = self.offsets["head.w"]
        p[off:off + n] = rng.normal(0.0, math.sqrt(1.0 / HEAD_C), size=n)
        self.params.copy_(torch.from_numpy(p))

    def pack(self):
        """BF16 forward and data-gradient weights from the FP32 master copy."""
        s = stream_handle()
        for name, c0, c1, co, _ in LAYERS:
            _lib.call("edge_unet_pack_weights", _ptr(self.view(name + ".w")), co, c0 + c1, _ptr(self._wf[name]),
                      _ptr(self._wt[name]), s)
        self._packed = True

    def save(self, path):
        np.savez(path, params=self.params.detach().cpu().numpy(), nparams=self.nparams)

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        if int(z["nparams"]) != self.nparams:
            raise ValueError(f"{path}: {int(z['nparams'])} parameters, this network has {self.nparams}")
        self.params.copy_(torch.from_numpy(z["params"].astype(np.float32)))
        self._packed = False
        return self

    # -- forward ------------------------------------------------------------------------
    def _buf(self, key, shape, dtype):
        t = self._acts.get(key)
        if t is None or t.shape != shape or t.dtype != dtype:
            t = torch.empty(shape, dtype=dtype, device=self.device)
            self._acts[key] = t
        return t

    @staticmethod
    def check_dims(n):
        n0, n1, n2 = n
        if n0 % 4 or n1 % 4 or n2 % 64:
            raise ValueError(f"U-Net needs nx, ny multiples of 4 and nz of 64, got {tuple(n)}")

    def forward(self, b, bm=None, n=None):
        """p0 (FP32, n cells) for the right-hand side b (FP32 tensor of n = (nx, ny, nz) cells)."""
        n = tuple(n if n is not None else b.shape)
        self.check_dims(n)
        if not self._packed:
            self.pack()
        s = stream_handle()
        dims = [n, tuple(x // 2 for x in n), tuple(x // 4 for x in n)]
        N = [int(np.prod(d)) for d in dims]
        bf = torch.bfloat16
        A = {}
        A["x"] = self._buf("x", (N[0], 8), bf)
        _lib.call("edge_unet_input", _ptr(b), as_ptr(bm), N[0], _ptr(A["x"]), _ptr(self._scale), _ptr(self._work), s)

        def conv(name, src0, src1=None):
            _, c0, c1, co, lev = next(L for L in LAYERS if L[0] == name)
            y = self._buf(name, (N[lev], co), bf)
            _lib.call("edge_conv3d_bf16_fwd", _ptr(src0), c0, _ptr(src1), c1, _ptr(self._wf[name]),
                      _ptr(self.view(name + ".b")), _ptr(y), co, 1, *dims[lev], s)
            A[name] = y
            return y

        def pool(x, C, lev, key):
            y = self._buf(key, (N[lev + 1], C), bf)
            _lib.call("edge_unet_pool", _ptr(x), _ptr(y), C, *dims[lev], s)
            A[key] = y
            return y

        conv("e0a", A["x"])
        s0 = conv("e0b", A["e0a"])
        p0 = pool(s0, 16, 0, "p0")
        conv("e1a", p0)
        s1 = conv("e1b", A["e1a"])
        p1 = pool(s1, 32, 1, "p1")
        conv("ba", p1)
        bb = conv("bb", A["ba"])
        conv("d1a", s1, bb)
        d1b = conv("d1b", A["d1a"])
        conv("d0a", s0, d1b)
        d0b = conv("d0b", A["d0a"])
        p = self._buf("p", (N[0],), torch.float32)
        _lib.call("edge_unet_head", _ptr(d0b), _ptr(self.view("head.w")), _ptr(self.view("head.b")),
                  _ptr(self._scale), N[0], HEAD_C, _ptr(p), s)
        self._fwd = dict(A=A, dims=dims, N=N)
        return p

    # -- backward -----------------------------------------------------------------------
    def backward(self, target, loss="mse", apply_A=None):
        """Gradients (into self.grads) of the last forward's p against the
        converged pressure `target`; returns the loss (host float).
        loss="mse": the mean-free squared error (in units of the input's RMS).
        loss="energy": the error's energy norm 1/2 e'Ae / (N rms^2), e = p - target,
        whose gradient A e / (N rms^2) is minus the guess's residual b - A p: what
        the solver's stop rule and CG's error bound measure.  `apply_A(x, out)`
        applies the projection's operator (the solver's edge_mg_apply_A)."""
        A, dims, N = self._fwd["A"], self._fwd["dims"], self._fwd["N"]
        s = stream_handle()
        f32, bf = torch.float32, torch.bfloat16
        _lib.call("edge_zero_f32", _ptr(self.grads), self.nparams, s)
        G = {}

        def gbuf(key, lev, C):
            t = self._buf("g_" + key, (N[lev], C), f32)
            _lib.call("edge_zero_f32", _ptr(t), t.numel(), s)
            G[key] = t
            return t

        for key, lev, C in (("e0a", 0, 16), ("e0b", 0, 16), ("p0", 1, 16), ("e1a", 1, 32), ("e1b", 1, 32),
                            ("p1", 2, 32), ("ba", 2, 64), ("bb", 2, 64), ("d1a", 1, 32), ("d1b", 1, 32),
                            ("d0a", 0, 16), ("d0b", 0, 16)):
            gbuf(key, lev, C)
        dp = self._buf("g_p", (N[0],), f32)
        if loss == "mse":
            _lib.call("edge_unet_loss_head_bwd", _ptr(self._acts["p"]), _ptr(target), _ptr(self._scale),
                      _ptr(A["d0b"]), _ptr(self.view("head.w")), N[0], HEAD_C, _ptr(dp), _ptr(G["d0b"]),
                      _ptr(self.view("head.w", True)), _ptr(self.view("head.b", True)), _ptr(self._work), s)
            value = None
        elif loss == "energy":
            if apply_A is None:
                raise ValueError("loss='energy' needs apply_A")
            e = self._buf("g_e", (N[0],), f32)
            Ae = self._buf("g_Ae", (N[0],), f32)
            _lib.call("edge_axpy", _ptr(self._acts["p"]), _ptr(target), _ptr(e), -1.0, N[0], s)   # e = p - p*
            apply_A(e, Ae)
            sc = float(self._scale.item())
            norm = 1.0 / (N[0] * sc * sc)
            _lib.call("edge_scale", _ptr(Ae), norm, _ptr(dp), N[0], s)
            _lib.call("edge_unet_head_bwd", _ptr(dp), _ptr(self._scale), _ptr(A["d0b"]), _ptr(self.view("head.w")),
                      N[0], HEAD_C, _ptr(G["d0b"]), _ptr(self.view("head.w", True)),
                      _ptr(self.view("head.b", True)), s)
            value = 0.5 * norm * float(torch.dot(e, Ae).item())
        else:
            raise ValueError(f"loss must be 'mse' or 'energy', got {loss!r}")

        def layer_bwd(name, src0, src1, g_out, g0, g1):
            _, c0, c1, co, lev = next(L for L in LAYERS if L[0] == name)
            gp = self._buf("gp_" + name, (N[lev], co), bf)
            _lib.call("edge_unet_relu_bwd", _ptr(g_out), _ptr(A[name]), _ptr(gp), _ptr(self.view(name + ".b", True)),
                      N[lev], co, 1, s)
            _lib.call("edge_conv3d_bf16_bwd_weight", _ptr(gp), co, _ptr(src0), c0, _ptr(src1), c1,
                      _ptr(self.view(name + ".w", True)), *dims[lev], s)
            if g0 is not None:
                _lib.call("edge_conv3d_bf16_bwd_data", _ptr(gp), co, _ptr(self._wt[name]), c0 + c1, c0, _ptr(g0),
                          _ptr(g1), *dims[lev], s)

        layer_bwd("d0b", A["d0a"], None, G["d0b"], G["d0a"], None)
        layer_bwd("d0a", A["e0b"], A["d1b"], G["d0a"], G["e0b"], G["d1b"])
        layer_bwd("d1b", A["d1a"], None, G["d1b"], G["d1a"], None)
        layer_bwd("d1a", A["e1b"], A["bb"], G["d1a"], G["e1b"], G["bb"])
        layer_bwd("bb", A["ba"], None, G["bb"], G["ba"], None)
        layer_bwd("ba", A["p1"], None, G["ba"], G["p1"], None)
        _lib.call("edge_unet_pool_bwd", _ptr(G["p1"]), _ptr(G["e1b"]), 32, *dims[1], s)
        layer_bwd("e1b", A["e1a"], None, G["e1b"], G["e1a"], None)
        layer_bwd("e1a", A["p0"], None, G["e1a"], G["p0"], None)
        _lib.call("edge_unet_pool_bwd", _ptr(G["p0"]), _ptr(G["e0b"]), 16, *dims[0], s)
        layer_bwd("e0b", A["e0a"], None, G["e0b"], G["e0a"], None)
        layer_bwd("e0a", A["x"], None, G["e0a"], None, None)
        return float(self._work[3].item()) if value is None else value

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        if self.m1 is None:
            self.m1 = torch.zeros_like(self.params)
            self.m2 = torch.zeros_like(self.params)
        self.t += 1
        _lib.call("edge_adam", _ptr(self.params), _ptr(self.grads), _ptr(self.m1), _ptr(self.m2), self.nparams,
                  float(lr), float(b1), float(b2), float(eps), float(1 - b1 ** self.t), float(1 - b2 ** self.t),
                  stream_handle())
        # the padding channels of the first layer stay zero: their inputs are zero, so are their gradients
        self._packed = False


def load_default():
    """The committed trained network, or None when the weight file is absent."""
    if not os.path.exists(WEIGHTS):
        return None
    return UNet().load(WEIGHTS)


# ---------------------------------------------------------------------------
# Online training on (b, p) pairs from the step loop
# ---------------------------------------------------------------------------
def train(res=(128, 64, 64), pairs=400, steps_per_sim=16, lr=1e-3, seed=0, out=WEIGHTS, log=print, budget_s=None,
          loss="energy", init=None):
    """Run the EDGE step loop on seeded ring flows at `res`; every projection's
    (b, converged p) is one training pair, used for one Adam step (mixed solver,
    ||r|| <= 1e-6 ||b||).  A new flow (seed) every `steps_per_sim` steps.
    loss: "energy" (the error's energy norm, see UNet.backward) or "mse";
    init: weights to start from (a path), else a fresh network."""
    _here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if _here not in sys.path:
        sys.path.insert(0, _here)
    from run_edge import Simulation
    net = UNet(seed=seed)
    if init:
        net.load(init)
    n = tuple(res)
    UNet.check_dims(n)
    b = torch.empty(n, dtype=torch.float32, device="cuda")
    t0 = time.time()
    done, sim_seed, losses = 0, seed, []
    while done < pairs and (budget_s is None or time.time() - t0 < budget_s):
        sim = Simulation("cfg2", solver_mode="mixed", stop="relative", rel_tol=1e-6, res=n, num_samples=500,
                         seed=sim_seed)
        sim.initialize()
        sim_seed += 1
        solver = sim.solver

        def projection(u_x, u_y, u_z, max_iters=400, verbose=False, x0=None, _s=solver):
            _lib.call("edge_mg_rhs", _s._h, as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), _ptr(b), stream_handle())
            rv = _s.__class__.Poisson(_s, u_x, u_y, u_z, max_iters=max_iters, verbose=verbose, x0=x0)
            pairs_seen.append(True)
            net.forward(b)
            val = net.backward(_s.p.data.reshape(-1), loss=loss,
                               apply_A=lambda x, out: _s.get_Ap(x, out))
            net.adam_step(lr=lr * (0.5 ** (len(losses) // max(1, pairs // 4))))
            losses.append(val)
            return rv

        pairs_seen = []
        solver.Poisson = projection
        for _ in range(steps_per_sim):
            sim.substep()
            if len(losses) >= pairs:
                break
        done = len(losses)
        log(f"[train] sim seed {sim_seed - 1}: pairs {done}, loss (last 8 mean) {np.mean(losses[-8:]):.4g}, "
            f"{time.time() - t0:.0f}s")
    if out:
        net.save(out)
    return net, losses


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the projection's learned warm start")
    ap.add_argument("--res", type=int, nargs=3, default=[128, 64, 64])
    ap.add_argument("--pairs", type=int, default=400)
    ap.add_argument("--steps-per-sim", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--budget-s", type=float, default=None)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--loss", default="energy", choices=["energy", "mse"])
    ap.add_argument("--init", default=None, help="weights to start from")
    a = ap.parse_args(argv)
    train(tuple(a.res), a.pairs, a.steps_per_sim, a.lr, out=a.out, budget_s=a.budget_s,
          log=lambda m: print(m, flush=True), loss=a.loss, init=a.init)


if __name__ == "__main__":
    main()
