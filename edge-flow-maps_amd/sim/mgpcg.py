"""Drop-in for the reference's sim/mgpcg.py: MGPCG pressure projection.

MGPCG_3 keeps the reference's constructor and Poisson() interface
(sim/mgpcg.py:364-491); the solver runs in libedge_hip.so.  Extensions, all
keyword-only with defaults that reproduce the reference:
  mode        "plain" (FP32, the reference V-cycle), "mixed" (BF16 multigrid
              storage, FP64 dense coarsest solve) or "dense" (FP32 multigrid
              storage, FP64 dense coarsest solve)
  stop        "reference" (rTr < max(abs_tol, rel_tol * z0'r0), sim/mgpcg.py:228)
              or "relative" (||r|| <= rel_tol ||b||, FP64 norms)
  check_every CG iterations between host convergence checks (the reference
              reads a scalar every iteration: 1; at most 8), or 0 for batches
              sized from the last two solves (default for "mixed", "dense")
  graph       replay each batch of CG iterations as captured HIP graphs
              (default: on for "mixed" and "dense", off for "plain")
  x0          warm start for Poisson(), used only if ||b - A x0|| <= ||b||
  warm_start  a learned initial guess (sim.learned.UNet, or True for the
              committed network): Poisson() computes b, runs the network
              and passes its p0 as x0 (the same acceptance rule applies)
After a solve, `self.p` holds the pressure (the reference's get_result(self.p),
sim/mgpcg.py:438) and `self.stats` the solve statistics.
"""
import ctypes
import math
import time

import numpy as np
import torch

from edge_hip import _lib
from edge_hip.field import Field, as_ptr, stream_handle, touch

STAT_NAMES = ("iters", "rTr", "b2", "div_sum", "corrected", "converged", "zTr0", "r02", "x0_used")


def _bt_array(boundary_types):
    if boundary_types is None:
        return None
    bt = np.asarray(boundary_types.to_numpy() if hasattr(boundary_types, "to_numpy") else boundary_types,
                    dtype=np.int32).reshape(6)
    return (ctypes.c_int * 6)(*bt.tolist())


_MODES = {"plain": 0, "mixed": 1, "dense": 2}


class _MGBase:
    def __init__(self, boundary_types, N, base_level=3, levels=None, mode="plain", stop="reference",
                 check_every=None, abs_tol=1e-12, rel_tol=1e-12, graph=None, warm_start=None):
        self.N = [int(n) for n in N]
        self.dim = 3
        self.n_mg_levels = int(levels) if levels is not None else int(math.log2(min(self.N))) - base_level + 1
        self.mode = mode
        self.stop = stop
        self.abs_tol = abs_tol
        self.rel_tol = rel_tol
        self.check_every = check_every if check_every is not None else (1 if mode == "plain" else 0)
        self.boundary_types = boundary_types
        self._bt = _bt_array(boundary_types)
        h = ctypes.c_void_p()
        if mode not in _MODES:
            raise ValueError(f"mode must be one of {sorted(_MODES)}, got {mode!r}")
        _lib.call("edge_mg_create", *self.N, self.n_mg_levels, _MODES[mode], self._bt, ctypes.byref(h))
        self._h = h
        self.graph = (mode != "plain") if graph is None else bool(graph)
        _lib.call("edge_mg_set_graph", self._h, int(self.graph))
        self.p = Field(None, float, self.N)
        self.stats = {}
        self.last_time = 0.0
        if warm_start is True:
            from sim.learned import load_default
            warm_start = load_default()
            if warm_start is None:
                raise RuntimeError("warm_start=True but the trained network (sim/unet_warmstart.npz) is missing")
        self.warm_start = warm_start
        self._b = None
        self._bm = None   # int32 solid mask for the network input (MGPCG_Solid_3 sets it)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().edge_mg_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def _run_poisson(self, u_x, u_y, u_z, f=(None, None, None), max_iters=400, stop_mode=None, abs_tol=None,
                     rel_tol=None, x0=None, verbose=False):
        stats = (ctypes.c_double * 9)()
        sm = {"reference": 0, "relative": 1}[self.stop] if stop_mode is None else stop_mode
        t0 = time.perf_counter()
        if self.warm_start is not None and x0 is None:
            if self._b is None:
                self._b = torch.empty(self.N, dtype=torch.float32, device=self.p.data.device)
            _lib.call("edge_mg_rhs", self._h, as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), self._b.data_ptr(),
                      stream_handle())
            x0 = self.warm_start.forward(self._b, self._bm, n=self.N)
        st = _lib.call("edge_poisson", self._h, as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), as_ptr(self.p),
                       as_ptr(f[0]), as_ptr(f[1]), as_ptr(f[2]), int(max_iters), sm,
                       float(self.abs_tol if abs_tol is None else abs_tol),
                       float(self.rel_tol if rel_tol is None else rel_tol), as_ptr(x0), int(self.check_every),
                       stats, stream_handle(), ok=(0, _lib.NOCONV))
        self.last_time = time.perf_counter() - t0
        touch(u_x, u_y, u_z, self.p, *[x for x in f if x is not None])
        self.stats = dict(zip(STAT_NAMES, list(stats)))
        self.stats["iters"] = int(self.stats["iters"])
        if verbose:
            word = "Converged" if st == 0 else "Return without converging"
            print(f"[MGPCG] {word} at iter: {self.stats['iters']} with final error: "
                  f"{math.sqrt(max(self.stats['rTr'], 0.0))} using time: {self.last_time}")
        return st

    def solve_b(self, b, x, max_iters=400, stop="relative", rel_tol=1e-6, abs_tol=1e-12, x0=None):
        """Solve A x = b (tensors or Fields of the grid shape); returns the stats dict."""
        stats = (ctypes.c_double * 9)()
        sm = {"reference": 0, "relative": 1}[stop]
        _lib.call("edge_mg_solve", self._h, as_ptr(b), as_ptr(x), int(max_iters), sm, float(abs_tol), float(rel_tol),
                  as_ptr(x0), int(self.check_every), stats, stream_handle(), ok=(0, _lib.NOCONV))
        touch(x)
        self.stats = dict(zip(STAT_NAMES, list(stats)))
        self.stats["iters"] = int(self.stats["iters"])
        return self.stats

    def apply_preconditioner_to(self, r, z):
        """One V-cycle z = M^-1 r (sim/mgpcg.py:177-195) on tensors/Fields."""
        _lib.call("edge_mg_vcycle", self._h, as_ptr(r), as_ptr(z), stream_handle())
        touch(z)

    def get_Ap(self, p, Ap):
        """sim/mgpcg.py:119-124."""
        _lib.call("edge_mg_apply_A", self._h, as_ptr(p), as_ptr(Ap), stream_handle())
        touch(Ap)

    @property
    def dense_bottom(self):
        return bool(_lib.lib().edge_mg_dense_bottom(self._h))

    def get_memory_usage(self):
        n = math.prod(self.N)
        elem = 2 if self.mode == "mixed" else 4
        levels = sum(n // 8 ** l for l in range(self.n_mg_levels))
        return n * 4 * 5 + levels * elem * 2


class MGPCG_3(_MGBase):
    """sim/mgpcg.py:364-491 (all-Neumann Poisson in grid units, 7-point stencil)."""

    def __init__(self, boundary_types, N, base_level=3, abs_tol=1e-12, rel_tol=1e-12, real=float, **kw):
        super().__init__(boundary_types, N, base_level=base_level, abs_tol=abs_tol, rel_tol=rel_tol, **kw)

    def Poisson(self, u_x, u_y, u_z, max_iters=400, verbose=False, x0=None):
        """apply_bc, divergence, solve (max 400 iterations, sim/mgpcg.py:436), subtract grad p, apply_bc."""
        return self._run_poisson(u_x, u_y, u_z, max_iters=max_iters, x0=x0, verbose=verbose)

    def Poisson_with_buffer(self, u_x, u_y, u_z, f_x, f_y, f_z, max_iters=400, verbose=False, x0=None):
        """sim/mgpcg.py:481-486: as Poisson, also f = -(grad p)."""
        return self._run_poisson(u_x, u_y, u_z, f=(f_x, f_y, f_z), max_iters=max_iters, x0=x0, verbose=verbose)
