"""Tensor-level wrappers of the point-evaluation and march entry points.

These take torch CUDA tensors in the shared SoA layout (see edge_hip.field)
and are what the parity tests and the benchmark drive; the drop-in modules in
../sim call the same C-ABI on Fields.
"""
import ctypes

import torch

from . import _lib
from .field import stream_handle, touch_ptr


def _p(t):
    if t is None:
        return None
    assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), "expected contiguous float32 CUDA tensor"
    return t.data_ptr()


def _cells(u):
    return u[0].shape[0] - 1, u[0].shape[1], u[0].shape[2]


def eval_bspline(f, BL, pts, dx, grad=False):
    """interp_2 / interp_grad_2 of one field at points (n, 3)."""
    n = pts.shape[0]
    val = torch.empty(n, device=pts.device)
    g = torch.empty((n, 3), device=pts.device) if grad else None
    bl = (ctypes.c_float * 3)(*BL)
    _lib.call("edge_eval_points", 0, _p(f), *f.shape, ctypes.cast(bl, ctypes.c_void_p), None, None, None, None,
              4, 4, 4, _p(pts), n, float(dx), _p(val), _p(g), stream_handle())
    return (val, g) if grad else val


def eval_umac(u, pts, dx, grad=False):
    nx, ny, nz = _cells(u)
    n = pts.shape[0]
    val = torch.empty((n, 3), device=pts.device)
    J = torch.empty((n, 3, 3), device=pts.device) if grad else None
    _lib.call("edge_eval_points", 1, None, 0, 0, 0, None, _p(u[0]), _p(u[1]), _p(u[2]), None, nx, ny, nz, _p(pts), n,
              float(dx), _p(val), _p(J), stream_handle())
    return (val, J) if grad else val


def pack_table(kind, psi, T, dT=None, d2T=None, dx=1.0):
    """Hermite table of a map given as tensors: psi (3,...), T (3,3,...), dT (3,3,...), d2T (3,...).

    kind: "garm" | "full" | "fd"."""
    nx, ny, nz = psi.shape[1:]
    N = nx * ny * nz
    tab = torch.empty(3 * N * 8, device=psi.device)
    s = stream_handle()
    if kind == "fd":
        dT = torch.empty((3, 3, nx, ny, nz), device=psi.device)
        d2T = torch.empty((3, nx, ny, nz), device=psi.device)
        _lib.call("edge_calc_fd3", 1, _p(T[0]), _p(T[1]), _p(T[2]), _p(dT[0]), _p(dT[1]), _p(dT[2]), _p(d2T[0]),
                  _p(d2T[1]), _p(d2T[2]), nx, ny, nz, float(dx), s)
        kind = "full"
    k = 0 if kind == "garm" else 1
    d2 = [_p(d2T[c]) for c in range(3)] if k == 1 else [None] * 3
    _lib.call("edge_pack_hermite_table", k, _p(psi), _p(T[0]), _p(T[1]), _p(T[2]), _p(dT[0]), _p(dT[1]), _p(dT[2]),
              *d2, _p(tab), nx, ny, nz, float(dx), s)
    return tab


def eval_map(tab, dims, pts, dx, grad=False):
    nx, ny, nz = dims
    n = pts.shape[0]
    val = torch.empty((n, 3), device=pts.device)
    G = torch.empty((n, 3, 3), device=pts.device) if grad else None
    _lib.call("edge_eval_points", 2, None, 0, 0, 0, None, None, None, None, _p(tab), nx, ny, nz, _p(pts), n,
              float(dx), _p(val), _p(G), stream_handle())
    return (val, G) if grad else val


def advance_points(mode, u, pts, T, dx, dt):
    """pts (3, n...) and T (3, 3, n...) marched in place (mode 0: psi only)."""
    nx, ny, nz = _cells(u)
    n = pts[0].numel()
    Tp = [_p(T[c]) for c in range(3)] if mode == 1 else [None] * 3
    _lib.call("edge_advance_points", mode, _p(u[0]), _p(u[1]), _p(u[2]), _p(pts), *Tp, n, nx, ny, nz, float(dx),
              float(dt), stream_handle())


def compose_points(tab, dims, pts, T, dx):
    nx, ny, nz = dims
    n = pts[0].numel()
    Tp = [_p(T[c]) for c in range(3)] if T is not None else [None] * 3
    _lib.call("edge_compose_points", _p(tab), _p(pts), *Tp, n, nx, ny, nz, float(dx), stream_handle())


def march_edge(u, tab_old, X, psi, T, dT, dx, dt, eps):
    """Fused EDGE march: writes the new map into psi (3,...), T (3,3,...), dT (3,3,...)."""
    nx, ny, nz = psi.shape[1:]
    _lib.call("edge_march_edge", _p(u[0]), _p(u[1]), _p(u[2]), _p(tab_old), _p(X), _p(psi), _p(T[0]), _p(T[1]),
              _p(T[2]), _p(dT[0]), _p(dT[1]), _p(dT[2]), nx, ny, nz, float(dx), float(dt), float(eps),
              stream_handle())
    for t in (psi, T, dT):
        touch_ptr(t.data_ptr())


def march_edge_table(u, tab_old, X, psi, T, dT, dx, dt, eps):
    """march_edge that also returns the new map's garm table (the march writes
    its low planes, edge_pack_table_hi the high ones)."""
    nx, ny, nz = psi.shape[1:]
    tab_new = torch.empty_like(tab_old)
    s = stream_handle()
    _lib.call("edge_march_edge_table", _p(u[0]), _p(u[1]), _p(u[2]), _p(tab_old), _p(X), _p(psi), _p(T[0]),
              _p(T[1]), _p(T[2]), _p(dT[0]), _p(dT[1]), _p(dT[2]), _p(tab_new), nx, ny, nz, float(dx), float(dt),
              float(eps), s)
    _lib.call("edge_pack_table_hi", _p(dT[0]), _p(dT[1]), _p(dT[2]), _p(tab_new), nx, ny, nz, float(dx), s)
    for t in (psi, T, dT):
        touch_ptr(t.data_ptr())
    return tab_new


def advance_faces(u, phi, F, dx, dt):
    nx, ny, nz = _cells(u)
    _lib.call("edge_advance_map_faces", _p(u[0]), _p(u[1]), _p(u[2]), *[_p(t) for t in phi], *[_p(t) for t in F],
              nx, ny, nz, float(dx), float(dt), stream_handle())


def map_u(u0, F, phi, dx):
    nx, ny, nz = _cells(u0)
    out = [torch.empty_like(c) for c in u0]
    _lib.call("edge_map_u", *[_p(c) for c in u0], *[_p(c) for c in out], *[_p(c) for c in F], *[_p(c) for c in phi],
              nx, ny, nz, float(dx), stream_handle())
    return out


def map_u_psi(u0, tab, dx):
    nx, ny, nz = _cells(u0)
    out = [torch.empty_like(c) for c in u0]
    _lib.call("edge_map_u_psi", *[_p(c) for c in u0], *[_p(c) for c in out], _p(tab), nx, ny, nz, float(dx),
              stream_handle())
    return out


def map_smoke(mode, s0, s1, phi, dx):
    """s0, s1: (nx, ny, nz, nc) channels per cell (nc 4), or (nx, ny, nz)."""
    nx, ny, nz = s0.shape[:3]
    nc = s0.shape[3] if s0.dim() == 4 else 1
    s1 = s1.clone() if s1 is not None else torch.empty_like(s0)
    _lib.call("edge_map_smoke", mode, _p(s0), _p(s1), nc, *[_p(c) for c in phi], nx, ny, nz, float(dx),
              stream_handle())
    return s1


def map_smoke_psi(mode, s0, s1, tab, dx):
    """s0, s1: (nx, ny, nz, nc) channels per cell (nc 4), or (nx, ny, nz)."""
    nx, ny, nz = s0.shape[:3]
    nc = s0.shape[3] if s0.dim() == 4 else 1
    s1 = s1.clone() if s1 is not None else torch.empty_like(s0)
    _lib.call("edge_map_smoke_psi", mode, _p(s0), _p(s1), nc, _p(tab), nx, ny, nz, float(dx), stream_handle())
    return s1


def clamp_u(u, u1):
    nx, ny, nz = _cells(u)
    out = [torch.empty_like(c) for c in u]
    _lib.call("edge_clamp_u", *[_p(c) for c in u], *[_p(c) for c in u1], *[_p(c) for c in out], nx, ny, nz,
              stream_handle())
    return out


def advect_u_sl(u0, dx, dt):
    nx, ny, nz = _cells(u0)
    out = [torch.empty_like(c) for c in u0]
    _lib.call("edge_advect_u_sl", *[_p(c) for c in u0], *[_p(c) for c in out], nx, ny, nz, float(dx), float(dt),
              stream_handle())
    return out


def max_speed(u):
    nx, ny, nz = _cells(u)
    buf = torch.zeros(1, device=u[0].device)
    _lib.call("edge_max_speed", *[_p(c) for c in u], _p(buf), nx, ny, nz, stream_handle())
    return float(buf.item())


def split_psi_hermite(tab, dims, dx, device):
    nx, ny, nz = dims
    shapes = [(3, nx + 1, ny, nz), (3, nx, ny + 1, nz), (3, nx, ny, nz + 1)]
    phi = [torch.empty(s, device=device) for s in shapes]
    F = [torch.empty(s, device=device) for s in shapes]
    _lib.call("edge_split_psi_hermite", _p(tab), *[_p(c) for c in phi], *[_p(c) for c in F], nx, ny, nz, float(dx),
              stream_handle())
    return phi, F


def calc_fd3(T, dx):
    nx, ny, nz = T.shape[2:]
    dT = torch.empty_like(T)
    d2T = torch.empty((3, nx, ny, nz), device=T.device)
    _lib.call("edge_calc_fd3", 1, _p(T[0]), _p(T[1]), _p(T[2]), _p(dT[0]), _p(dT[1]), _p(dT[2]), _p(d2T[0]),
              _p(d2T[1]), _p(d2T[2]), nx, ny, nz, float(dx), stream_handle())
    return dT, d2T


def vortex_ring(w, rad, delta, c, ex, ey, X, vf, smoke, color, num_samples):
    """vf: (3, nx, ny, nz) planes; smoke: (nx, ny, nz, 4) channels per cell."""
    nx, ny, nz = X.shape[1:]
    arr = lambda v: (ctypes.c_float * 3)(*[float(x) for x in v])  # noqa: E731
    a_c, a_ex, a_ey, a_col = arr(c), arr(ex), arr(ey), arr(color)
    _lib.call("edge_vortex_ring", float(w), float(rad), float(delta), ctypes.cast(a_c, ctypes.c_void_p),
              ctypes.cast(a_ex, ctypes.c_void_p), ctypes.cast(a_ey, ctypes.c_void_p), _p(X), _p(vf), _p(smoke),
              ctypes.cast(a_col, ctypes.c_void_p), int(num_samples), nx, ny, nz, stream_handle())
