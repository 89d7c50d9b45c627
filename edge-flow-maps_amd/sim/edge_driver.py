"""The driver-local kernels of the reference's run_edge.py / run_ed8.py / run_ge.py.

In the reference these @ti.kernel functions read module globals (res_*, dx,
X_x..., psi_garm...); here the same computations take those fields as
arguments.  Names match the reference; argument lists append what the
reference took from globals.
"""
import torch

from edge_hip import _lib
from edge_hip.field import as_ptr, stream_handle, touch, version_of
from edge_hip import timing
from edge_hip.tables import hermite_table, table_version


def _cells_of_mac(u_x):
    return u_x.shape[0] - 1, u_x.shape[1], u_x.shape[2]


_speed_buf = {}
# (kind, device) -> (key, event, pinned host value): a maximum launched ahead
# by prefetch_max_speed, valid while the fields keep the storage and version
# they had then (every kernel wrapper touches what it writes)
_prefetched = {}
_pinned = {}


def _key(fields):
    return tuple((as_ptr(f), version_of(as_ptr(f))) for f in fields)


def _take(kind, dev, key):
    pre = _prefetched.pop((kind, dev), None)
    if pre is None or pre[0] != key:
        return None
    pre[1].synchronize()
    return float(pre[2][0])


def _launch_max_speed(u_x, u_y, u_z, buf):
    nx, ny, nz = _cells_of_mac(u_x)
    _lib.call("edge_max_speed", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), buf.data_ptr(), nx, ny, nz, stream_handle())


def _dev_buf(table, dev):
    buf = table.get(dev)
    if buf is None:
        buf = table[dev] = torch.zeros(1, dtype=torch.float32, device=dev)
    return buf


def calc_max_speed(u_x, u_y, u_z):
    """run_edge.py:295-303; returns max(1e-3, max cell-centred speed) (host sync, as max_speed[None]).
    When prefetch_max_speed ran on the same, unmodified fields, its value (the same
    kernel's) is returned once its copy has landed, without another launch."""
    dev = u_x.data.device
    got = _take("u", dev, _key((u_x, u_y, u_z)))
    if got is not None:
        return got
    buf = _dev_buf(_speed_buf, dev)
    _launch_max_speed(u_x, u_y, u_z, buf)
    return float(buf.item())


def _prefetch(kind, dev, key, buf):
    host = _pinned.get((kind, dev))
    if host is None:
        host = _pinned[(kind, dev)] = torch.zeros(1, dtype=torch.float32, pin_memory=True)
    host.copy_(buf, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    _prefetched[(kind, dev)] = (key, ev, host)


def prefetch_max_speed(u_x, u_y, u_z):
    """Launch calc_max_speed's kernel now and copy its result to pinned host memory
    without waiting: the step loop calls this at the end of a step, so the next
    step's calc_max_speed finds the value ready instead of launching and waiting
    at its start (the GPU would otherwise idle through that round trip)."""
    dev = u_x.data.device
    buf = _dev_buf(_speed_buf, dev)
    _launch_max_speed(u_x, u_y, u_z, buf)
    _prefetch("u", dev, _key((u_x, u_y, u_z)), buf)


_pspeed_buf = {}


def calc_max_particle_speed(laden_particles_vel):
    """run_edge.py:258-262; returns max(1e-3, max |v|) over every particle slot (host sync;
    a prefetched value as in calc_max_speed)."""
    dev = laden_particles_vel.data.device
    got = _take("p", dev, _key((laden_particles_vel,)))
    if got is not None:
        return got
    buf = _dev_buf(_pspeed_buf, dev)
    _lib.call("edge_lp_max_speed", as_ptr(laden_particles_vel), laden_particles_vel.shape[0], buf.data_ptr(),
              stream_handle())
    return float(buf.item())


def prefetch_max_particle_speed(laden_particles_vel):
    """calc_max_particle_speed launched ahead (see prefetch_max_speed)."""
    dev = laden_particles_vel.data.device
    buf = _dev_buf(_pspeed_buf, dev)
    _lib.call("edge_lp_max_speed", as_ptr(laden_particles_vel), laden_particles_vel.shape[0], buf.data_ptr(),
              stream_handle())
    _prefetch("p", dev, _key((laden_particles_vel,)), buf)


def reset_to_identity(psi_x, psi_y, psi_z, T_x, T_y, T_z, dx):
    """run_edge.py:306-320: face maps phi_a = X_a, F_a = e_a."""
    nx, ny, nz = _cells_of_mac(psi_x)
    _lib.call("edge_reset_faces", as_ptr(psi_x), as_ptr(psi_y), as_ptr(psi_z), as_ptr(T_x), as_ptr(T_y), as_ptr(T_z),
              nx, ny, nz, float(dx), stream_handle())
    touch(psi_x, psi_y, psi_z, T_x, T_y, T_z)


def march_phi_grid(phi_x, phi_y, phi_z, F_x, F_y, F_z, u_x, u_y, u_z, dx, dt):
    """run_edge.py:330-357: advance_map on the three face grids (+dt), one launch."""
    nx, ny, nz = _cells_of_mac(u_x)
    with timing.span("k_advance_faces"):
        _lib.call("edge_advance_map_faces", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), as_ptr(phi_x), as_ptr(phi_y),
                  as_ptr(phi_z), as_ptr(F_x), as_ptr(F_y), as_ptr(F_z), nx, ny, nz, float(dx), float(dt),
                  stream_handle())
    touch(phi_x, phi_y, phi_z, F_x, F_y, F_z)


def backtrack_faces(psi_x, psi_y, psi_z, T_x, T_y, T_z, velocities, dts, dx):
    """run_efm.py:170-185 backtrack_psi_grid in one pass: reset_to_identity, then
    advance_map through velocities[s] = (u_x, u_y, u_z) with dt dts[s], s = 0, 1, ..
    (the reference passes the current velocity first, then its cache newest first)."""
    import ctypes
    n = len(velocities)
    assert len(dts) == n
    nx, ny, nz = _cells_of_mac(psi_x)
    arr = [(ctypes.c_void_p * max(n, 1))(*[as_ptr(v[c]) for v in velocities]) for c in range(3)]
    dta = (ctypes.c_float * max(n, 1))(*[float(d) for d in dts])
    with timing.span("k_backtrack_faces"):
        _lib.call("edge_backtrack_faces", *[ctypes.cast(a, ctypes.c_void_p) for a in arr],
                  ctypes.cast(dta, ctypes.c_void_p), n, as_ptr(psi_x), as_ptr(psi_y), as_ptr(psi_z), as_ptr(T_x),
                  as_ptr(T_y), as_ptr(T_z), nx, ny, nz, float(dx), stream_handle())
    touch(psi_x, psi_y, psi_z, T_x, T_y, T_z)


def map_u(u_x0, u_y0, u_z0, u_x1, u_y1, u_z1, T_x, T_y, T_z, psi_x, psi_y, psi_z, dx, add_scale=None,
          sub_clamp=None):
    """run_edge.py:360-373.

    add_scale = (f_x, f_y, f_z, multiplier, scale) folds the BFECC arithmetic
    into the same pass: u1 = scale * (map_u(u0) + multiplier * f), the values of
    map_u, add_fields(u1, f, u1, multiplier), scale_field(u1, scale, u1)
    (run_edge.py:671-681).
    sub_clamp = (c_x, c_y, c_z) stores clamp(c - map_u(u0)) to the min and max
    of c's six neighbours: map_u into tmp_u, add_fields(c, tmp_u, err, -1) and
    clamp_u(err, c, tmp_u) (run_edge.py:683-693) in one pass, the same values."""
    nx, ny, nz = _cells_of_mac(u_x0)
    _map_u_ptrs([as_ptr(u_x0), as_ptr(u_y0), as_ptr(u_z0), as_ptr(u_x1), as_ptr(u_y1), as_ptr(u_z1), as_ptr(T_x),
                 as_ptr(T_y), as_ptr(T_z), as_ptr(psi_x), as_ptr(psi_y), as_ptr(psi_z)], nx, ny, nz, dx,
                add_scale, sub_clamp)
    touch(u_x1, u_y1, u_z1)


def _map_u_ptrs(args, nx, ny, nz, dx, add_scale=None, sub_clamp=None):
    if add_scale is not None and sub_clamp is not None:
        raise ValueError("give add_scale or sub_clamp, not both")
    with timing.span("k_map_u"):
        if add_scale is not None:
            f_x, f_y, f_z, m, sc = add_scale
            _lib.call("edge_map_u_add_scale", *args, as_ptr(f_x), as_ptr(f_y), as_ptr(f_z), float(m), float(sc), nx,
                      ny, nz, float(dx), stream_handle())
        elif sub_clamp is not None:
            _lib.call("edge_map_u_sub_clamp", *args, *[as_ptr(c) for c in sub_clamp], nx, ny, nz, float(dx),
                      stream_handle())
        else:
            _lib.call("edge_map_u", *args, nx, ny, nz, float(dx), stream_handle())


def _garm_table(m, dx):
    psi, T_x, T_y, T_z, dT_x, dT_y, dT_z = m
    return hermite_table("garm", psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), None, dx)


# Data derived from one map state and shared by the step's pull-backs: the
# face map (psi and the column d psi / d x_a at every face, i.e.
# split_psi_hermite of the garm map), read by map_u_psi, map_smoke_psi and
# clamp_smoke_psi.  Keyed by the map's storage and version.
_derived = {"key": None, "face": None, "bufs": {}}


def _derived_for(m, dx):
    psi, T_x, T_y, T_z, dT_x, dT_y, dT_z = m
    key = table_version("garm", psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), None, dx)
    if _derived["key"] != key:
        _derived["key"] = key
        _derived["face"] = None
    return _derived


def _bufs(shape, device):
    b = _derived["bufs"].get((shape, device))
    if b is None:
        nx, ny, nz = shape
        fs = [(3, nx + 1, ny, nz), (3, nx, ny + 1, nz), (3, nx, ny, nz + 1)]
        b = dict(phi=[torch.empty(f, device=device) for f in fs], F=[torch.empty(f, device=device) for f in fs])
        _derived["bufs"][(shape, device)] = b
    return b


def face_map_of(m, dx):
    """(phi[3], F[3]) tensors: the map and its d/dx_a column at the faces (cached)."""
    d = _derived_for(m, dx)
    if d["face"] is None:
        psi = m[0]
        nx, ny, nz = psi.shape
        b = _bufs(psi.shape, psi.data.device)
        tab = _garm_table(m, dx)
        _lib.call("edge_split_psi_hermite", tab.data_ptr(), *[t.data_ptr() for t in b["phi"]],
                  *[t.data_ptr() for t in b["F"]], nx, ny, nz, float(dx), stream_handle())
        d["face"] = (b["phi"], b["F"])
    return d["face"]


def map_u_psi(u_x0, u_y0, u_z0, u_x1, u_y1, u_z1, dx, garm_map, sub_clamp=None):
    """run_edge.py:376-417; garm_map = (psi_garm, T_x/y/z_garm, dT_x/y/z_garm).

    u1_a(X_a) = sum_c d psi_c / d x_a (X_a) u0_c(psi(X_a)): the map's value and
    column at the faces are evaluated once per map state and shared by the
    step's two calls; the pull-back itself is map_u's kernel.  sub_clamp as in
    map_u (the BFECC correction and clamp_u of run_edge.py:683-693 in the same
    pass)."""
    nx, ny, nz = _cells_of_mac(u_x0)
    phi, F = face_map_of(garm_map, dx)
    _map_u_ptrs([as_ptr(u_x0), as_ptr(u_y0), as_ptr(u_z0), as_ptr(u_x1), as_ptr(u_y1), as_ptr(u_z1),
                 *[t.data_ptr() for t in F], *[t.data_ptr() for t in phi]], nx, ny, nz, dx, None, sub_clamp)
    touch(u_x1, u_y1, u_z1)


def _smoke_ptrs(mode, smoke0, smoke1, phi_ptrs, dx, add):
    nx, ny, nz = smoke0.shape
    nc = smoke0.n or 1
    if add is None:
        _lib.call("edge_map_smoke", mode, as_ptr(smoke0), as_ptr(smoke1), nc, *phi_ptrs, nx, ny, nz, float(dx),
                  stream_handle())
    else:
        s2, m, sc = add
        _lib.call("edge_map_smoke_add_scale", mode, as_ptr(smoke0), as_ptr(smoke1), as_ptr(s2), float(m), float(sc),
                  nc, *phi_ptrs, nx, ny, nz, float(dx), stream_handle())
    touch(smoke1)


def _smoke(mode, smoke0, smoke1, psi_x, psi_y, psi_z, dx, add=None):
    _smoke_ptrs(mode, smoke0, smoke1, (as_ptr(psi_x), as_ptr(psi_y), as_ptr(psi_z)), dx, add)


def map_smoke(smoke0, smoke1, psi_x, psi_y, psi_z, dx, add_scale=None):
    """run_edge.py:420-427.  add_scale = (f, multiplier, scale) stores
    scale * (map + multiplier * f), the values of map_smoke, add_fields and
    scale_field (run_edge.py:672-682) in one pass."""
    _smoke(0, smoke0, smoke1, psi_x, psi_y, psi_z, dx, add_scale)


def _sub_args(sub, sub_map):
    """(mode, add) of a clamp: smoke1 - sub (mode 1) or smoke1 - map(sub_map) (mode 2)."""
    if sub is not None and sub_map is not None:
        raise ValueError("give sub or sub_map, not both")
    if sub_map is not None:
        return 2, (sub_map, -1., 1.)
    return 1, None if sub is None else (sub, -1., 1.)


def clamp_smoke(smoke0, smoke1, psi_x, psi_y, psi_z, dx, sub=None, sub_map=None):
    """run_edge.py:469-477 (smoke1 clamped in place).  With sub, smoke1 - sub is
    clamped: add_fields(smoke1, sub, smoke1, -1) (run_edge.py:688) in the same pass.
    With sub_map, smoke1 - map_smoke(sub_map) is clamped: run_edge.py:684-689
    (the BFECC error pulled back, subtracted and clamped) in one pass, the
    pulled-back error kept in registers (tmp_smoke is not written)."""
    mode, add = _sub_args(sub, sub_map)
    _smoke(mode, smoke0, smoke1, psi_x, psi_y, psi_z, dx, add)


def _smoke_psi(mode, smoke0, smoke1, dx, garm_map, add=None):
    # psi_c = 1/6 sum of psi at the cell's six faces (run_edge.py:448-465) is
    # map_smoke's average of a face map, here the cached psi at the faces
    phi, _ = face_map_of(garm_map, dx)
    _smoke_ptrs(mode, smoke0, smoke1, [t.data_ptr() for t in phi], dx, add)


def map_smoke_psi(smoke0, smoke1, dx, garm_map):
    """run_edge.py:443-466 (face positions shared with map_u_psi)."""
    _smoke_psi(0, smoke0, smoke1, dx, garm_map)


def clamp_smoke_psi(smoke0, smoke1, dx, garm_map, sub=None, sub_map=None):
    """run_edge.py:480-505 (smoke1 clamped in place).  With sub, smoke1 - sub is
    clamped: add_fields(smoke1, sub, smoke1, -1) (run_edge.py:688) in the same pass.
    With sub_map, smoke1 - map_smoke_psi(sub_map) is clamped (run_edge.py:684-689
    in one pass; tmp_smoke is not written)."""
    mode, add = _sub_args(sub, sub_map)
    _smoke_psi(mode, smoke0, smoke1, dx, garm_map, add)


def clamp_u3(u_x, u_y, u_z, u1_x, u1_y, u1_z, u2_x, u2_y, u2_z):
    """run_edge.py:508-519 clamp_u on the three face grids in one launch."""
    nx, ny, nz = _cells_of_mac(u_x)
    _lib.call("edge_clamp_u", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), as_ptr(u1_x), as_ptr(u1_y), as_ptr(u1_z),
              as_ptr(u2_x), as_ptr(u2_y), as_ptr(u2_z), nx, ny, nz, stream_handle())
    touch(u2_x, u2_y, u2_z)


def sub_clamp_u3(u_x, u_y, u_z, t_x, t_y, t_z):
    """run_edge.py:685-693 add_fields(u, t, err, -1) then clamp_u(err, u, t), one launch over the
    three face grids, in place on t; err is not written."""
    nx, ny, nz = _cells_of_mac(u_x)
    _lib.call("edge_sub_clamp_u", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), as_ptr(t_x), as_ptr(t_y), as_ptr(t_z),
              nx, ny, nz, stream_handle())
    touch(t_x, t_y, t_z)


def split_psi_hermite(psi_x, psi_y, psi_z, T_x, T_y, T_z, full_map, dx):
    """run_ed8.py:190-223 / run_ge.py:184-217; full_map = (psi, T_x/y/z, dT_x/y/z, d2T_x/y/z).

    full_map = (psi, T_x/y/z) alone is run_ed4.py's map (FD-unrolled Hermite):
    the map's value and column at every face, which run_ed4.py's map_u_psi
    (:219-262, interp_grad_hermite_fd_unroll) and map_smoke_psi / clamp_smoke_psi
    (:285-337, interp_hermite_fd_unroll at the six faces) evaluate in place."""
    if len(full_map) == 4:
        psi, Tx, Ty, Tz = full_map
        tab = hermite_table("fd", psi, (Tx, Ty, Tz), None, None, dx)
    else:
        psi, Tx, Ty, Tz, dTx, dTy, dTz, d2x, d2y, d2z = full_map
        tab = hermite_table("full", psi, (Tx, Ty, Tz), (dTx, dTy, dTz), (d2x, d2y, d2z), dx)
    nx, ny, nz = psi.shape
    _lib.call("edge_split_psi_hermite", tab.data_ptr(), as_ptr(psi_x), as_ptr(psi_y), as_ptr(psi_z), as_ptr(T_x),
              as_ptr(T_y), as_ptr(T_z), nx, ny, nz, float(dx), stream_handle())
    touch(psi_x, psi_y, psi_z, T_x, T_y, T_z)
