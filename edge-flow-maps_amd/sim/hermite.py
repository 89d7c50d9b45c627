"""Drop-in for the reference's sim/hermite.py: flow-map marches and Hermite maps.

Same function names, argument order and meaning as sim/hermite.py; every
function runs hand-written gfx950 kernels of libedge_hip.so on the fields'
GPU storage (fields are edge_hip.Field).  Differences from the reference that
a caller can observe are listed per function:

* The fused marches (advect_psi_garm_ed4_fusion_unroll, advect_ps_hermite_ed8,
  advect_ps_garm, advect_ps_hermite_ed4_fd_unroll) read the old map through a
  packed Hermite table and write the new map in place; their scratch
  arguments (psi_tmp, T_*_tmp, dT_*_tmp, the eps-point fields) are accepted
  for signature compatibility and left untouched.  In the reference they end
  up holding a copy of the new map (sim/hermite.py:1515-1521); nothing reads
  them before they are overwritten again.
* The per-point @ti.func helpers (interp_*, hermite_*) are device code here;
  edge_hip.ops exposes batch point evaluation for tests.
"""
import ctypes

from edge_hip import _lib, timing
from edge_hip.field import Field, as_ptr, stream_handle, touch
from edge_hip.tables import hermite_table, install_table, spare_table

__all__ = [
    "reset_to_identity_1", "reset_to_identity_2", "reset_to_identity_3", "reset_to_identity_ed8",
    "reset_to_identity_ed4", "reset_to_identity_garm_ed4", "advance_map_0", "advance_map_1_transpose",
    "interp_psi_hermite", "interp_psi_hermite_fd_unroll", "interp_psi_hermite_garm_fd_unroll",
    "interp_psi_hermite_garm", "calc_ed4_garm", "calc_ed8_3", "calc_ed4_1", "calc_fd_3_from_grad",
    "calc_fd_3_from_grad_grad", "advect_ps_hermite_ed8", "advect_ps_hermite_ed4",
    "advect_ps_hermite_ed4_fd_unroll", "advect_ps_garm", "advect_psi_garm_ed4_fusion_unroll",
    "advect_psi_garm_ed4", "advect_u_semi_lagrangian", "transpose_matrix",
]


def _dims(f):
    nx, ny, nz = f.shape
    return nx, ny, nz


def _cells(f):
    nx, ny, nz = f.shape
    return nx * ny * nz


def _check_mac(u_x, u_y, u_z, nx, ny, nz):
    if u_x.shape != (nx + 1, ny, nz) or u_y.shape != (nx, ny + 1, nz) or u_z.shape != (nx, ny, nz + 1):
        raise ValueError(f"MAC velocity shapes {u_x.shape},{u_y.shape},{u_z.shape} do not match cells {(nx, ny, nz)}")


def transpose_matrix(T_x, T_y, T_z):
    """sim/hermite.py:623-630 on 3-vectors (host helper)."""
    return ((T_x[0], T_y[0], T_z[0]), (T_x[1], T_y[1], T_z[1]), (T_x[2], T_y[2], T_z[2]))


# ---------------------------------------------------------------------------
# resets (sim/hermite.py:642-783)
# ---------------------------------------------------------------------------
def _reset(X, psi, T=(None, None, None), dT=(None, None, None), d2T=(None, None, None)):
    nx, ny, nz = _dims(X)
    _lib.call("edge_reset_identity", as_ptr(X), as_ptr(psi), *[as_ptr(t) for t in T], *[as_ptr(t) for t in dT],
              *[as_ptr(t) for t in d2T], nx, ny, nz, stream_handle())
    touch(psi, *T, *dT, *d2T)


def reset_to_identity_1(psi, T_x, T_y, T_z, X):
    _reset(X, psi, (T_x, T_y, T_z))


def reset_to_identity_2(psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, X):
    _reset(X, psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z))


def reset_to_identity_3(psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, X):
    _reset(X, psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), (d2T_x, d2T_y, d2T_z))


def _offset(X, dst, o):
    _lib.call("edge_offset_points", as_ptr(X), as_ptr(dst), _cells(X), float(o[0]), float(o[1]), float(o[2]),
              stream_handle())
    touch(dst)


ED8_SIGNS = [(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, 1), (1, 1, 1)]
ED4_SIGNS = [(-1, -1, -1), (1, 1, -1), (-1, 1, 1), (1, -1, 1)]


def reset_to_identity_ed8(psi_000, psi_001, psi_010, psi_011, psi_100, psi_101, psi_110, psi_111, X, epsilon):
    """sim/hermite.py:704-725 (the last digit of the name is the x sign)."""
    for f, s in zip((psi_000, psi_001, psi_010, psi_011, psi_100, psi_101, psi_110, psi_111), ED8_SIGNS):
        _offset(X, f, [si * epsilon for si in s])


def reset_to_identity_ed4(psi_0, psi_1, psi_2, psi_3, X, epsilon):
    """sim/hermite.py:728-742."""
    for f, s in zip((psi_0, psi_1, psi_2, psi_3), ED4_SIGNS):
        _offset(X, f, [si * epsilon for si in s])


def reset_to_identity_garm_ed4(psi_0, T_x_0, T_y_0, T_z_0, psi_1, T_x_1, T_y_1, T_z_1, psi_2, T_x_2, T_y_2, T_z_2,
                               psi_3, T_x_3, T_y_3, T_z_3, X, epsilon):
    """sim/hermite.py:745-783."""
    reset_to_identity_ed4(psi_0, psi_1, psi_2, psi_3, X, epsilon)
    for T in ((T_x_0, T_y_0, T_z_0), (T_x_1, T_y_1, T_z_1), (T_x_2, T_y_2, T_z_2), (T_x_3, T_y_3, T_z_3)):
        _reset(X, None, T)


# ---------------------------------------------------------------------------
# point marches (sim/hermite.py:786-918)
# ---------------------------------------------------------------------------
def advance_map_0(psi, u_x0, u_y0, u_z0, dx, dt):
    """sim/hermite.py:786-803: RK4 of every point of psi through u."""
    nx, ny, nz = u_x0.shape[0] - 1, u_x0.shape[1], u_x0.shape[2]
    _check_mac(u_x0, u_y0, u_z0, nx, ny, nz)
    _lib.call("edge_advance_points", 0, as_ptr(u_x0), as_ptr(u_y0), as_ptr(u_z0), as_ptr(psi), None, None, None,
              _cells(psi), nx, ny, nz, float(dx), float(dt), stream_handle())
    touch(psi)


def advance_map_1_transpose(psi, T_x, T_y, T_z, u_x0, u_y0, u_z0, dx, dt):
    """sim/hermite.py:863-918: RK4 of psi with its Jacobian rows T."""
    nx, ny, nz = u_x0.shape[0] - 1, u_x0.shape[1], u_x0.shape[2]
    _check_mac(u_x0, u_y0, u_z0, nx, ny, nz)
    _lib.call("edge_advance_points", 1, as_ptr(u_x0), as_ptr(u_y0), as_ptr(u_z0), as_ptr(psi), as_ptr(T_x),
              as_ptr(T_y), as_ptr(T_z), _cells(psi), nx, ny, nz, float(dx), float(dt), stream_handle())
    touch(psi, T_x, T_y, T_z)


# ---------------------------------------------------------------------------
# compositions through an old map (sim/hermite.py:921-1046)
# ---------------------------------------------------------------------------
def _compose(tab, map_psi, pts, T=None, dx=None):
    nx, ny, nz = _dims(map_psi)
    Tp = [as_ptr(t) for t in T] if T is not None else [None, None, None]
    _lib.call("edge_compose_points", tab.data_ptr(), as_ptr(pts), *Tp, _cells(pts), nx, ny, nz, float(dx),
              stream_handle())
    touch(pts, *(T or ()))


def interp_psi_hermite(T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, psi, psi_tmp, dx):
    """sim/hermite.py:921-940: psi_tmp <- psi_old(psi_tmp) (full Hermite)."""
    tab = hermite_table("full", psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), (d2T_x, d2T_y, d2T_z), dx)
    _compose(tab, psi, psi_tmp, None, dx)


def interp_psi_hermite_fd_unroll(T_x, T_y, T_z, psi, psi_tmp, dx):
    """sim/hermite.py:943-956 (T only; cross terms by finite differences)."""
    tab = hermite_table("fd", psi, (T_x, T_y, T_z), None, None, dx)
    _compose(tab, psi, psi_tmp, None, dx)


def interp_psi_hermite_garm_fd_unroll(psi_tmp, T_x_tmp, T_y_tmp, T_z_tmp, psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, dx):
    """sim/hermite.py:983-995: compose (psi_tmp, T_tmp) with the garm map."""
    tab = hermite_table("garm", psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), None, dx)
    _compose(tab, psi, psi_tmp, (T_x_tmp, T_y_tmp, T_z_tmp), dx)


def interp_psi_hermite_garm(psi_tmp, T_x_tmp, T_y_tmp, T_z_tmp, psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y,
                            d2T_z, dx):
    """sim/hermite.py:1023-1046: compose (psi_tmp, T_tmp) with the full map."""
    tab = hermite_table("full", psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), (d2T_x, d2T_y, d2T_z), dx)
    _compose(tab, psi, psi_tmp, (T_x_tmp, T_y_tmp, T_z_tmp), dx)


# ---------------------------------------------------------------------------
# eps-difference assembly and finite differences (sim/hermite.py:1049-1287)
# ---------------------------------------------------------------------------
def calc_ed4_garm(T_x_0, T_y_0, T_z_0, T_x_1, T_y_1, T_z_1, T_x_2, T_y_2, T_z_2, T_x_3, T_y_3, T_z_3, dT_x, dT_y, dT_z,
                  epsilon):
    """sim/hermite.py:1078-1095."""
    ins = [T_x_0, T_y_0, T_z_0, T_x_1, T_y_1, T_z_1, T_x_2, T_y_2, T_z_2, T_x_3, T_y_3, T_z_3]
    _lib.call("edge_calc_ed4_garm", *[as_ptr(t) for t in ins], as_ptr(dT_x), as_ptr(dT_y), as_ptr(dT_z),
              _cells(T_x_0), float(epsilon), stream_handle())
    touch(dT_x, dT_y, dT_z)


def _calc_ed(pts, psi, T, dT, d2T, epsilon):
    arr = (ctypes.c_void_p * 8)(*([as_ptr(p) for p in pts] + [None] * (8 - len(pts))))
    _lib.call("edge_calc_ed", len(pts), ctypes.cast(arr, ctypes.c_void_p), as_ptr(psi), *[as_ptr(t) for t in T],
              *[as_ptr(t) for t in dT], *[as_ptr(t) for t in d2T], _cells(psi), float(epsilon), stream_handle())
    touch(psi, *T, *dT, *d2T)


def calc_ed8_3(psi_000, psi_001, psi_010, psi_011, psi_100, psi_101, psi_110, psi_111, psi, T_x, T_y, T_z, dT_x, dT_y,
               dT_z, d2T_x, d2T_y, d2T_z, epsilon):
    """sim/hermite.py:1098-1151."""
    _calc_ed((psi_000, psi_001, psi_010, psi_011, psi_100, psi_101, psi_110, psi_111), psi, (T_x, T_y, T_z),
             (dT_x, dT_y, dT_z), (d2T_x, d2T_y, d2T_z), epsilon)


def calc_ed4_1(psi_0, psi_1, psi_2, psi_3, psi, T_x, T_y, T_z, epsilon):
    """sim/hermite.py:1154-1177."""
    _calc_ed((psi_0, psi_1, psi_2, psi_3), psi, (T_x, T_y, T_z), (None, None, None), (None, None, None), epsilon)


def calc_fd_3_from_grad(T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, dx):
    """sim/hermite.py:1180-1255."""
    nx, ny, nz = _dims(T_x)
    _lib.call("edge_calc_fd3", 1, as_ptr(T_x), as_ptr(T_y), as_ptr(T_z), as_ptr(dT_x), as_ptr(dT_y), as_ptr(dT_z),
              as_ptr(d2T_x), as_ptr(d2T_y), as_ptr(d2T_z), nx, ny, nz, float(dx), stream_handle())
    touch(dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z)


def calc_fd_3_from_grad_grad(dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, dx):
    """sim/hermite.py:1258-1287."""
    nx, ny, nz = _dims(dT_x)
    _lib.call("edge_calc_fd3", 0, None, None, None, as_ptr(dT_x), as_ptr(dT_y), as_ptr(dT_z), as_ptr(d2T_x),
              as_ptr(d2T_y), as_ptr(d2T_z), nx, ny, nz, float(dx), stream_handle())
    touch(d2T_x, d2T_y, d2T_z)


# ---------------------------------------------------------------------------
# composite marches
# ---------------------------------------------------------------------------
def advect_ps_hermite_ed8(psi_000, psi_001, psi_010, psi_011, psi_100, psi_101, psi_110, psi_111, psi, T_x, T_y, T_z,
                          dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, u_x, u_y, u_z, X, dx, dt, epsilon):
    """sim/hermite.py:1290-1338, fused into one kernel (no eps-point fields written)."""
    nx, ny, nz = _dims(psi)
    _check_mac(u_x, u_y, u_z, nx, ny, nz)
    tab = hermite_table("full", psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), (d2T_x, d2T_y, d2T_z), dx)
    with timing.span("k_march_ed8"):
        _lib.call("edge_march_ed8", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), tab.data_ptr(), as_ptr(X), as_ptr(psi),
                  as_ptr(T_x), as_ptr(T_y), as_ptr(T_z), as_ptr(dT_x), as_ptr(dT_y), as_ptr(dT_z), as_ptr(d2T_x),
                  as_ptr(d2T_y), as_ptr(d2T_z), nx, ny, nz, float(dx), float(dt), float(epsilon), stream_handle())
    touch(psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z)


def advect_ps_hermite_ed4(psi_0, psi_1, psi_2, psi_3, psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, u_x,
                          u_y, u_z, X, dx, dt, epsilon):
    """sim/hermite.py:1341-1372 (unfused, as the reference composes it)."""
    reset_to_identity_ed4(psi_0, psi_1, psi_2, psi_3, X, epsilon)
    for p in (psi_0, psi_1, psi_2, psi_3):
        advance_map_0(p, u_x, u_y, u_z, dx, -dt)
    for p in (psi_0, psi_1, psi_2, psi_3):
        interp_psi_hermite(T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, psi, p, dx)
    calc_ed4_1(psi_0, psi_1, psi_2, psi_3, psi, T_x, T_y, T_z, epsilon)
    calc_fd_3_from_grad(T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, dx)


def advect_ps_hermite_ed4_fd_unroll(psi_0, psi_1, psi_2, psi_3, psi, T_x, T_y, T_z, u_x, u_y, u_z, X, dx, dt,
                                    epsilon):
    """sim/hermite.py:1375-1395, fused into one kernel."""
    nx, ny, nz = _dims(psi)
    _check_mac(u_x, u_y, u_z, nx, ny, nz)
    tab = hermite_table("fd", psi, (T_x, T_y, T_z), None, None, dx)
    with timing.span("k_march_ed4"):
        _lib.call("edge_march_ed4", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), tab.data_ptr(), as_ptr(X), as_ptr(psi),
                  as_ptr(T_x), as_ptr(T_y), as_ptr(T_z), nx, ny, nz, float(dx), float(dt), float(epsilon),
                  stream_handle())
    touch(psi, T_x, T_y, T_z)


def advect_ps_garm(psi_tmp, T_x_tmp, T_y_tmp, T_z_tmp, psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, u_x,
                   u_y, u_z, X, dx, dt):
    """sim/hermite.py:1398-1424: GE march (fused march + compose, then finite differences)."""
    nx, ny, nz = _dims(psi)
    _check_mac(u_x, u_y, u_z, nx, ny, nz)
    tab = hermite_table("full", psi, (T_x, T_y, T_z), (dT_x, dT_y, dT_z), (d2T_x, d2T_y, d2T_z), dx)
    with timing.span("k_march_ge"):
        _lib.call("edge_march_ge", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), tab.data_ptr(), as_ptr(X), as_ptr(psi),
                  as_ptr(T_x), as_ptr(T_y), as_ptr(T_z), nx, ny, nz, float(dx), float(dt), stream_handle())
    touch(psi, T_x, T_y, T_z)
    calc_fd_3_from_grad(T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, dx)


def advect_psi_garm_ed4_fusion_unroll(psi_tmp, T_x_tmp, T_y_tmp, T_z_tmp, dT_x_tmp, dT_y_tmp, dT_z_tmp, psi, T_x, T_y,
                                      T_z, dT_x, dT_y, dT_z, u_x, u_y, u_z, X, dx, dt, epsilon):
    """sim/hermite.py:1426-1521: the EDGE march, one fused kernel, new map written in place."""
    nx, ny, nz = _dims(psi)
    _check_mac(u_x, u_y, u_z, nx, ny, nz)
    T, dT = (T_x, T_y, T_z), (dT_x, dT_y, dT_z)
    tab = hermite_table("garm", psi, T, dT, None, dx)
    # the march also writes the low planes {psi, T} of the new map's table and
    # the high planes {dT, d2T} are packed from the new dT, so the map's next use
    # needs no pass over the stored psi and T
    tab_new = spare_table("garm", psi, T, dT, None, dx)
    with timing.span("k_march_edge"):
        _lib.call("edge_march_edge_table", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), tab.data_ptr(), as_ptr(X),
                  as_ptr(psi), as_ptr(T_x), as_ptr(T_y), as_ptr(T_z), as_ptr(dT_x), as_ptr(dT_y), as_ptr(dT_z),
                  tab_new.data_ptr(), nx, ny, nz, float(dx), float(dt), float(epsilon), stream_handle())
    _lib.call("edge_pack_table_hi", as_ptr(dT_x), as_ptr(dT_y), as_ptr(dT_z), tab_new.data_ptr(), nx, ny, nz,
              float(dx), stream_handle())
    touch(psi, T_x, T_y, T_z, dT_x, dT_y, dT_z)
    install_table("garm", tab_new, psi, T, dT, None, dx)


def advect_psi_garm_ed4(psi_0, T_x_0, T_y_0, T_z_0, psi_1, T_x_1, T_y_1, T_z_1, psi_2, T_x_2, T_y_2, T_z_2, psi_3,
                        T_x_3, T_y_3, T_z_3, psi_tmp, T_x_tmp, T_y_tmp, T_z_tmp, psi, T_x, T_y, T_z, dT_x, dT_y, dT_z,
                        d2T_x, d2T_y, d2T_z, u_x, u_y, u_z, X, dx, dt, epsilon):
    """sim/hermite.py:1523-1576 (unfused, as the reference composes it)."""
    reset_to_identity_1(psi_tmp, T_x_tmp, T_y_tmp, T_z_tmp, X)
    reset_to_identity_garm_ed4(psi_0, T_x_0, T_y_0, T_z_0, psi_1, T_x_1, T_y_1, T_z_1, psi_2, T_x_2, T_y_2, T_z_2,
                               psi_3, T_x_3, T_y_3, T_z_3, X, epsilon)
    pts = [(psi_tmp, T_x_tmp, T_y_tmp, T_z_tmp), (psi_0, T_x_0, T_y_0, T_z_0), (psi_1, T_x_1, T_y_1, T_z_1),
           (psi_2, T_x_2, T_y_2, T_z_2), (psi_3, T_x_3, T_y_3, T_z_3)]
    for p, a, b, c in pts:
        advance_map_1_transpose(p, a, b, c, u_x, u_y, u_z, dx, -dt)
    for p, a, b, c in pts:
        interp_psi_hermite_garm(p, a, b, c, psi, T_x, T_y, T_z, dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, dx)
    psi.copy_from(psi_tmp)
    T_x.copy_from(T_x_tmp)
    T_y.copy_from(T_y_tmp)
    T_z.copy_from(T_z_tmp)
    calc_ed4_garm(T_x_0, T_y_0, T_z_0, T_x_1, T_y_1, T_z_1, T_x_2, T_y_2, T_z_2, T_x_3, T_y_3, T_z_3, dT_x, dT_y, dT_z,
                  epsilon)
    calc_fd_3_from_grad_grad(dT_x, dT_y, dT_z, d2T_x, d2T_y, d2T_z, dx)


def advect_u_semi_lagrangian(u_x0, u_y0, u_z0, u_x1, u_y1, u_z1, dx, dt, X_x, X_y, X_z):
    """sim/hermite.py:1578-1621.  X_x/X_y/X_z must be the face coordinates
    (x_coords_func etc.); the kernel recomputes them from the indices."""
    nx, ny, nz = u_x0.shape[0] - 1, u_x0.shape[1], u_x0.shape[2]
    _check_mac(u_x0, u_y0, u_z0, nx, ny, nz)
    with timing.span("k_advect_sl"):
        _lib.call("edge_advect_u_sl", as_ptr(u_x0), as_ptr(u_y0), as_ptr(u_z0), as_ptr(u_x1), as_ptr(u_y1),
                  as_ptr(u_z1), nx, ny, nz, float(dx), float(dt), stream_handle())
    touch(u_x1, u_y1, u_z1)
