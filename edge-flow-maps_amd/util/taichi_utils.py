"""Drop-in for the kernels of the reference's util/taichi_utils.py.

The @ti.func interpolators of that file (sample, interp_2, interp_grad_2,
interp_u_MAC(_grad), interp_1, sample_min_max_1) are device code in
libedge_hip.so (csrc/edge_device.h); the @ti.kernel entry points the step
loop calls are here with the reference's names and arguments.
"""
from edge_hip import _lib
from edge_hip.field import Field, as_ptr, stream_handle, touch, ti  # noqa: F401  (ti re-exported)

eps = 1.e-6  # util/taichi_utils.py:3


def scale_field(a, alpha, result):
    """util/taichi_utils.py:6-9: result = alpha * a."""
    assert a.data.shape == result.data.shape
    _lib.call("edge_scale", as_ptr(a), float(alpha), as_ptr(result), result.numel, stream_handle())
    touch(result)


def add_fields(f1, f2, dest, multiplier):
    """util/taichi_utils.py:12-15: dest = f1 + multiplier * f2."""
    assert f1.data.shape == f2.data.shape == dest.data.shape
    _lib.call("edge_axpy", as_ptr(f1), as_ptr(f2), as_ptr(dest), float(multiplier), dest.numel, stream_handle())
    touch(dest)


def add_scale_fields(f1, f2, dest, multiplier, scale):
    """add_fields(f1, f2, dest, multiplier) then scale_field(dest, scale, dest)
    (run_edge.py:675-681) in one pass: the same values."""
    assert f1.data.shape == f2.data.shape == dest.data.shape
    _lib.call("edge_axpy_scale", as_ptr(f1), as_ptr(f2), as_ptr(dest), float(multiplier), float(scale), dest.numel,
              stream_handle())
    touch(dest)


def _coords(mode, pf, dx, cell_shape):
    nx, ny, nz = cell_shape
    _lib.call("edge_coords", mode, as_ptr(pf), nx, ny, nz, float(dx), stream_handle())
    touch(pf)


def center_coords_func(pf, dx):
    """util/taichi_utils.py:23-26."""
    _coords(0, pf, dx, pf.shape)


def x_coords_func(pf, dx):
    """util/taichi_utils.py:29-32 (pf on the x-face grid)."""
    n0, n1, n2 = pf.shape
    _coords(1, pf, dx, (n0 - 1, n1, n2))


def y_coords_func(pf, dx):
    """util/taichi_utils.py:35-38."""
    n0, n1, n2 = pf.shape
    _coords(2, pf, dx, (n0, n1 - 1, n2))


def z_coords_func(pf, dx):
    """util/taichi_utils.py:41-44."""
    n0, n1, n2 = pf.shape
    _coords(3, pf, dx, (n0, n1, n2 - 1))


def split_central_vector(vc, vx, vy, vz):
    """util/taichi_utils.py:103-116: cell-centred vector -> MAC faces."""
    nx, ny, nz = vc.shape
    _lib.call("edge_split_central_vector", as_ptr(vc), as_ptr(vx), as_ptr(vy), as_ptr(vz), nx, ny, nz,
              stream_handle())
    touch(vx, vy, vz)


def get_central_vector(vx, vy, vz, vc):
    """util/taichi_utils.py:95-100: MAC faces -> cell-centred vector (face means)."""
    nx, ny, nz = vc.shape
    assert vx.shape == (nx + 1, ny, nz) and vy.shape == (nx, ny + 1, nz) and vz.shape == (nx, ny, nz + 1)
    _lib.call("edge_get_central_vector", as_ptr(vx), as_ptr(vy), as_ptr(vz), as_ptr(vc), nx, ny, nz,
              stream_handle())
    touch(vc)


def curl(vf, cf, dx):
    """util/taichi_utils.py:70-92: cf = curl of the cell-centred vector field vf
    (central differences, neighbour index clamped to the grid as sample() does)."""
    nx, ny, nz = cf.shape
    assert vf.shape == cf.shape and vf.n == cf.n == 3
    _lib.call("edge_curl", as_ptr(vf), as_ptr(cf), nx, ny, nz, float(dx), stream_handle())
    touch(cf)
