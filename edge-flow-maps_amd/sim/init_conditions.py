"""Initial conditions (sim/init_conditions.py of the reference), on the GPU.

add_vortex_ring_and_smoke (sim/init_conditions.py:13-36) is the HIP kernel
edge_vortex_ring; the presets init_four_vorts (38-67) and init_vorts_leapfrog
(70-84) keep the reference's parameters.  `init_rings` builds the seeded
vortex-ring flows used by the benchmark configurations.
"""
import ctypes
import math

import numpy as np

from edge_hip import _lib
from edge_hip.field import as_ptr, stream_handle, touch
from util.taichi_utils import add_fields


def _f3(v):
    return ctypes.cast((ctypes.c_float * 3)(*[float(x) for x in v]), ctypes.c_void_p)


def add_vortex_ring_and_smoke(w, rad, delta, c, unit_x, unit_y, pf, vf, smokef, color, num_samples):
    """sim/init_conditions.py:13-36: adds the ring's velocity to vf; smokef[3]
    accumulates the core density then is thresholded (colour where > 0.002)."""
    nx, ny, nz = pf.shape
    _lib.call("edge_vortex_ring", float(w), float(rad), float(delta), _f3(c), _f3(unit_x), _f3(unit_y), as_ptr(pf),
              as_ptr(vf), as_ptr(smokef), _f3(color), int(num_samples), nx, ny, nz, stream_handle())
    touch(vf, smokef)


def _normalized(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def init_four_vorts(X, u, smoke1, smoke2):
    """sim/init_conditions.py:38-67."""
    smoke1.fill(0.)
    smoke2.fill(0.)
    x_offset, y_offset, size = 0.16, 0.16, 0.15
    c45 = math.cos(math.pi / 4)
    rings = [(2.e-2, (0.5 - x_offset, 0.5 - y_offset, 1), (-c45, c45, 0.), (1, 0, 0)),
             (-2.e-2, (0.5 + x_offset, 0.5 - y_offset, 1), (c45, c45, 0.), (0, 1, 0)),
             (2.e-2, (0.5 - x_offset, 0.5 + y_offset, 1), (c45, c45, 0.), (0, 0, 1)),
             (-2.e-2, (0.5 + x_offset, 0.5 + y_offset, 1), (-c45, c45, 0.), (0, 0, 0))]
    for q, (w, c, ux, col) in enumerate(rings):
        target = smoke1 if q == 0 else smoke2
        add_vortex_ring_and_smoke(w, size, 0.024, c, _normalized(ux), (0., 0., 1.), X, u, target, col, 500)
        if q > 0:
            add_fields(smoke1, smoke2, smoke1, 1.0)
            smoke2.fill(0.)


def init_vorts_leapfrog(X, u, smoke1, smoke2, radius=0.21, x_gap=None, x_start=0.16, delta=None, w=None,
                        num_samples=2000):
    """sim/init_conditions.py:70-84 (defaults = the reference's parameters)."""
    x_gap = 0.625 * radius if x_gap is None else x_gap
    delta = 0.08 * radius if delta is None else delta
    w = radius * 0.1 if w is None else w
    add_vortex_ring_and_smoke(w, radius, delta, (x_start, 0.5, 0.5), (0., 0., -1.), (0., 1, 0.), X, u, smoke1,
                              (1., 0, 0), num_samples)
    add_vortex_ring_and_smoke(w, radius, delta, (x_start + x_gap, 0.5, 0.5), (0., 0., -1.), (0., 1, 0.), X, u,
                              smoke2, (0, 0, 1.), num_samples)
    add_fields(smoke1, smoke2, smoke1, 1.0)


def ring_frame(axis):
    """Two unit vectors spanning the plane normal to `axis`."""
    a = _normalized(axis)
    t = np.array([0., 0., 1.]) if abs(a[2]) < 0.9 else np.array([1., 0., 0.])
    ex = _normalized(np.cross(a, t))
    ey = np.cross(a, ex)
    return ex, ey


def seeded_rings(n_rings, domain, seed=0, radius=0.15, delta=0.025, strength=0.02):
    """Ring centres uniform in the middle of the domain, axes pointing at its centre."""
    rng = np.random.default_rng(seed)
    L = np.asarray(domain, float)
    out = []
    centre = L / 2
    for q in range(n_rings):
        c = np.array([rng.uniform(0.25, 0.75) * L[0], rng.uniform(0.3, 0.7) * L[1], rng.uniform(0.3, 0.7) * L[2]])
        axis = centre - c
        if np.linalg.norm(axis) < 1e-6:
            axis = np.array([1., 0., 0.])
        ex, ey = ring_frame(axis)
        sgn = 1.0 if q % 2 == 0 else -1.0
        out.append(dict(w=sgn * strength, rad=radius, delta=delta, c=c, ex=ex, ey=ey))
    return out


def init_rings(X, u, smoke1, smoke2, rings, num_samples=2000):
    """Superpose rings (each as add_vortex_ring_and_smoke) and their dye."""
    smoke1.fill(0.)
    colours = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0)]
    for q, r in enumerate(rings):
        smoke2.fill(0.)
        add_vortex_ring_and_smoke(r["w"], r["rad"], r["delta"], r["c"], r["ex"], r["ey"], X, u, smoke2,
                                  colours[q % 4], num_samples)
        add_fields(smoke1, smoke2, smoke1, 1.0)


# ---------------------------------------------------------------------------
# Delta-wing plate in an oblique inflow (sim/init_conditions.py:127-175)
# ---------------------------------------------------------------------------
DELTA_SPEED = 0.6
DELTA_ATTACK_DEG = 20.0


def voxelize_triangle(vertices, dx, subsamples=4):
    """Cells (i, j) of a pitch-dx grid covered by a triangle in the z = 0 plane.

    The reference voxelizes its plate with trimesh (`mesh.voxelized(dx)`, not
    available here); this marks every cell of the triangle's bounding box with
    at least one of subsamples^2 sample points inside the triangle.  Returns
    the (nx, ny) bool array and the grid origin (x0, y0) of cell (0, 0)."""
    v = np.asarray(vertices, np.float64)[:, :2]
    lo = np.floor(v.min(0) / dx) * dx
    hi = np.ceil(v.max(0) / dx) * dx
    n = np.maximum(np.round((hi - lo) / dx).astype(int), 1)
    s = (np.arange(subsamples) + 0.5) / subsamples
    gi, gj, si, sj = np.meshgrid(np.arange(n[0]), np.arange(n[1]), s, s, indexing="ij")
    px = lo[0] + (gi + si) * dx
    py = lo[1] + (gj + sj) * dx

    def side(a, b):
        return (b[0] - a[0]) * (py - a[1]) - (b[1] - a[1]) * (px - a[0])
    d0, d1, d2 = side(v[0], v[1]), side(v[1], v[2]), side(v[2], v[0])
    inside = ((d0 >= 0) & (d1 >= 0) & (d2 >= 0)) | ((d0 <= 0) & (d1 <= 0) & (d2 <= 0))
    return inside.any(axis=(2, 3)), lo


def set_boundary_mask_delta(X, boundary_mask, boundary_vel, _t, dx):
    """sim/init_conditions.py:127-167: the triangular plate (half-angle 15
    degrees, length 1, four cells thick, leading edge at (0.1, 0.5, 0.2)) and
    the box's outer cell layer as solid; the outer layer moves with the
    inflow (0.6 cos 20, 0, 0.6 sin 20), the plate is at rest."""
    shape = boundary_mask.shape
    ang = np.radians(15.0)
    vert = np.array([[0.0, 0.0, 0.0], [np.cos(ang), np.sin(ang), 0.0], [np.cos(ang), -np.sin(ang), 0.0]])
    plate, _ = voxelize_triangle(vert, dx)
    size = plate.shape
    # a flat triangle voxelizes one cell thick: z offset 0.2 // dx (- 1 // 2)
    off = np.array([int(0.1 // dx), int(0.5 // dx) - size[1] // 2, int(0.2 // dx)], np.int64)
    mask = np.zeros(shape, np.int32)
    sx = min(size[0], shape[0] - off[0])
    sy = min(size[1], shape[1] - off[1])
    for t in range(4):
        mask[off[0]:off[0] + sx, off[1]:off[1] + sy, off[2] + t] += plate[:sx, :sy]
    mask[mask > 0] = 1
    vel = np.zeros(shape + (3,), np.float32)
    a = np.radians(DELTA_ATTACK_DEG)
    edge = np.zeros(shape, bool)
    edge[[0, -1], :, :] = True
    edge[:, [0, -1], :] = True
    edge[:, :, [0, -1]] = True
    mask[edge] = 1
    vel[edge] = (DELTA_SPEED * np.cos(a), 0.0, DELTA_SPEED * np.sin(a))
    boundary_mask.from_numpy(mask)
    boundary_vel.from_numpy(vel)


def init_delta(X, u, smoke1, smoke2):
    """sim/init_conditions.py:170-175: uniform oblique inflow, no dye."""
    a = np.radians(DELTA_ATTACK_DEG)
    n = u.shape
    v = np.empty(n + (3,), np.float32)
    v[...] = (DELTA_SPEED * np.cos(a), 0.0, DELTA_SPEED * np.sin(a))
    u.from_numpy(v)
    smoke1.fill(0.0)
