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


def voxelize_triangle(vertices, pitch, max_iter=10):
    """The reference's `trimesh.Trimesh(vertices, [[0, 1, 2]]).voxelized(pitch).matrix`
    (sim/init_conditions.py:139-150), computed here without trimesh: trimesh's
    "subdivide" voxelizer splits every triangle with an edge longer than pitch/2
    into four at its edge midpoints ((a + b) / 2, up to max_iter times), keeps
    the corners of the triangles that are small enough, and marks the voxel
    round(corner / pitch) of each (round half to even).  Triangles are carried
    as coordinate triples; edge lengths are sqrt((dx^2 + dy^2) + dz^2) like
    trimesh's, so the same triangles stop at the same pass.  Returns the dense
    bool matrix over the occupied index range and its origin index."""
    tri = np.asarray(vertices, np.float64)[None]          # (t, 3 corners, 3)
    kept = []
    max_edge = pitch / 2.0
    for _ in range(max_iter + 1):
        d = tri[:, [1, 2, 0]] - tri                       # the three edge vectors
        long_ = ((d ** 2).sum(axis=2) ** 0.5 > max_edge).any(axis=1)
        kept.append(tri[~long_].reshape(-1, 3))
        if not long_.any():
            break
        t = tri[long_]
        m01 = (t[:, 0] + t[:, 1]) / 2.0
        m12 = (t[:, 1] + t[:, 2]) / 2.0
        m20 = (t[:, 2] + t[:, 0]) / 2.0
        tri = np.concatenate([np.stack([t[:, 0], m01, m20], 1), np.stack([m01, t[:, 1], m12], 1),
                              np.stack([m20, m12, t[:, 2]], 1), np.stack([m01, m12, m20], 1)])
    hit = np.unique(np.round(np.concatenate(kept) / pitch).astype(np.int64), axis=0)
    lo = hit.min(axis=0)
    m = np.zeros(tuple(hit.max(axis=0) - lo + 1), bool)
    m[tuple((hit - lo).T)] = True
    return m, lo


def set_boundary_mask_delta(X, boundary_mask, boundary_vel, _t, dx):
    """sim/init_conditions.py:127-167: the triangular plate (half-angle 15
    degrees, length 1, four cells thick, leading edge at (0.1, 0.5, 0.2)) and
    the box's outer cell layer as solid; the outer layer moves with the
    inflow (0.6 cos 20, 0, 0.6 sin 20), the plate is at rest."""
    shape = boundary_mask.shape
    ang = np.radians(15)
    x, y = np.cos(ang) * 1.0, np.sin(ang) * 1.0
    vert = np.array([[0, 0, 0], [x, y, 0], [x, -y, 0]])
    plate, _ = voxelize_triangle(vert, dx)   # (sx, sy, 1): the plate is flat
    size = plate.shape
    off = np.array([0.1 // dx, 0.5 // dx - size[1] // 2, 0.2 // dx - size[2] // 2], np.int64)
    mask = np.zeros(shape, np.int32)
    sx = min(size[0], shape[0] - off[0])
    sy = min(size[1], shape[1] - off[1])
    for t in range(4):   # extruded four cells thick
        mask[off[0]:off[0] + sx, off[1]:off[1] + sy, off[2] + t] += plate[:sx, :sy, 0]
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


# ---------------------------------------------------------------------------
# The triangular plate of the 256x128x128 solid workload: the reference's
# delta shape (half-angle 15 degrees, flat, four cells thick, the outer cell
# layer moving with the inflow) at span 0.3, in an inflow of speed 1 at a
# 20-degree angle of attack, added to the four seeded rings
# ---------------------------------------------------------------------------
PLATE_SPAN = 0.3
PLATE_SPEED = 1.0
PLATE_ATTACK_DEG = 20.0
PLATE_HALF_ANGLE_DEG = 15.0
PLATE_APEX = (0.3, 0.5, 0.5)   # leading-edge apex; the plate spans y symmetrically, centred in z


def plate_vertices(span=PLATE_SPAN, half_angle_deg=PLATE_HALF_ANGLE_DEG):
    """Apex at the origin, trailing edge of width `span` downstream (+x)."""
    a = np.radians(half_angle_deg)
    length = span / (2.0 * np.sin(a))
    x, y = np.cos(a) * length, np.sin(a) * length
    return np.array([[0.0, 0.0, 0.0], [x, y, 0.0], [x, -y, 0.0]])


def plate_mask(shape, dx, span=PLATE_SPAN, apex=PLATE_APEX, thickness=4):
    """The voxelised plate (the trimesh rule of voxelize_triangle), apex cell at
    apex // dx, centred in y and z, `thickness` cells thick in z; the box's
    outer layer is not included."""
    plate, _ = voxelize_triangle(plate_vertices(span), dx)
    size = plate.shape
    off = np.array([apex[0] // dx, apex[1] // dx - size[1] // 2, round(apex[2] / dx) - thickness // 2], np.int64)
    if np.any(off < 1) or off[0] + size[0] > shape[0] - 1 or off[1] + size[1] > shape[1] - 1 or \
            off[2] + thickness > shape[2] - 1:
        raise ValueError(f"the plate (span {span}) does not fit the {shape} grid at dx {dx}")
    mask = np.zeros(shape, np.int32)
    for t in range(thickness):
        mask[off[0]:off[0] + size[0], off[1]:off[1] + size[1], off[2] + t] |= plate[:, :, 0].astype(np.int32)
    return mask


def plate_inflow(speed=PLATE_SPEED, attack_deg=PLATE_ATTACK_DEG):
    a = np.radians(attack_deg)
    return np.array([speed * np.cos(a), 0.0, speed * np.sin(a)], np.float32)


def set_boundary_mask_plate(boundary_mask, boundary_vel, dx):
    """The plate at rest and the box's outer cell layer, which moves with the
    inflow (the reference's delta-case boundary, sim/init_conditions.py:127-167)."""
    shape = boundary_mask.shape
    mask = plate_mask(shape, dx)
    vel = np.zeros(shape + (3,), np.float32)
    edge = np.zeros(shape, bool)
    edge[[0, -1], :, :] = True
    edge[:, [0, -1], :] = True
    edge[:, :, [0, -1]] = True
    mask[edge] = 1
    vel[edge] = plate_inflow()
    boundary_mask.from_numpy(mask)
    boundary_vel.from_numpy(vel)


def add_inflow(u):
    """u += the uniform inflow (cell-centred field)."""
    v = u.to_numpy()
    v += plate_inflow()
    u.from_numpy(v)
