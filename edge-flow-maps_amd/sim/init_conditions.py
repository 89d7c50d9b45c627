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
