"""Restatement of trimesh's surface voxelizer, the one the reference's delta
case calls (`mesh.voxelized(dx)`, sim/init_conditions.py:139-150).

TEST INFRASTRUCTURE ONLY.  trimesh is a dependency of the reference (its
README installs it unpinned with `pip install trimesh`) and is not available
here.  Its `Trimesh.voxelized(pitch)` uses the "subdivide" method, unchanged
across the 3.x and 4.x releases:
  * trimesh.remesh.subdivide_to_size(vertices, faces, max_edge=pitch / 2,
    max_iter=10): up to max_iter + 1 passes; in each pass the faces whose
    three edge lengths (Euclidean, from the face's vertex cycle 0-1-2-0) are
    all <= max_edge are kept as done (with their vertices), the others are
    split by trimesh.remesh.subdivide into 4 at the edge midpoints (the mean of
    the edge's two endpoints, one midpoint per unique edge);
  * trimesh.voxel.creation.voxelize_subdivide: every kept vertex v goes to the
    voxel round(v / pitch) (numpy's round half to even), duplicates merged,
    and the dense `matrix` spans the occupied indices from their minimum.
The arithmetic below is the same numpy expressions, so the vertex positions,
the edge-length tests and the rounding are bit for bit those of trimesh.
"""
import numpy as np


def _subdivide(vertices, faces):
    """trimesh.remesh.subdivide: each face split into 4 at its edge midpoints."""
    edges = np.sort(faces[:, [0, 1, 1, 2, 2, 0]].reshape((-1, 2)), axis=1)
    unique, inverse = np.unique(edges, axis=0, return_inverse=True)
    mid = vertices[unique].mean(axis=1)
    mid_idx = inverse.reshape((-1, 3)) + len(vertices)
    f = np.column_stack([faces[:, 0], mid_idx[:, 0], mid_idx[:, 2],
                         mid_idx[:, 0], faces[:, 1], mid_idx[:, 1],
                         mid_idx[:, 2], mid_idx[:, 1], faces[:, 2],
                         mid_idx[:, 0], mid_idx[:, 1], mid_idx[:, 2]]).reshape((-1, 3))
    return np.vstack((vertices, mid)), f


def subdivide_to_size(vertices, faces, max_edge, max_iter=10):
    """trimesh.remesh.subdivide_to_size: the vertices of the faces kept."""
    done = []
    cur_v = np.array(vertices, dtype=np.float64, copy=True)
    cur_f = np.array(faces, dtype=np.int64, copy=True)
    for _ in range(max_iter + 1):
        edge_length = (np.diff(cur_v[cur_f[:, [0, 1, 2, 0]]], axis=1) ** 2).sum(axis=2) ** 0.5
        too_long = (edge_length > max_edge).any(axis=1)
        ok = ~too_long
        done.append(cur_v[np.unique(cur_f[ok].ravel())])
        if not too_long.any():
            break
        cur_v, cur_f = _subdivide(cur_v, cur_f[too_long])
    return np.vstack(done)


def voxelize_subdivide(vertices, faces, pitch, max_iter=10, edge_factor=2.0):
    """trimesh.voxel.creation.voxelize_subdivide: (dense bool matrix, origin index)."""
    v = subdivide_to_size(vertices, faces, max_edge=pitch / edge_factor, max_iter=max_iter)
    hit = np.round(v / pitch).astype(int)
    occ = np.unique(hit, axis=0)
    origin = occ.min(axis=0)
    idx = occ - origin
    m = np.zeros(idx.max(axis=0) + 1, dtype=bool)
    m[tuple(idx.T)] = True
    return m, origin


def delta_plate_mask(shape, dx):
    """sim/init_conditions.py:139-165: the reference's plate mask (without the
    box's outer layer), its voxelized triangle placed four cells thick."""
    angle = np.radians(15)
    length = 1.0
    x = np.cos(angle) * length
    y = np.sin(angle) * length
    vert = np.array([[0, 0, 0], [x, y, 0], [x, -y, 0]])
    face = np.array([[0, 1, 2]])
    obj, _ = voxelize_subdivide(vert, face, dx)
    size = obj.shape
    offset = np.array((0.1 // dx, 0.5 // dx - size[1] // 2, 0.2 // dx - size[2] // 2), dtype=np.int32)
    mask = np.zeros(shape, dtype=np.int32)
    for t in range(4):
        mask[offset[0]:offset[0] + size[0], offset[1]:offset[1] + size[1], offset[2] + t:offset[2] + t + 1] += obj
    mask[mask > 0] = 1
    return mask
