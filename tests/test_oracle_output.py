"""Pins of the oracle's frame-output kernels (util/taichi_utils.py:70-100):
closed forms on linear and quadratic fields, the clamped boundary rule, and
an independent NumPy restatement on random fields."""
import numpy as np


def _cell_centres(n, dx):
    i, j, k = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    return (i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx


def _face_centres(a, n, dx):
    dims = [m + (b == a) for b, m in enumerate(n)]
    idx = np.meshgrid(*[np.arange(m) for m in dims], indexing="ij")
    return [(idx[b] + (0.0 if b == a else 0.5)) * dx for b in range(3)]


def test_central_vector_of_a_linear_field_is_its_value_at_the_centre(orc64):
    n, dx = (7, 5, 6), 0.1
    G = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [-0.6, 0.9, 1.5]])
    c0 = np.array([0.2, -0.5, 1.1])
    u = []
    for a in range(3):
        x = _face_centres(a, n, dx)
        u.append(c0[a] + G[a, 0] * x[0] + G[a, 1] * x[1] + G[a, 2] * x[2])
    vc = orc64.central_vector(u)
    X = _cell_centres(n, dx)
    for a in range(3):
        want = c0[a] + G[a, 0] * X[0] + G[a, 1] * X[1] + G[a, 2] * X[2]
        assert np.max(np.abs(vc[a] - want)) < 1e-13


def test_curl_of_a_rotation_with_the_clamped_edges(orc64):
    # v = w (-y, x, 0): curl = (0, 0, 2w) inside; on a boundary layer the
    # clamped neighbour halves that derivative: each of d vy/dx and -d vx/dy
    # contributes w inside and w/2 on its boundary layers.
    n, dx, w = (6, 5, 4), 0.25, 1.3
    X = _cell_centres(n, dx)
    vf = np.stack([-w * X[1], w * X[0], np.zeros(n)])
    cf = orc64.curl(vf, dx)
    i, j, _ = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    dvy_dx = np.where((i == 0) | (i == n[0] - 1), w / 2, w)
    dvx_dy = np.where((j == 0) | (j == n[1] - 1), -w / 2, -w)
    assert np.max(np.abs(cf[0])) < 1e-13 and np.max(np.abs(cf[1])) < 1e-13
    assert np.max(np.abs(cf[2] - (dvy_dx - dvx_dy))) < 1e-12


def test_curl_exact_on_quadratics_inside(orc64):
    # central differences are exact for quadratics: v = (y^2, z^2, x^2) has
    # curl (-2z, -2x, -2y)
    n, dx = (8, 7, 6), 0.2
    X = _cell_centres(n, dx)
    vf = np.stack([X[1] ** 2, X[2] ** 2, X[0] ** 2])
    cf = orc64.curl(vf, dx)
    inner = (slice(1, -1),) * 3
    want = [-2 * X[2], -2 * X[0], -2 * X[1]]
    for a in range(3):
        assert np.max(np.abs(cf[a][inner] - want[a][inner])) < 1e-11


def _np_curl(v, dx):
    """Independent restatement: neighbours by clipped index arrays."""
    n = v.shape[1:]
    idx = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")

    def nb(c, a, s):
        ii = [x.copy() for x in idx]
        ii[a] = np.clip(ii[a] + s, 0, n[a] - 1)
        return v[c][tuple(ii)]

    h = 1.0 / (2 * dx)
    d = lambda c, a: h * (nb(c, a, 1) - nb(c, a, -1))  # noqa: E731
    return np.stack([d(2, 1) - d(1, 2), d(0, 2) - d(2, 0), d(1, 0) - d(0, 1)])


def test_curl_and_central_vector_equal_numpy_on_random_fields(orc64, orc32):
    rng = np.random.default_rng(3)
    n, dx = (9, 6, 5), 1.0 / 16
    u = [rng.normal(size=(n[0] + 1, n[1], n[2])), rng.normal(size=(n[0], n[1] + 1, n[2])),
         rng.normal(size=(n[0], n[1], n[2] + 1))]
    vc = orc64.central_vector(u)
    want = np.stack([0.5 * (u[0][1:] + u[0][:-1]), 0.5 * (u[1][:, 1:] + u[1][:, :-1]),
                     0.5 * (u[2][:, :, 1:] + u[2][:, :, :-1])])
    assert np.max(np.abs(vc - want)) < 1e-14
    assert np.max(np.abs(orc64.curl(vc, dx) - _np_curl(vc, dx))) < 1e-12
    # the FP32 restatement follows the same formula within FP32 rounding
    c32 = orc32.curl(vc.astype(np.float32), dx)
    assert np.max(np.abs(c32 - _np_curl(vc, dx))) < 1e-4 * np.max(np.abs(_np_curl(vc, dx)))
