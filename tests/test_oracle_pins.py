"""Pin the CPU oracle: closed forms, known answers and an independent NumPy
restatement, chosen so that a dropped term, a wrong sign or a transposed
operand in the oracle fails at least one of them.  CPU only."""
import numpy as np
import pytest

from helpers import (PolyMap, any_points, cell_coords, identity_plus, interior_points, linear_velocity,
                     smooth_velocity, taylor_P)
from oracle import np_ref
from oracle.oracle import HK_FD, HK_FULL, HK_GARM_FD

NX, NY, NZ, DX = 12, 10, 8, 1.0 / 10


# ---------------------------------------------------------------- B-spline
@pytest.mark.parametrize("BL", [(0, .5, .5), (.5, 0, .5), (.5, .5, 0), (.5, .5, .5)])
def test_bspline_matches_independent_restatement(orc64, BL):
    rng = np.random.default_rng(1)
    f = rng.normal(size=(NX + 1, NY, NZ))
    pts = any_points(NX, NY, NZ, DX, 400, seed=2)
    v, g = orc64.eval_bspline(f, BL, pts, DX, grad=True)
    v2, g2 = np_ref.bspline(f, BL, pts, DX, grad=True)
    np.testing.assert_allclose(v, v2, rtol=0, atol=1e-12)
    np.testing.assert_allclose(g, g2, rtol=0, atol=1e-10)


def test_bspline_reproduces_affine_and_constants(orc64):
    A = np.array([[0.3, -1.2, 0.5], [2.0, 0.1, -0.7], [-0.4, 0.9, 1.1]])
    b = np.array([0.2, -0.5, 1.0])
    u = linear_velocity(A, b, NX, NY, NZ, DX)
    pts = interior_points(NX, NY, NZ, DX, 300, margin=2.0)
    val, J = orc64.eval_umac(u, pts, DX, grad=True)
    np.testing.assert_allclose(val, pts @ A.T + b, atol=1e-12)
    np.testing.assert_allclose(J, np.broadcast_to(A, J.shape), atol=1e-10)
    # partition of unity, including the clamp zone
    c = np.full((NX + 1, NY, NZ), 3.25)
    v, g = orc64.eval_bspline(c, (0, .5, .5), any_points(NX, NY, NZ, DX, 200), DX, grad=True)
    np.testing.assert_allclose(v, 3.25, atol=1e-13)
    np.testing.assert_allclose(g, 0, atol=1e-11)


def test_bspline_is_not_interpolatory(orc64):
    # the reference's quadratic B-spline smooths: at a grid point the weights are
    # 1/8, 3/4, 1/8 per axis (util/taichi_utils.py:230-238), not a Kronecker delta
    f = np.zeros((NX, NY, NZ))
    f[5, 5, 4] = 1.0
    p = np.array([[(5 + .5) * DX, (5 + .5) * DX, (4 + .5) * DX]])
    assert orc64.eval_bspline(f, (.5, .5, .5), p, DX)[0] == pytest.approx(0.75 ** 3, abs=1e-14)
    p2 = np.array([[(6 + .5) * DX, (5 + .5) * DX, (4 + .5) * DX]])
    assert orc64.eval_bspline(f, (.5, .5, .5), p2, DX)[0] == pytest.approx(0.125 * 0.75 ** 2, abs=1e-14)


def test_linear_interp_and_minmax(orc64):
    rng = np.random.default_rng(3)
    s = rng.normal(size=(4, NX, NY, NZ))
    pts = interior_points(NX, NY, NZ, DX, 100, margin=1.6)
    val, mn, mx = orc64.eval_linear(s, pts, DX)
    # trilinear of an affine field is exact
    X = cell_coords(NX, NY, NZ, DX)
    aff = 1.0 + 2 * X[0] - X[1] + 0.5 * X[2]
    v2, _, _ = orc64.eval_linear(aff[None], pts, DX)
    np.testing.assert_allclose(v2[:, 0], 1.0 + 2 * pts[:, 0] - pts[:, 1] + 0.5 * pts[:, 2], atol=1e-12)
    assert np.all(mn <= val + 1e-12) and np.all(val <= mx + 1e-12)


# ---------------------------------------------------------------- Hermite
def _poly_tricubic(seed=0):
    rng = np.random.default_rng(seed)
    E = [(a, b, c) for a in range(4) for b in range(4) for c in range(4) if a + b + c <= 5]
    C = rng.normal(size=(3, len(E))) * 0.3
    return PolyMap(E, C)


def test_full_hermite_reproduces_tricubics(orc64):
    pm = _poly_tricubic()
    psi, T, dT, d2T = pm.tables(NX, NY, NZ, DX)
    pts = interior_points(NX, NY, NZ, DX, 300, margin=1.6)
    v, G = orc64.eval_hermite(HK_FULL, psi, T, dT, d2T, pts, DX, grad=True)
    P = pts.T.reshape(3, -1, 1, 1)
    np.testing.assert_allclose(v, pm.value(P)[:, :, 0, 0].T, atol=1e-9)
    np.testing.assert_allclose(G, np.moveaxis(pm.grad_rows(P)[:, :, :, 0, 0], 2, 0), atol=1e-7)


def test_hermite_matches_independent_restatement(orc64):
    rng = np.random.default_rng(4)
    psi = rng.normal(size=(3, NX, NY, NZ))
    T = rng.normal(size=(3, 3, NX, NY, NZ))
    dT = rng.normal(size=(3, 3, NX, NY, NZ))
    d2T = rng.normal(size=(3, NX, NY, NZ))
    pts = any_points(NX, NY, NZ, DX, 300, seed=5)
    v, G = orc64.eval_hermite(HK_FULL, psi, T, dT, d2T, pts, DX, grad=True)
    v2, G2 = np_ref.hermite_full(psi, T, dT, d2T, pts, DX, grad=True)
    np.testing.assert_allclose(v, v2, atol=1e-10)
    np.testing.assert_allclose(G, G2, atol=1e-8)


def test_garm_equals_full_with_fd_d2T(orc64):
    # sim/hermite.py:340-345: the garm xyz term is d2T of calc_fd_3_from_grad_grad
    rng = np.random.default_rng(6)
    psi = rng.normal(size=(3, NX, NY, NZ))
    T = rng.normal(size=(3, 3, NX, NY, NZ))
    dT = rng.normal(size=(3, 3, NX, NY, NZ))
    _, d2T = orc64.fd3(None, dT, DX, want_dT=False)
    pts = any_points(NX, NY, NZ, DX, 300, seed=7)
    a, Ga = orc64.eval_hermite(HK_GARM_FD, psi, T, dT, None, pts, DX, grad=True)
    b, Gb = orc64.eval_hermite(HK_FULL, psi, T, dT, d2T, pts, DX, grad=True)
    np.testing.assert_allclose(a, b, atol=1e-11)
    np.testing.assert_allclose(Ga, Gb, atol=1e-9)


def test_fd_unroll_equals_full_with_fd_data(orc64):
    # sim/hermite.py:111-158: the FD variant's cross terms are calc_fd_3_from_grad of T
    rng = np.random.default_rng(8)
    psi = rng.normal(size=(3, NX, NY, NZ))
    T = rng.normal(size=(3, 3, NX, NY, NZ))
    dT, d2T = orc64.fd3(T, None, DX, want_dT=True)
    pts = any_points(NX, NY, NZ, DX, 300, seed=9)
    a, Ga = orc64.eval_hermite(HK_FD, psi, T, None, None, pts, DX, grad=True)
    b, Gb = orc64.eval_hermite(HK_FULL, psi, T, dT, d2T, pts, DX, grad=True)
    np.testing.assert_allclose(a, b, atol=1e-11)
    np.testing.assert_allclose(Ga, Gb, atol=1e-9)


def test_fd3_exact_on_polynomials(orc64):
    # central differences are exact for quadratics: dT of a cubic's gradient
    pm = PolyMap([(1, 1, 0), (0, 1, 1), (1, 0, 1), (2, 1, 0), (0, 2, 1), (1, 1, 1)],
                 np.random.default_rng(10).normal(size=(3, 6)))
    X = cell_coords(NX, NY, NZ, DX)
    dT, d2T = orc64.fd3(pm.grad_rows(X), None, DX, want_dT=True)
    s = (slice(2, -2),) * 3
    np.testing.assert_allclose(dT[(slice(None), slice(None)) + s], pm.cross(X)[(slice(None), slice(None)) + s],
                               atol=1e-10)
    np.testing.assert_allclose(d2T[(slice(None),) + s], pm.xyz(X)[(slice(None),) + s], atol=1e-8)


def test_identity_map_returns_points(orc32):
    X = cell_coords(NX, NY, NZ, DX)
    T = np.zeros((3, 3, NX, NY, NZ))
    for c in range(3):
        T[c, c] = 1
    dT = np.zeros_like(T)
    pts = interior_points(NX, NY, NZ, DX, 200, margin=1.6)
    for kind in (HK_GARM_FD, HK_FULL, HK_FD):
        v = orc32.eval_hermite(kind, X, T, dT, np.zeros((3, NX, NY, NZ)), pts, DX)
        np.testing.assert_allclose(v, pts, atol=3e-7)


# ---------------------------------------------------------------- RK4 marches
A_LIN = np.array([[0.2, -0.8, 0.3], [0.6, -0.1, -0.4], [-0.3, 0.5, -0.1]])


def test_rk4_linear_flow_known_answer(orc64):
    u = linear_velocity(A_LIN, np.zeros(3), NX, NY, NZ, DX)
    pts = interior_points(NX, NY, NZ, DX, 100, margin=3.5).T.copy()
    dt = 0.07
    P = taylor_P(A_LIN * dt)
    out = orc64.advance_map_0(pts, u, DX, dt)
    np.testing.assert_allclose(out, P @ pts, atol=1e-12)
    T0 = np.broadcast_to(np.eye(3)[:, :, None], (3, 3, pts.shape[1])).copy()
    psi, T = orc64.advance_map_1_transpose(pts, T0, u, DX, -dt)
    Pm = taylor_P(-A_LIN * dt)
    np.testing.assert_allclose(psi, Pm @ pts, atol=1e-12)
    np.testing.assert_allclose(np.moveaxis(T, 2, 0), np.broadcast_to(Pm, (pts.shape[1], 3, 3)), atol=1e-12)
    phi, F = orc64.advance_map_col(pts[:, None, None, :].reshape(3, 1, 1, -1),
                                   np.tile(np.array([0., 1., 0.])[:, None, None, None], (1, 1, 1, pts.shape[1])),
                                   u, DX, dt)
    np.testing.assert_allclose(phi.reshape(3, -1), P @ pts, atol=1e-12)
    np.testing.assert_allclose(F.reshape(3, -1), np.broadcast_to((P @ np.array([0., 1., 0.]))[:, None],
                                                                 (3, pts.shape[1])), atol=1e-12)


def test_edge_march_linear_flow_from_identity(orc64):
    n = (16, 14, 12)
    dx = 1.0 / 14
    u = linear_velocity(0.3 * A_LIN, np.zeros(3), *n, dx)
    X = cell_coords(*n, dx)
    T = np.zeros((3, 3) + n)
    for c in range(3):
        T[c, c] = 1
    dt = 0.05
    psi, T1, dT1 = orc64.march_edge(X, T, np.zeros_like(T), u, X, dx, dt, 0.025 * dx)
    Pm = taylor_P(-0.3 * A_LIN * dt)
    s = (slice(4, -4),) * 3
    np.testing.assert_allclose(psi[(slice(None),) + s], np.einsum("ab,b...->a...", Pm, X)[(slice(None),) + s],
                               atol=1e-12)
    np.testing.assert_allclose(T1[(slice(None), slice(None)) + s],
                               np.broadcast_to(Pm[:, :, None, None, None], T1[(slice(None), slice(None)) + s].shape),
                               atol=1e-10)
    assert np.max(np.abs(dT1[(slice(None), slice(None)) + s])) < 1e-6


def test_edge_march_zero_flow_quadratic_map(orc64):
    # u = 0: each point stays put, so the new map is the old one sampled at the
    # cell centre (value, rows) and dT is the eps-difference of the four eps
    # points' rows (sim/hermite.py:1049-1075) -- exact for a quadratic map.
    n = (12, 12, 12)
    dx = 1.0 / 12
    E = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (0, 1, 1), (1, 0, 1)]
    C = np.random.default_rng(11).normal(size=(3, len(E))) * 0.2
    C[:, :3] += np.eye(3)
    pm = PolyMap(E, C)
    psi, T, dT, _ = pm.tables(*n, dx)
    u = [np.zeros((n[0] + 1, n[1], n[2])), np.zeros((n[0], n[1] + 1, n[2])), np.zeros((n[0], n[1], n[2] + 1))]
    X = cell_coords(*n, dx)
    p1, T1, dT1 = orc64.march_edge(psi, T, dT, u, X, dx, 0.01, 0.02 * dx)
    s = (slice(2, -2),) * 3
    S2 = (slice(None), slice(None)) + s
    np.testing.assert_allclose(p1[(slice(None),) + s], psi[(slice(None),) + s], atol=1e-12)
    np.testing.assert_allclose(T1[S2], T[S2], atol=1e-10)
    np.testing.assert_allclose(dT1[S2], dT[S2], atol=1e-6)


def test_ed8_assembly_multilinear_map(orc64):
    # calc_ed8_3 (sim/hermite.py:1098-1151) is exact for per-axis-linear maps
    n = (6, 5, 4)
    dx = 0.1
    eps = 0.01
    E = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1)]
    pm = PolyMap(E, np.random.default_rng(12).normal(size=(3, len(E))))
    X = cell_coords(*n, dx)
    P8 = []
    for e in range(8):
        s = np.array([1 if e & 1 else -1, 1 if e & 2 else -1, 1 if e & 4 else -1])
        P8.append(pm.value(X + eps * s.reshape(3, 1, 1, 1)))
    psi, T, dT, d2T = orc64.calc_ed8(np.stack(P8), eps)
    np.testing.assert_allclose(psi, pm.value(X), atol=1e-10)
    np.testing.assert_allclose(T, pm.grad_rows(X), atol=1e-8)
    np.testing.assert_allclose(dT, pm.cross(X), atol=1e-6)
    np.testing.assert_allclose(d2T, pm.xyz(X), atol=1e-4)
    # ED4 (sim/hermite.py:1154-1177) is exact for affine maps
    pa = PolyMap(E[:4], np.random.default_rng(13).normal(size=(3, 4)))
    S4 = [(-1, -1, -1), (1, 1, -1), (-1, 1, 1), (1, -1, 1)]
    P4 = np.stack([pa.value(X + eps * np.array(s).reshape(3, 1, 1, 1)) for s in S4])
    psi4, T4 = orc64.calc_ed4_1(P4, eps)
    np.testing.assert_allclose(psi4, pa.value(X), atol=1e-10)
    np.testing.assert_allclose(T4, pa.grad_rows(X), atol=1e-8)


# ---------------------------------------------------------------- driver kernels
def test_map_u_psi_identity_is_bspline_of_u(orc64):
    n = (10, 8, 8)
    dx = 0.125
    u = smooth_velocity(*n, dx, seed=3)
    X = cell_coords(*n, dx)
    T = np.zeros((3, 3) + n)
    for c in range(3):
        T[c, c] = 1
    out = orc64.map_u_psi(u, X, T, np.zeros_like(T), dx)
    from helpers import face_coords
    # away from the walls, where the Hermite clamp (sim/hermite.py:367-369) is inactive
    s = (slice(2, -2),) * 3
    for a in range(3):
        Xf = face_coords(a, *n, dx)[(slice(None),) + s].reshape(3, -1).T
        ref = orc64.eval_umac(u, Xf, dx)[:, a]
        np.testing.assert_allclose(out[a][s].reshape(-1), ref, atol=1e-12)


def test_map_u_psi_affine_map_pulls_back_covector(orc64):
    # psi(x) = M x + c (exact in every Hermite variant): u1_a(X) = sum_c M[c,a] u0_c(psi(X))
    n = (12, 12, 12)
    dx = 1.0 / 12
    M = np.eye(3) + 0.05 * np.random.default_rng(14).normal(size=(3, 3))
    cvec = np.array([0.01, -0.02, 0.015])
    pm = PolyMap([(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, 0)], np.concatenate([M, cvec[:, None]], 1))
    psi, T, dT, _ = pm.tables(*n, dx)
    A = 0.4 * A_LIN
    u = linear_velocity(A, np.array([0.1, 0.2, -0.1]), *n, dx)
    out = orc64.map_u_psi(u, psi, T, dT, dx)
    from helpers import face_coords
    for a in range(3):
        Xf = face_coords(a, *n, dx)
        P = np.einsum("cb,b...->c...", M, Xf) + cvec.reshape(3, 1, 1, 1)
        u0 = np.einsum("cb,b...->c...", A, P) + np.array([0.1, 0.2, -0.1]).reshape(3, 1, 1, 1)
        ref = np.einsum("c,c...->...", M[:, a], u0)
        s = (slice(3, -3),) * 3
        np.testing.assert_allclose(out[a][s], ref[s], atol=1e-10)


def test_semi_lagrangian_constant_and_linear(orc64):
    n = (10, 10, 10)
    dx = 0.1
    c = np.array([0.3, -0.2, 0.1])
    u = linear_velocity(np.zeros((3, 3)), c, *n, dx)
    out = orc64.advect_u_sl(u, dx, 0.05)
    for a in range(3):
        np.testing.assert_allclose(out[a], c[a], atol=1e-12)
    A = 0.5 * A_LIN
    u = linear_velocity(A, np.zeros(3), *n, dx)
    dt = 0.04
    out = orc64.advect_u_sl(u, dx, dt)
    from helpers import face_coords
    Pm = taylor_P(-A * dt)
    for a in range(3):
        Xf = face_coords(a, *n, dx)
        foot = np.einsum("ab,b...->a...", Pm, Xf)
        ref = np.einsum("b,b...->...", A[a], foot)
        s = (slice(3, -3),) * 3
        np.testing.assert_allclose(out[a][s], ref[s], atol=1e-12)


def test_clamp_u_and_max_speed(orc64):
    rng = np.random.default_rng(15)
    u = rng.normal(size=(7, 6, 5))
    u1 = rng.normal(size=(7, 6, 5))
    out = orc64.clamp_u(u, u1)
    i, j, k = 3, 2, 2
    nb = [u1[i - 1, j, k], u1[i + 1, j, k], u1[i, j - 1, k], u1[i, j + 1, k], u1[i, j, k - 1], u1[i, j, k + 1]]
    assert out[i, j, k] == pytest.approx(min(max(u[i, j, k], min(nb)), max(nb)))
    uu = [rng.normal(size=(7, 6, 5)), rng.normal(size=(6, 7, 5)), rng.normal(size=(6, 6, 6))]
    ms = orc64.max_speed(uu)
    cu = 0.5 * (uu[0][1:] + uu[0][:-1])
    cv = 0.5 * (uu[1][:, 1:] + uu[1][:, :-1])
    cw = 0.5 * (uu[2][:, :, 1:] + uu[2][:, :, :-1])
    assert ms == pytest.approx(max(1e-3, float(np.sqrt(cu ** 2 + cv ** 2 + cw ** 2).max())))
    zero = [np.zeros_like(x) for x in uu]
    assert orc64.max_speed(zero) == pytest.approx(1e-3)


def test_vortex_ring_biot_savart_on_the_axis(orc64):
    """sim/init_conditions.py:13-36 against the closed form of a circular vortex
    ring (circulation w, radius R) on its axis: u = w R^2 / (2 (R^2 + z^2)^(3/2))
    along ex x ey.  Far from the core the regularisation 1 - exp(-(r/delta)^3)
    is 1 to machine precision, and the 2000-sample trapezoid rule on the circle
    is exact to rounding for this periodic integrand; the smoke there stays 0,
    while a point on the core line is dyed with the ring's colour.  (A point
    exactly on a sample gives 0/0, in the reference as here.)"""
    w, R, delta = 1.3, 0.2, 0.03
    c = np.array([0.5, 0.4, 0.6])
    ex, ey = np.array([0.0, 0.0, -1.0]), np.array([0.0, 1.0, 0.0])
    n_hat = np.cross(ex, ey)
    z = np.linspace(-0.3, 0.3, 13)
    pts = c[:, None] + n_hat[:, None] * z[None, :]
    th = np.pi / 2000                      # on the core line, between samples 0 and 1
    core = c + R * (np.cos(th) * ex + np.sin(th) * ey)
    X = np.concatenate([pts, core[:, None]], axis=1).reshape(3, -1, 1, 1)
    vf = np.zeros(X.shape, np.float64)
    smoke = np.zeros((4,) + X.shape[1:], np.float64)
    orc64.vortex_ring(w, R, delta, c, ex, ey, X, vf, smoke, (0.25, 0.5, 1.0), 2000)
    v = vf.reshape(3, -1)
    want = w * R ** 2 / (2 * (R ** 2 + z ** 2) ** 1.5)
    assert np.max(np.abs(v[:, :13] - n_hat[:, None] * want[None, :])) < 1e-12 * want.max()
    s = smoke.reshape(4, -1)
    assert np.all(s[:, :13] == 0)
    assert np.allclose(s[:, 13], [0.25, 0.5, 1.0, 1.0])
    assert np.all(np.isfinite(v[:, 13]))


def _plate():
    ang = np.radians(15)
    x, y = np.cos(ang) * 1.0, np.sin(ang) * 1.0
    return np.array([[0, 0, 0], [x, y, 0], [x, -y, 0]], np.float64)


@pytest.mark.parametrize("res", [32, 128])
def test_voxelizer_restatement_is_pinned_by_geometry(res):
    """The restated trimesh voxelizer (the reference's delta plate): every voxel
    whose cube lies inside the triangle is marked, and every marked voxel's
    cube touches the triangle (a half-voxel shift or a floor instead of a round
    breaks one of the two); mirror symmetric in y like the triangle; and no
    kept vertex lies within 1e-9 of a rounding boundary, so the bit-level order
    of the midpoint sums cannot change the result."""
    from oracle import voxel_ref as V
    dx = 1.0 / res
    tri = _plate()
    m, lo = V.voxelize_subdivide(tri, np.array([[0, 1, 2]]), dx)
    assert m.shape[2] == 1
    m2 = m[:, :, 0]
    assert np.array_equal(m2, m2[:, ::-1])
    ii, jj = np.meshgrid(np.arange(m2.shape[0]) + lo[0], np.arange(m2.shape[1]) + lo[1], indexing="ij")
    a, b = np.radians(15), 1.0
    x1 = np.cos(a) * b

    def inside(px, py):   # the closed triangle (0,0), (x1, +-y1)
        return (px >= 0) & (px <= x1) & (np.abs(py) <= px * np.tan(a))
    corners = [(sx, sy) for sx in (-0.5, 0.5) for sy in (-0.5, 0.5)]
    full = np.ones_like(m2)
    for sx, sy in corners:
        full &= inside((ii + sx) * dx, (jj + sy) * dx)
    assert np.all(m2[full])
    # a marked voxel's square meets the triangle: a point of a 41 x 41 sample of it is inside
    s = (np.arange(41) / 40.0) - 0.5
    touch = np.zeros_like(m2)
    for sx in s:
        for sy in s:
            touch |= inside((ii + sx) * dx, (jj + sy) * dx)
    assert not np.any(m2 & ~touch)
    v = V.subdivide_to_size(tri, np.array([[0, 1, 2]]), dx / 2)
    assert np.min(np.abs(np.abs(v[:, :2] / dx - np.round(v[:, :2] / dx)) - 0.5)) > 1e-9


@pytest.mark.parametrize("res", [16, 48, 128, 256])
def test_delta_plate_voxelizer_equals_the_restatement(res):
    """sim/init_conditions.voxelize_triangle (coordinate triangles) against the
    restatement of trimesh's indexed subdivision: the same voxels."""
    from oracle import voxel_ref as V
    from sim.init_conditions import voxelize_triangle
    dx = 1.0 / res
    a, lo_a = voxelize_triangle(_plate(), dx)
    b, lo_b = V.voxelize_subdivide(_plate(), np.array([[0, 1, 2]]), dx)
    assert np.array_equal(a, b) and np.array_equal(lo_a, lo_b)


@pytest.mark.parametrize("res", [32, 128])
def test_plate_mask_is_the_restated_voxelizer_placed(res):
    """The 256x128x128 solid workload's plate (span 0.3, half-angle 15 degrees):
    its mask equals the restated trimesh voxelizer's matrix of the same
    triangle, placed at the apex cell and extruded four cells in z; its y
    extent is the span to within a cell and it is mirror symmetric in y."""
    from oracle import voxel_ref as V
    from sim.init_conditions import PLATE_APEX, plate_mask, plate_vertices
    dx = 1.0 / res
    shape = (2 * res, res, res)
    m = plate_mask(shape, dx)
    obj, _ = V.voxelize_subdivide(plate_vertices(), np.array([[0, 1, 2]]), dx)
    ii, jj, kk = np.nonzero(m)
    i0, j0, k0 = ii.min(), jj.min(), kk.min()
    assert i0 == int(PLATE_APEX[0] // dx) and set(np.unique(kk)) == set(range(k0, k0 + 4))
    assert k0 + 2 == round(PLATE_APEX[2] / dx)
    for t in range(4):
        sl = m[i0:i0 + obj.shape[0], j0:j0 + obj.shape[1], k0 + t]
        assert np.array_equal(sl.astype(bool), obj[:, :, 0])
    assert m.sum() == 4 * obj.sum()
    assert abs((jj.max() - jj.min() + 1) * dx - 0.3) <= 1.5 * dx
    jc = (jj.max() + jj.min()) / 2
    assert abs(jc + 0.5 - 0.5 / dx) <= 0.5
