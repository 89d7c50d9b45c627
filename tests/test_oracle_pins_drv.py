"""Pin the oracle's step-loop driver kernels and its map composition with
closed forms: identity and affine maps (which every Hermite variant and the
B-spline reproduce exactly away from the clamps), an independent NumPy
restatement of the trilinear min/max clamp, and the whole BFECC correction
through affine maps.  Each test names the mistake it is built to catch.  CPU
only."""
import numpy as np
import pytest

from helpers import (PolyMap, cell_coords, face_coords, interior_points, linear_velocity, smooth_velocity,
                     taylor_P)
from oracle.oracle import HK_FD, HK_FULL, HK_GARM_FD
from oracle.step import OracleStep

# non-symmetric, non-commuting with the flow matrices below
M_AFF = np.array([[1.04, 0.07, -0.03], [-0.05, 0.97, 0.06], [0.08, -0.02, 1.02]])
C_AFF = np.array([0.0131371, -0.0213417, 0.0172931])
N_AFF = np.array([[0.98, -0.04, 0.05], [0.06, 1.03, -0.02], [-0.03, 0.05, 0.99]])
D_AFF = np.array([-0.0112713, 0.0163291, -0.0091237])
A_FLOW = np.array([[0.2, -0.8, 0.3], [0.6, -0.1, -0.4], [-0.3, 0.5, -0.1]])
KINDS = (HK_GARM_FD, HK_FULL, HK_FD)


def affine_map(M, c, n, dx):
    pm = PolyMap([(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, 0)], np.concatenate([M, c[:, None]], 1))
    return pm.tables(*n, dx)


def affine_face_map(M, c, n, dx):
    """phi_a = M X_a + c on each face grid, F_a = column a of M (run_edge.py:306-320 layout)."""
    phi, F = [], []
    for a in range(3):
        X = face_coords(a, *n, dx)
        phi.append(np.einsum("cb,b...->c...", M, X) + c.reshape(3, 1, 1, 1))
        F.append(np.ascontiguousarray(np.broadcast_to(M[:, a].reshape(3, 1, 1, 1), X.shape)))
    return phi, F


def identity_face_map(n, dx):
    return affine_face_map(np.eye(3), np.zeros(3), n, dx)


def affine_eval(A, b, P):
    """A P + b for points P of shape (3, ...)."""
    return np.einsum("ab,b...->a...", A, P) + b.reshape((3,) + (1,) * (P.ndim - 1))


def trilinear_min_max_np(s0, pc, dx):
    """util/taichi_utils.py:200-227 sample_min_max_1, restated with NumPy index
    arrays: u = p/dx - 0.5 clamped to [1, dim - 1 - 1e-6], floor, the 2x2x2 taps
    clipped into the grid (sample(), :61-67)."""
    nc = s0.shape[0]
    dims = s0.shape[1:]
    base = []
    for d in range(3):
        u = np.clip(pc[d] / dx - 0.5, 1.0, dims[d] - 1 - 1e-6)
        base.append(np.floor(u).astype(int))
    mn = np.full((nc,) + pc.shape[1:], np.inf)
    mx = np.full((nc,) + pc.shape[1:], -np.inf)
    for o in np.ndindex(2, 2, 2):
        idx = [np.clip(base[d] + o[d], 0, dims[d] - 1) for d in range(3)]
        v = s0[:, idx[0], idx[1], idx[2]]
        mn = np.minimum(mn, v)
        mx = np.maximum(mx, v)
    return mn, mx


# ---------------------------------------------------------------- composition
def test_compose_multiplies_old_rows_on_the_left(orc64):
    """sim/hermite.py:976 (and :1016): T = T_old @ T_new.  Through an affine old
    map psi(x) = M x + c a point p with rows Tp composes to (M p + c, M Tp).  M
    and Tp do not commute, so the swapped product Tp @ M fails, as does a
    transposed M (rows read as columns)."""
    n, dx = (10, 9, 8), 0.1
    psi, T, dT, d2T = affine_map(M_AFF, C_AFF, n, dx)
    pts = interior_points(*n, dx, 60, margin=2.5).T.copy()
    Tp = np.random.default_rng(1).normal(size=(3, 3, pts.shape[1]))
    for kind in KINDS:
        p2, T2 = orc64.compose(kind, psi, T, dT, d2T, pts, Tp, dx)
        np.testing.assert_allclose(p2, affine_eval(M_AFF, C_AFF, pts), atol=1e-12)
        np.testing.assert_allclose(T2, np.einsum("ac,cbq->abq", M_AFF, Tp), atol=1e-12)
        assert np.max(np.abs(T2 - np.einsum("acq,cb->abq", Tp, M_AFF))) > 1e-2


def test_edge_march_composes_linear_flow_through_affine_map(orc64):
    """The fused EDGE march (sim/hermite.py:1426-1492) from an affine old map
    psi = M x + c in the linear flow u = A x: the RK4 step is P = P(-A dt)
    (4th-order Taylor), so psi_new = M P x + c and T_new = M P exactly, dT = 0.
    A march that composed as P M (the swapped product of :976) fails."""
    n = (16, 14, 12)
    dx = 1.0 / 14
    A = 0.3 * A_FLOW
    u = linear_velocity(A, np.zeros(3), *n, dx)
    X = cell_coords(*n, dx)
    psi, T, dT, _ = affine_map(M_AFF, C_AFF, n, dx)
    dt = 0.05
    p1, T1, dT1 = orc64.march_edge(psi, T, dT, u, X, dx, dt, 0.025 * dx)
    P = taylor_P(-A * dt)
    MP = M_AFF @ P
    s = (slice(4, -4),) * 3
    S2 = (slice(None), slice(None)) + s
    np.testing.assert_allclose(p1[(slice(None),) + s], affine_eval(MP, C_AFF, X)[(slice(None),) + s], atol=1e-12)
    np.testing.assert_allclose(T1[S2], np.broadcast_to(MP[:, :, None, None, None], T1[S2].shape), atol=1e-10)
    assert np.max(np.abs(dT1[S2])) < 1e-6
    assert np.max(np.abs(P @ M_AFF - MP)) > 1e-3


# ---------------------------------------------------------------- map_u
def test_map_u_identity_face_map_is_bspline_of_u(orc64):
    """run_edge.py:360-373 with phi_a = X_a and F_a = e_a: u1_a = interp_u_MAC(u0, X_a)_a
    on every face, walls included (the B-spline's own clamp decides there).  A
    map that dotted F with the wrong face grid's velocity fails."""
    n, dx = (10, 9, 8), 0.1
    u = smooth_velocity(*n, dx, seed=5)
    phi, F = identity_face_map(n, dx)
    out = orc64.map_u(u, F, phi, dx)
    for a in range(3):
        ref = orc64.eval_umac(u, phi[a].reshape(3, -1).T, dx)[:, a]
        np.testing.assert_allclose(out[a].reshape(-1), ref, atol=1e-13)


def test_map_u_affine_face_map_pulls_back_covector(orc64):
    """Face map phi(x) = N x + d with columns F_a = N[:, a] and an affine u0 = A x + b:
    u1_a(X) = sum_c N[c, a] u0_c(N X + d) exactly away from the walls.  Using the
    row N[a, :] instead of the column (a transposed operand) fails: N is not
    symmetric."""
    n, dx = (14, 12, 12), 1.0 / 12
    A, b = 0.4 * A_FLOW, np.array([0.1, 0.2, -0.1])
    u = linear_velocity(A, b, *n, dx)
    phi, F = affine_face_map(N_AFF, D_AFF, n, dx)
    out = orc64.map_u(u, F, phi, dx)
    s = (slice(3, -3),) * 3
    for a in range(3):
        u0 = affine_eval(A, b, phi[a])
        ref = np.einsum("c,c...->...", N_AFF[:, a], u0)
        wrong = np.einsum("c,c...->...", N_AFF[a, :], u0)
        np.testing.assert_allclose(out[a][s], ref[s], atol=1e-12)
        assert np.max(np.abs(out[a][s] - wrong[s])) > 1e-3


# ---------------------------------------------------------------- smoke maps
def _affine_smoke(n, dx, X=None):
    G = np.array([[0.5, -1.0, 0.25], [1.5, 0.3, -0.7], [-0.2, 0.8, 0.4], [0.0, 0.0, 0.0]])
    h = np.array([0.3, -0.1, 0.2, 1.0])
    X = cell_coords(*n, dx) if X is None else X
    return G, h, np.einsum("cb,b...->c...", G, X) + h.reshape(4, 1, 1, 1)


def test_map_smoke_identity_and_affine(orc64):
    """run_edge.py:420-427: psi_c = the mean of the six face values of phi, then
    the trilinear sample.  Identity faces: s1 = s0 exactly where the trilinear
    clamp [1, dim-1-eps] is inactive (a node is reproduced).  Affine faces and an
    affine smoke: s1 = s0(N X_c + d), since the six faces average to the cell
    centre.  A mean over the wrong faces (i-1 instead of i+1) or a missing 1/6
    fails."""
    n, dx = (12, 10, 9), 0.1
    s0 = np.random.default_rng(2).normal(size=(4,) + n)
    phi, _ = identity_face_map(n, dx)
    out = orc64.map_smoke(0, s0, None, phi, dx)
    s = (slice(None),) + (slice(1, -1),) * 3
    np.testing.assert_allclose(out[s], s0[s], atol=1e-12)
    G, h, sa = _affine_smoke(n, dx)
    phi, _ = affine_face_map(N_AFF, D_AFF, n, dx)
    out = orc64.map_smoke(0, sa, None, phi, dx)
    ref = np.einsum("cb,b...->c...", G, affine_eval(N_AFF, D_AFF, cell_coords(*n, dx))) + h.reshape(4, 1, 1, 1)
    s = (slice(None),) + (slice(2, -2),) * 3
    np.testing.assert_allclose(out[s], ref[s], atol=1e-12)


def test_map_smoke_psi_identity_and_affine(orc64):
    """run_edge.py:443-466: six garm-Hermite evaluations of the cell map at the
    cell's faces (split_to_mac, :430-440), averaged, then trilinear.  Identity map:
    s1 = s0 where the Hermite clamp [1, dim-2-eps] is inactive at all six faces.
    Affine map psi = M x + c with an affine smoke: s1 = s0(M X_c + c).  A face
    taken at X_x[i-1] instead of X_x[i+1], or the mean over five faces, fails."""
    n, dx = (12, 10, 9), 0.1
    s0 = np.random.default_rng(3).normal(size=(4,) + n)
    psi, T, dT, _ = affine_map(np.eye(3), np.zeros(3), n, dx)
    out = orc64.map_smoke_psi(0, s0, None, psi, T, dT, dx)
    s = (slice(None),) + (slice(2, -2),) * 3
    np.testing.assert_allclose(out[s], s0[s], atol=1e-12)
    G, h, sa = _affine_smoke(n, dx)
    psi, T, dT, _ = affine_map(M_AFF, C_AFF, n, dx)
    out = orc64.map_smoke_psi(0, sa, None, psi, T, dT, dx)
    ref = np.einsum("cb,b...->c...", G, affine_eval(M_AFF, C_AFF, cell_coords(*n, dx))) + h.reshape(4, 1, 1, 1)
    np.testing.assert_allclose(out[s], ref[s], atol=1e-12)


def test_clamp_smoke_against_numpy_min_max(orc64):
    """run_edge.py:469-477: clamp s1 to the min/max of the 8 trilinear taps of s0
    around the mean face position.  Faces displaced at random by up to 2.5 cells
    (some beyond the walls, so the position and index clamps act) are checked
    against an independent NumPy restatement; the output lies within the taps'
    range and equals s1 wherever s1 already does.  A tap window shifted by one
    cell, or a clamp of the position to [0, dim-1], fails the comparison."""
    n, dx = (11, 9, 8), 0.1
    rng = np.random.default_rng(4)
    s0 = rng.normal(size=(4,) + n)
    s1 = 1.3 * rng.normal(size=(4,) + n)
    phi, _ = identity_face_map(n, dx)
    phi = [p + rng.uniform(-2.5, 2.5, size=p.shape) * dx for p in phi]
    out = orc64.map_smoke(1, s0, s1, phi, dx)
    pc = (phi[0][:, :-1] + phi[0][:, 1:] + phi[1][:, :, :-1] + phi[1][:, :, 1:] + phi[2][..., :-1] +
          phi[2][..., 1:]) / 6.0
    mn, mx = trilinear_min_max_np(s0, pc, dx)
    np.testing.assert_array_equal(out, np.minimum(np.maximum(s1, mn), mx))
    assert np.all(out >= mn) and np.all(out <= mx)
    inside = (s1 >= mn) & (s1 <= mx)
    assert 0.1 < inside.mean() < 0.9
    np.testing.assert_array_equal(out[inside], s1[inside])


def test_clamp_smoke_psi_against_numpy_min_max(orc64):
    """run_edge.py:480-505: the same clamp at the mean of the six garm-Hermite
    face evaluations.  Through an affine map that mean is M X_c + c exactly, so
    the taps follow from the closed form; the affine map's shift moves some
    cells' tap window by one cell from the identity's, which a window taken at
    X_c fails."""
    n, dx = (12, 10, 9), 0.1
    rng = np.random.default_rng(5)
    s0 = rng.normal(size=(4,) + n)
    s1 = 1.3 * rng.normal(size=(4,) + n)
    M = np.eye(3) + 0.3 * (M_AFF - np.eye(3))
    c = np.array([0.6, -0.4, 0.3]) * dx
    psi, T, dT, _ = affine_map(M, c, n, dx)
    out = orc64.map_smoke_psi(1, s0, s1, psi, T, dT, dx)
    pc = affine_eval(M, c, cell_coords(*n, dx))
    mn, mx = trilinear_min_max_np(s0, pc, dx)
    s = (slice(None),) + (slice(2, -2),) * 3
    ref = np.minimum(np.maximum(s1, mn), mx)
    np.testing.assert_allclose(out[s], ref[s], atol=0)
    mn0, mx0 = trilinear_min_max_np(s0, cell_coords(*n, dx), dx)
    assert np.any(mn0[s] != mn[s])
    assert np.all(out[s] >= mn[s]) and np.all(out[s] <= mx[s])


# ---------------------------------------------------------------- split_psi_hermite
@pytest.mark.parametrize("kind", KINDS)
def test_split_psi_hermite_returns_map_and_column_at_faces(orc64, kind):
    """run_ed8.py:190-223: phi_a = psi(X_a) and F_a = d psi / d x_a (column a,
    transposed from the Hermite rows) at every face of grid a.  With exact data
    of a polynomial the variant reproduces (full: tricubic; garm-FD and FD:
    per-axis quadratic, their finite-difference cross terms are exact there) the
    result is the polynomial's value and column away from the clamp.  Returning
    the row (d psi_a / d x) instead of the column fails: the polynomial's
    Jacobian is not symmetric."""
    n, dx = (10, 9, 8), 0.1
    rng = np.random.default_rng(6 + kind)
    if kind == HK_FULL:
        E = [(a, b, c) for a in range(4) for b in range(4) for c in range(4) if a + b + c <= 5]
    else:
        E = [(a, b, c) for a in range(3) for b in range(3) for c in range(3)]
    C = rng.normal(size=(3, len(E))) * 0.3
    C[:, [E.index((1, 0, 0)), E.index((0, 1, 0)), E.index((0, 0, 1))]] += np.eye(3)
    pm = PolyMap(E, C)
    psi, T, dT, d2T = pm.tables(*n, dx)
    phi, F = orc64.split_psi_hermite(kind, psi, T, dT, d2T, dx)
    s = (slice(None),) + (slice(2, -2),) * 3
    for a in range(3):
        Xf = face_coords(a, *n, dx)
        G = pm.grad_rows(Xf)          # G[c, b] = d psi_c / d x_b
        np.testing.assert_allclose(phi[a][s], pm.value(Xf)[s], atol=1e-9)
        np.testing.assert_allclose(F[a][s], G[:, a][s], atol=1e-7)
        assert np.max(np.abs(F[a][s] - G[a][s])) > 1e-2


# ---------------------------------------------------------------- BFECC
def _affine_bfecc_state(clamp):
    n = (22, 20, 18)
    dx = 1.0 / 20
    A0 = 0.5 * A_FLOW
    b0 = np.array([0.8, -0.6, 0.7])
    u0 = linear_velocity(A0, b0, *n, dx)
    G, h, s0 = _affine_smoke(n, dx)
    st = OracleStep(u0, s0, dx, 12, 0.5, 0.025 * dx, 2, clamp, dtype=np.float64)
    st.psi, st.T, st.dT, st.d2T = affine_map(M_AFF, C_AFF, n, dx)
    st.phi, st.F = affine_face_map(N_AFF, D_AFF, n, dx)
    st.init_u = [c.copy() for c in u0]
    st.init_smoke = s0.copy()
    return st, n, dx, (A0, b0), (G, h, s0)


def _bfecc_closed_form(A0, b0, G, h):
    """Every pull-back and push-forward of an affine field through an affine map is
    affine: pull u (x) = M^T u(M x + c), push u (x) = N^T u(N x + d), and the same
    for a scalar without the transpose."""
    def pull_v(A, b):
        return M_AFF.T @ A @ M_AFF, M_AFF.T @ (A @ C_AFF + b)

    def push_v(A, b):
        return N_AFF.T @ A @ N_AFF, N_AFF.T @ (A @ D_AFF + b)
    A1, b1 = pull_v(A0, b0)
    A2, b2 = push_v(A1, b1)
    Ae, be = (A2 - A0) / 2, (b2 - b0) / 2
    Ap, bp = pull_v(Ae, be)
    vel = (A1, b1, A1 - Ap, b1 - bp)

    def pull_s(g, c0):
        return g @ M_AFF, g @ C_AFF + c0

    def push_s(g, c0):
        return g @ N_AFF, g @ D_AFF + c0
    g1, c1 = pull_s(G, h)
    g2, c2 = push_s(g1, c1)
    ge, ce = (g2 - G) / 2, (c2 - h) / 2
    gp, cp = pull_s(ge, ce)
    return vel, (g1 - gp, c1 - cp)


def test_bfecc_through_affine_maps_closed_form(orc64):
    """run_edge.py:669-700 without BFECC_clamp: u = pull(u0) - pull((push(pull(u0))
    - u0) / 2), each pull/push of an affine field through an affine map being
    affine (the B-spline and the garm Hermite reproduce them), so the corrected
    velocity and smoke are known in closed form away from the walls.  The maps
    are not inverses of each other, so the error term is not zero: a wrong sign,
    a missing 1/2 or a push through the backward map fails.  Identity maps give
    u = u0 (the SURVEY's 'BFECC through identity maps returns affine u')."""
    st, n, dx, (A0, b0), (G, h, s0) = _affine_bfecc_state(False)
    st.bfecc()
    (A1, b1, Au, bu), (gs, cs) = _bfecc_closed_form(A0, b0, G, h)
    s = (slice(7, -7),) * 3
    for a in range(3):
        ref = affine_eval(Au, bu, face_coords(a, *n, dx))[a]
        np.testing.assert_allclose(st.u[a][s], ref[s], atol=1e-11)
        plain = affine_eval(A1, b1, face_coords(a, *n, dx))[a]
        assert np.max(np.abs(plain[s] - ref[s])) > 1e-3
    Xc = cell_coords(*n, dx)
    sm_ref = np.einsum("cb,b...->c...", gs, Xc) + cs.reshape(4, 1, 1, 1)
    pc = affine_eval(M_AFF, C_AFF, Xc)
    mn, mx = trilinear_min_max_np(s0, pc, dx)
    sm_ref = np.minimum(np.maximum(sm_ref, mn), mx)
    # a cell whose clamp position lies within rounding of a grid line may take
    # either tap window; those few are left out
    u = pc / dx - 0.5
    clear = np.all(np.abs(u - np.round(u)) > 1e-9, axis=0)[s]
    assert clear.mean() > 0.99
    got = st.smoke[(slice(None),) + s]
    np.testing.assert_allclose(got[:, clear], sm_ref[(slice(None),) + s][:, clear], atol=1e-11)
    # identity maps: affine u and smoke come back unchanged
    st.psi, st.T, st.dT, _ = affine_map(np.eye(3), np.zeros(3), n, dx)
    st.phi, st.F = identity_face_map(n, dx)
    st.bfecc()
    for a in range(3):
        np.testing.assert_allclose(st.u[a][s], st.init_u[a][s], atol=1e-12)
    np.testing.assert_allclose(st.smoke[(slice(None),) + s], s0[(slice(None),) + s], atol=1e-12)


def test_bfecc_clamp_limits_to_uncorrected_neighbours(orc64):
    """run_edge.py:694-700 with BFECC_clamp: the corrected velocity is clamped to
    the min/max of the six neighbours of the uncorrected pull-back pull(u0)
    (clamp_u, :508-519), which for an affine field is its value plus or minus
    the largest one-cell difference along an axis."""
    st, n, dx, (A0, b0), (G, h, s0) = _affine_bfecc_state(True)
    st.bfecc()
    (A1, b1, Au, bu), _ = _bfecc_closed_form(A0, b0, G, h)
    s = (slice(7, -7),) * 3
    clamped_any = False
    for a in range(3):
        X = face_coords(a, *n, dx)
        corr = affine_eval(Au, bu, X)[a]
        u1 = affine_eval(A1, b1, X)[a]
        nb = [u1 + A1[a, d] * dx * sg for d in range(3) for sg in (-1, 1)]
        ref = np.minimum(np.maximum(corr, np.min(nb, 0)), np.max(nb, 0))
        np.testing.assert_allclose(st.u[a][s], ref[s], atol=1e-11)
        clamped_any |= bool(np.any(np.abs(ref[s] - corr[s]) > 1e-6))
    assert clamped_any
