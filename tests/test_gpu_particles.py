"""GPU parity of the particle kernels (sim/lp.py drop-in) against the FP32 oracle.

Per-particle and per-face results agree to 1e-5 of scale; the P2G scatter
sums with FP32 atomics in an order that differs from the oracle's sequential
loop, so it is held to 1e-5 of the largest drag value.
"""
import numpy as np
import pytest
import torch

from helpers import smooth_velocity

pytestmark = pytest.mark.gpu

NX, NY, NZ, DX = 32, 24, 20, 1.0 / 24
L = (NX * DX, NY * DX, NZ * DX)
NMAX, NUM = 6000, 5500


def scale_err(gpu, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(np.asarray(gpu, np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-30))


def vec_field(arr):
    from edge_hip import Field
    f = Field(3, float, arr.shape[1:])
    f.from_numpy(np.moveaxis(arr.astype(np.float32), 0, -1))
    return f


def soa(f):
    return np.moveaxis(f.to_numpy(), -1, 0).astype(np.float64)


def mac_fields(u):
    from edge_hip import Field
    out = []
    for c in u:
        f = Field(None, float, c.shape)
        f.from_numpy(c.astype(np.float32))
        out.append(f)
    return out


@pytest.fixture(scope="module")
def state():
    rng = np.random.default_rng(0)
    u = [c.astype(np.float32) for c in smooth_velocity(NX, NY, NZ, DX, seed=9)]
    pos = np.stack([rng.uniform(0, L[a], NMAX) for a in range(3)]).astype(np.float32)
    vel = (0.3 * rng.normal(size=(3, NMAX))).astype(np.float32)
    rr = rng.uniform(0.5, 1.5, NMAX).astype(np.float32)
    return u, pos, vel, rr


def test_particle_kernels_match_oracle(orc32, dev, state):
    import sim.lp as lp
    from sim.edge_driver import calc_max_particle_speed
    u, pos, vel, rr = state
    U = mac_fields(u)
    P, V = vec_field(pos), vec_field(vel)
    A, D = vec_field(np.zeros_like(vel)), vec_field(np.zeros_like(vel))
    St, g, dt, dr, er = 0.05, 1.0, 0.004, 2.0 / 3, 35.0
    lp.update_particles_acc(*U, V, P, A, DX, dt, NUM, St, g)
    acc = orc32.lp_acc(u, vel, pos, NUM, DX, dt, St, g)
    assert scale_err(soa(A), acc) < 1e-5
    lp.update_particles_velocity(A, V, P, dt, DX, NUM, *L)
    vel1 = orc32.lp_vel(acc, vel, pos, NUM, dt, DX, L)
    assert scale_err(soa(V), vel1) < 1e-5
    lp.calculate_velocity_diff(*U, V, D, P, DX, NUM, St, dr, er)
    diff = orc32.lp_vel_diff(u, vel1, pos, NUM, DX, St, dr, er)
    assert scale_err(soa(D)[:, :NUM], diff[:, :NUM]) < 1e-5

    from edge_hip import Field
    drag = [Field(None, float, c.shape) for c in u]
    R = Field(None, float, (NMAX,))
    R.from_numpy(rr)
    lp.laden_particles_back_with_computing_particle(*drag, *U, V, D, P, R, DX, NUM, 0.01, *L, NX, NY, NZ)
    want = orc32.lp_p2g(soa(D), pos, rr, NUM, DX, 0.01, L, (NX, NY, NZ))
    for a in range(3):
        got = drag[a].to_numpy()
        assert np.abs(want[a]).max() > 0
        assert scale_err(got, want[a]) < 1e-5

    lp.update_particles_position(V, P, dt, NUM, *L)
    pos1 = orc32.lp_pos(vel1, pos, NUM, dt, L)
    assert scale_err(soa(P), pos1) < 1e-6

    vis = [Field(None, float, c.shape) for c in u]
    Re = 5.5
    lp.calculate_viscous_force(*U, *vis, DX, Re)
    for a in range(3):
        assert scale_err(vis[a].to_numpy(), orc32.lp_viscous(u[a], DX, Re)) < 1e-5

    m = calc_max_particle_speed(V)
    assert m == pytest.approx(max(1e-3, float(np.sqrt((soa(V) ** 2).sum(0)).max())), rel=1e-6)


def test_p2g_skips_nan_and_wall_particles(dev):
    import sim.lp as lp
    from edge_hip import Field
    n = 4
    pos = np.full((3, n), 0.5 * L[2], np.float32)
    pos[0, 0] = np.nan          # skipped (would be a garbage index otherwise)
    pos[1, 1] = 2.0 * DX        # within 3 dx of the lower y wall: skipped
    diff = np.ones((3, n), np.float32)
    P, D = vec_field(pos), vec_field(diff)
    R = Field(None, float, (n,))
    R.from_numpy(np.ones(n, np.float32))
    drag = [Field(None, float, s) for s in [(NX + 1, NY, NZ), (NX, NY + 1, NZ), (NX, NY, NZ + 1)]]
    lp.laden_particles_back_with_computing_particle(*drag, None, None, None, None, D, P, R, DX, n, 1.0, *L,
                                                    NX, NY, NZ)
    for a in range(3):
        np.testing.assert_allclose(drag[a].to_numpy().sum(), 2.0, rtol=1e-5)  # particles 2 and 3 only


def test_p2g_dense_cells_and_grid_aligned_positions(orc32, dev):
    """Many particles per cell (a tight cluster) plus positions on cell corners,
    faces and centres (multiples of dx/2, where a tap window can start one
    cell lower than the particle's cell), against the oracle's sequential
    scatter."""
    import sim.lp as lp
    from edge_hip import Field
    rng = np.random.default_rng(12)
    c = np.array([NX / 2, NY / 2, NZ / 2]) * DX
    cluster = c[:, None] + rng.uniform(-1.5 * DX, 1.5 * DX, size=(3, 40000))
    grid = np.stack([rng.integers(6, 2 * n - 6, 4000) * (DX / 2) for n in (NX, NY, NZ)])
    pos = np.concatenate([cluster, grid], axis=1).astype(np.float32)
    n = pos.shape[1]
    diff = rng.normal(size=(3, n)).astype(np.float32)
    rr = rng.uniform(0.5, 1.5, n).astype(np.float32)
    P, D = vec_field(pos), vec_field(diff)
    R = Field(None, float, (n,))
    R.from_numpy(rr)
    drag = [Field(None, float, s) for s in [(NX + 1, NY, NZ), (NX, NY + 1, NZ), (NX, NY, NZ + 1)]]
    lp.laden_particles_back_with_computing_particle(*drag, None, None, None, None, D, P, R, DX, n, 0.01, *L,
                                                    NX, NY, NZ)
    want = orc32.lp_p2g(diff, pos, rr, n, DX, 0.01, L, (NX, NY, NZ))
    for a in range(3):
        assert scale_err(drag[a].to_numpy(), want[a]) < 1e-5


def test_gathers_in_cell_order_give_the_same_values(dev, state):
    """After a scatter the gathers visit particles in its cell order: the
    per-particle results are bitwise those of the identity order."""
    import sim.lp as lp
    from edge_hip import Field
    u, pos, vel, rr = state
    U = mac_fields(u)
    P, V = vec_field(pos), vec_field(vel)
    St, g, dt, dr, er = 0.05, 1.0, 0.004, 2.0 / 3, 35.0

    def gathers():
        A, D = vec_field(np.zeros_like(vel)), vec_field(np.zeros_like(vel))
        lp.update_particles_acc(*U, V, P, A, DX, dt, NUM, St, g)
        lp.calculate_velocity_diff(*U, V, D, P, DX, NUM, St, dr, er)
        return soa(A), soa(D)

    # a scatter of a different particle count leaves no usable order
    R = Field(None, float, (NMAX,))
    R.from_numpy(rr)
    drag = [Field(None, float, c.shape) for c in u]
    lp.laden_particles_back_with_computing_particle(*drag, *U, V, vec_field(np.zeros_like(vel)), P, R, DX, NUM - 7,
                                                    0.01, *L, NX, NY, NZ)
    a0, d0 = gathers()
    # a scatter of the same particles: its cell order is used
    lp.laden_particles_back_with_computing_particle(*drag, *U, V, vec_field(np.zeros_like(vel)), P, R, DX, NUM, 0.01,
                                                    *L, NX, NY, NZ)
    a1, d1 = gathers()
    assert np.array_equal(a0, a1) and np.array_equal(d0, d1)
