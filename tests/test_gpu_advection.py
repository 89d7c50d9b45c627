"""GPU parity of the advection kernels against the CPU oracle (FP32, and FP64
where a quantity is ill-conditioned).  Tolerances are written per test:
  * 1e-5 of the field's scale for FP32 quantities computed in a different
    summation order (the reference's 64-tap loops vs the 27-tap kernels);
  * for eps-differences (dT), the GPU's error against the FP64 oracle must be
    within a small factor of the FP32 oracle's own error against FP64.
"""
import numpy as np
import pytest
import torch

from helpers import (any_points, cell_coords, identity_plus, interior_points, linear_velocity, smooth_velocity)
from oracle.oracle import HK_FD, HK_FULL, HK_GARM_FD

pytestmark = pytest.mark.gpu

N3 = (24, 20, 16)
DX = 1.0 / 20


def T_(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def N_(t):
    return t.detach().cpu().numpy().astype(np.float64)


def cl(a):
    """(4, nx, ny, nz) channel planes (the oracle's smoke) -> channels per cell (the kernels')."""
    return np.moveaxis(a, 0, -1)


def cf(a):
    return np.moveaxis(a, -1, 0)


def scale_err(gpu, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(np.asarray(gpu, np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-30))


def realistic_map(orc, n, dx, steps=2, seed=0, amp=1.0):
    """A map produced by the oracle's own EDGE march from identity."""
    u = smooth_velocity(*n, dx, seed=seed, amp=amp)
    X = cell_coords(*n, dx)
    psi = X.copy()
    T = np.zeros((3, 3) + n)
    for c in range(3):
        T[c, c] = 1
    dT = np.zeros_like(T)
    for _ in range(steps):
        psi, T, dT = orc.march_edge(psi, T, dT, u, X, dx, 0.5 * dx / amp, 0.025 * dx)
    return u, X, psi, T, dT


def test_bspline_points(orc32, dev):
    from edge_hip import ops
    rng = np.random.default_rng(0)
    f = rng.normal(size=(N3[0] + 1, N3[1], N3[2]))
    pts = any_points(*N3, DX, 5000, seed=1)
    for BL in [(0, .5, .5), (.5, .5, .5)]:
        v, g = ops.eval_bspline(T_(f, dev), BL, T_(pts, dev), DX, grad=True)
        rv, rg = orc32.eval_bspline(f, BL, pts, DX, grad=True)
        assert scale_err(N_(v), rv) < 1e-5
        assert scale_err(N_(g), rg) < 1e-5


def test_umac_points(orc32, dev):
    from edge_hip import ops
    u = smooth_velocity(*N3, DX, seed=2)
    pts = any_points(*N3, DX, 5000, seed=3)
    v, J = ops.eval_umac([T_(c, dev) for c in u], T_(pts, dev), DX, grad=True)
    rv, rJ = orc32.eval_umac(u, pts, DX, grad=True)
    assert scale_err(N_(v), rv) < 1e-5
    assert scale_err(N_(J), rJ) < 1e-5


@pytest.mark.parametrize("kind", ["garm", "full", "fd"])
def test_hermite_table_points(orc32, dev, kind):
    from edge_hip import ops
    pm = identity_plus(*N3, DX, amp=0.3, seed=4)
    psi, T, dT, d2T = pm.tables(*N3, DX)
    rng = np.random.default_rng(5)
    dT = dT + 0.1 * rng.normal(size=dT.shape)  # arbitrary derivative data
    d2T = d2T + 0.1 * rng.normal(size=d2T.shape)
    pts = any_points(*N3, DX, 5000, seed=6)
    tab = ops.pack_table(kind, T_(psi, dev), T_(T, dev), T_(dT, dev), T_(d2T, dev), DX)
    v, G = ops.eval_map(tab, N3, T_(pts, dev), DX, grad=True)
    hk = {"garm": HK_GARM_FD, "full": HK_FULL, "fd": HK_FD}[kind]
    rv, rG = orc32.eval_hermite(hk, psi, T, dT, d2T, pts, DX, grad=True)
    assert scale_err(N_(v), rv) < 2e-6
    assert scale_err(N_(G), rG) < 1e-5


def test_point_marches(orc32, dev):
    from edge_hip import ops
    u = smooth_velocity(*N3, DX, seed=7)
    pts = any_points(*N3, DX, 4000, seed=8).T.copy()
    gu = [T_(c, dev) for c in u]
    p0 = T_(pts, dev)
    ops.advance_points(0, gu, p0, None, DX, -0.03)
    ref = orc32.advance_map_0(pts, u, DX, -0.03)
    assert scale_err(N_(p0), ref) < 1e-6
    Tm = np.broadcast_to(np.eye(3)[:, :, None], (3, 3, pts.shape[1])).copy()
    p1, t1 = T_(pts, dev), T_(Tm, dev)
    ops.advance_points(1, gu, p1, t1, DX, -0.03)
    rp, rT = orc32.advance_map_1_transpose(pts, Tm, u, DX, -0.03)
    assert scale_err(N_(p1), rp) < 1e-6
    assert np.max(np.abs(N_(t1) - rT)) < 1e-5


def test_face_march(orc32, dev):
    from edge_hip import ops
    from helpers import face_coords
    u = smooth_velocity(*N3, DX, seed=9)
    gu = [T_(c, dev) for c in u]
    phi = [face_coords(a, *N3, DX) for a in range(3)]
    F = [np.broadcast_to(np.eye(3)[a][:, None, None, None], phi[a].shape).copy() for a in range(3)]
    gphi, gF = [T_(p, dev) for p in phi], [T_(f, dev) for f in F]
    ops.advance_faces(gu, gphi, gF, DX, 0.02)
    for a in range(3):
        rphi, rF = orc32.advance_map_col(phi[a], F[a], u, DX, 0.02)
        assert scale_err(N_(gphi[a]), rphi) < 1e-6
        assert np.max(np.abs(N_(gF[a]) - rF)) < 1e-5


def test_edge_march_parity(orc32, orc64, dev):
    from edge_hip import ops
    n, dx = (32, 24, 20), 1.0 / 24
    u, X, psi, T, dT = realistic_map(orc64, n, dx, steps=2, seed=10)
    dt, eps = 0.5 * dx, 0.025 * dx
    r32 = orc32.march_edge(psi, T, dT, u, X, dx, dt, eps)
    r64 = orc64.march_edge(psi, T, dT, u, X, dx, dt, eps)
    tab = ops.pack_table("garm", T_(psi, dev), T_(T, dev), T_(dT, dev), None, dx)
    gp, gT, gdT = torch.empty((3,) + n, device=dev), torch.empty((3, 3) + n, device=dev), \
        torch.empty((3, 3) + n, device=dev)
    ops.march_edge([T_(c, dev) for c in u], tab, T_(X, dev), gp, gT, gdT, dx, dt, eps)
    assert scale_err(N_(gp), r32[0]) < 1e-6
    # T: the FP32 oracle is itself ~2e-5 from FP64 here (dx = 1/24: p / dx and the
    # kernels' p * (1/dx) round differently); GPU against FP64 with that yardstick
    eT_orc = np.max(np.abs(r32[1] - r64[1]))
    eT_gpu = np.max(np.abs(N_(gT) - r64[1]))
    print(f"march T: GPU vs FP64 {eT_gpu:.3g}, FP32 oracle vs FP64 {eT_orc:.3g}")
    assert eT_gpu <= max(2e-5, 3 * eT_orc), (eT_gpu, eT_orc)
    # dT is an eps-difference: compare against FP64 with the FP32 oracle's own error as the yardstick
    e_orc = np.max(np.abs(r32[2] - r64[2]))
    e_gpu = np.max(np.abs(N_(gdT) - r64[2]))
    assert e_gpu <= 4 * e_orc + 1e-3 * np.max(np.abs(r64[2])), (e_gpu, e_orc)


def test_edge_march_writes_the_packed_table(orc64, dev):
    """The march's new-table output (plus the d2T fill) is bit for bit the table
    edge_pack_hermite_table builds from the map the march stores, and the stored
    map is bit for bit the plain march's (32x24x20: a partial last block)."""
    from edge_hip import ops
    n, dx = (32, 24, 20), 1.0 / 24
    u, X, psi, T, dT = realistic_map(orc64, n, dx, steps=2, seed=11)
    dt, eps = 0.5 * dx, 0.025 * dx
    tab = ops.pack_table("garm", T_(psi, dev), T_(T, dev), T_(dT, dev), None, dx)
    ug, Xg = [T_(c, dev) for c in u], T_(X, dev)
    out = []
    for with_table in (False, True):
        gp, gT, gdT = torch.empty((3,) + n, device=dev), torch.empty((3, 3) + n, device=dev), \
            torch.empty((3, 3) + n, device=dev)
        if with_table:
            tab_new = ops.march_edge_table(ug, tab, Xg, gp, gT, gdT, dx, dt, eps)
        else:
            ops.march_edge(ug, tab, Xg, gp, gT, gdT, dx, dt, eps)
        out.append((gp, gT, gdT))
    for a, b in zip(*out):
        assert torch.equal(a, b)
    packed = ops.pack_table("garm", *out[1], None, dx)
    assert torch.equal(tab_new, packed)
    # the old table is left as it was (the march reads it while writing the new one)
    assert torch.equal(tab, ops.pack_table("garm", T_(psi, dev), T_(T, dev), T_(dT, dev), None, dx))


def test_driver_maps_parity(orc32, orc64, dev):
    from edge_hip import ops
    n, dx = (32, 24, 20), 1.0 / 24
    u, X, psi, T, dT = realistic_map(orc64, n, dx, steps=2, seed=11)
    gu = [T_(c, dev) for c in u]
    tab = ops.pack_table("garm", T_(psi, dev), T_(T, dev), T_(dT, dev), None, dx)
    # map_u_psi
    out = ops.map_u_psi(gu, tab, dx)
    ref = orc32.map_u_psi(u, psi, T, dT, dx)
    ref64 = orc64.map_u_psi(u, psi, T, dT, dx)
    for a in range(3):
        # the FP32 oracle itself is ~1e-5 of scale from FP64 here (sum of
        # d psi_c/dx_a u_c with cancellation): compare both against FP64
        e32 = scale_err(ref[a], ref64[a])
        assert scale_err(N_(out[a]), ref64[a]) <= max(2.0 * e32, 1e-5), (a, e32)
        assert scale_err(N_(out[a]), ref[a]) < 5e-5
    # smoke maps and clamps through the map
    rng = np.random.default_rng(12)
    s0 = rng.uniform(0, 1, size=(4,) + n)
    s1 = rng.uniform(-0.5, 1.5, size=(4,) + n)
    for mode in (0, 1):
        g = ops.map_smoke_psi(mode, T_(cl(s0), dev), T_(cl(s1), dev) if mode else None, tab, dx)
        r = orc32.map_smoke_psi(mode, s0, s1 if mode else None, psi, T, dT, dx)
        assert scale_err(cf(N_(g)), r) < 1e-5
    # face maps phi/F from an oracle face march, then map_u / map_smoke / clamp_smoke
    from helpers import face_coords
    phi = [face_coords(a, *n, dx) for a in range(3)]
    F = [np.broadcast_to(np.eye(3)[a][:, None, None, None], phi[a].shape).copy() for a in range(3)]
    for a in range(3):
        phi[a], F[a] = orc64.advance_map_col(phi[a], F[a], u, dx, 0.3 * dx)
    gphi, gF = [T_(p, dev) for p in phi], [T_(f, dev) for f in F]
    out = ops.map_u(gu, gF, gphi, dx)
    ref = orc32.map_u(u, F, phi, dx)
    for a in range(3):
        assert scale_err(N_(out[a]), ref[a]) < 1e-5
    for mode in (0, 1):
        g = ops.map_smoke(mode, T_(cl(s0), dev), T_(cl(s1), dev) if mode else None, gphi, dx)
        r = orc32.map_smoke(mode, s0, s1 if mode else None, phi, dx)
        assert scale_err(cf(N_(g)), r) < 1e-5
    # the BFECC arithmetic folded into the maps: the bits of the separate passes
    from edge_hip import _lib
    from edge_hip.field import stream_handle
    _p = ops._p
    gi = [T_(c, dev) for c in smooth_velocity(*n, dx, seed=16)]
    fused = [torch.empty_like(c) for c in gi]
    _lib.call("edge_map_u_add_scale", *[_p(c) for c in gu], *[_p(c) for c in fused], *[_p(c) for c in gF],
              *[_p(c) for c in gphi], *[_p(c) for c in gi], -1., 0.5, *n, dx, stream_handle())
    for a in range(3):
        seq = torch.empty_like(out[a])
        _lib.call("edge_axpy", _p(out[a]), _p(gi[a]), _p(seq), -1., seq.numel(), stream_handle())
        _lib.call("edge_scale", _p(seq), 0.5, _p(seq), seq.numel(), stream_handle())
        assert torch.equal(fused[a], seq), a
    gs0, gs1, gs2 = T_(cl(s0), dev), T_(cl(s1), dev), T_(rng.uniform(-0.5, 1.5, size=n + (4,)), dev)
    for mode in (0, 1):
        sm = gs1.clone()
        _lib.call("edge_map_smoke_add_scale", mode, _p(gs0), _p(sm), _p(gs2), -1., 0.5, 4,
                  *[_p(c) for c in gphi], *n, dx, stream_handle())
        if mode == 0:
            seq = ops.map_smoke(0, gs0, None, gphi, dx)
            _lib.call("edge_axpy", _p(seq), _p(gs2), _p(seq), -1., seq.numel(), stream_handle())
            _lib.call("edge_scale", _p(seq), 0.5, _p(seq), seq.numel(), stream_handle())
        else:
            seq = gs1.clone()
            _lib.call("edge_axpy", _p(seq), _p(gs2), _p(seq), -1., seq.numel(), stream_handle())
            seq = ops.map_smoke(1, gs0, seq, gphi, dx)
        assert torch.equal(sm, seq), mode
    # clamp_u, semi-Lagrangian, max speed
    u1 = smooth_velocity(*n, dx, seed=13)
    gc = ops.clamp_u(gu, [T_(c, dev) for c in u1])
    for a in range(3):
        assert np.array_equal(N_(gc[a]).astype(np.float32), orc32.clamp_u(u[a], u1[a]))
    # the fused BFECC tail: t = clamp(u1 - t) in place, the same bits as add_fields + clamp_u
    gu1 = [T_(c, dev) for c in u1]
    gt = [T_(c, dev) for c in u]
    err = [(a_ - b_) for a_, b_ in zip(gu1, gt)]
    two = ops.clamp_u(err, gu1)
    _lib.call("edge_sub_clamp_u", *[_p(c) for c in gu1], *[_p(c) for c in gt], *n, stream_handle())
    torch.cuda.synchronize()
    for a in range(3):
        assert torch.equal(gt[a], two[a])
        u1f, uf = u1[a].astype(np.float32), u[a].astype(np.float32)
        ref = orc32.clamp_u(u1f + np.float32(-1.) * uf, u1f)
        assert np.array_equal(N_(gt[a]).astype(np.float32), ref)
    # the error's pull-back folded in too: clamp(u1 - map_u(err)) in one pass, the
    # bits of map_u into t and then the fused tail above
    gt = ops.map_u(gi, gF, gphi, dx)
    _lib.call("edge_sub_clamp_u", *[_p(c) for c in gu1], *[_p(c) for c in gt], *n, stream_handle())
    one = [torch.empty_like(c) for c in gu1]
    _lib.call("edge_map_u_sub_clamp", *[_p(c) for c in gi], *[_p(c) for c in one], *[_p(c) for c in gF],
              *[_p(c) for c in gphi], *[_p(c) for c in gu1], *n, dx, stream_handle())
    torch.cuda.synchronize()
    for a in range(3):
        assert torch.equal(one[a], gt[a]), a
    # and for the smoke: clamp(s1 - map_smoke(s2)) with s0's range (mode 2), the
    # bits of map_smoke into tmp then mode 1 with tmp
    tmp = ops.map_smoke(0, gs2, None, gphi, dx)
    seq = gs1.clone()
    _lib.call("edge_map_smoke_add_scale", 1, _p(gs0), _p(seq), _p(tmp), -1., 1., 4, *[_p(c) for c in gphi], *n, dx,
              stream_handle())
    sm = gs1.clone()
    _lib.call("edge_map_smoke_add_scale", 2, _p(gs0), _p(sm), _p(gs2), -1., 1., 4, *[_p(c) for c in gphi], *n, dx,
              stream_handle())
    torch.cuda.synchronize()
    assert torch.equal(sm, seq)
    sl = ops.advect_u_sl(gu, dx, 0.25 * dx)
    rsl = orc32.advect_u_sl(u, dx, 0.25 * dx)
    for a in range(3):
        assert scale_err(N_(sl[a]), rsl[a]) < 1e-5
    assert ops.max_speed(gu) == pytest.approx(orc32.max_speed(u), rel=1e-6)


@pytest.mark.parametrize("disp_cells", [0.25, 0.6, 3.0])
def test_advect_sl_displacements(orc32, dev, disp_cells):
    """k_advect_sl reads the face's own velocity component from 27 taps held in
    registers (and the other two components' RK4 stage-2 taps for stages 3
    and 4), gathering again wherever a stage point's taps differ.  The largest
    displacement sweeps from a quarter cell (the held taps serve every stage),
    through 0.6 cells (some stages leave them) to 3 cells (nearly every stage
    leaves them, and the walls clamp): every case within 1e-5 of scale of the
    FP32 oracle's semi-Lagrangian step (sim/hermite.py:1578-1621)."""
    from edge_hip import ops
    n, dx = (32, 24, 20), 1.0 / 24
    u = smooth_velocity(*n, dx, seed=21)
    gu = [T_(c, dev) for c in u]
    umax = max(float(np.abs(c).max()) for c in u)
    dt = disp_cells * dx / umax
    sl = ops.advect_u_sl(gu, dx, dt)
    rsl = orc32.advect_u_sl(u, dx, dt)
    for a in range(3):
        e = scale_err(N_(sl[a]), rsl[a])
        print(f"displacement {disp_cells} cells, u_{'xyz'[a]}: {e:.2e}")
        assert e < 1e-5


def test_prefetched_max_speed(orc32, dev):
    """prefetch_max_speed (launched at the end of a step) gives calc_max_speed's
    value; a field written since (every wrapper bumps its version) is measured again."""
    from edge_hip import Field
    from sim.edge_driver import calc_max_speed, prefetch_max_speed
    from util.taichi_utils import scale_field
    u = smooth_velocity(*N3, DX, seed=11)
    f = []
    for c in u:
        g = Field(None, float, c.shape)
        g.from_numpy(c.astype(np.float32))
        f.append(g)
    direct = calc_max_speed(*f)
    assert direct == pytest.approx(orc32.max_speed(u), rel=1e-6)
    prefetch_max_speed(*f)
    assert calc_max_speed(*f) == direct
    prefetch_max_speed(*f)
    scale_field(f[0], 3.0, f[0])
    u3 = [u[0] * 3.0, u[1], u[2]]
    assert calc_max_speed(*f) == pytest.approx(orc32.max_speed(u3), rel=1e-6)
    assert calc_max_speed(*f) != direct


def test_split_psi_and_variant_marches(orc32, orc64, dev):
    """ED8 (sim/hermite.py:1290-1338), GE (1398-1424), ED4 (1375-1395) and
    split_psi_hermite (run_ed8.py:190-223) against the oracle's unfused sequence."""
    from edge_hip import ops, _lib
    from edge_hip.field import stream_handle
    n, dx = (24, 20, 16), 1.0 / 20
    u = smooth_velocity(*n, dx, seed=14)
    pm = identity_plus(*n, dx, amp=0.2, seed=15)
    psi, T, dT, d2T = pm.tables(*n, dx)
    X = cell_coords(*n, dx)
    dt, eps = 0.4 * dx, 0.025 * dx
    gu = [T_(c, dev) for c in u]
    s = stream_handle()
    # --- ED8 fused vs oracle: reset, 8x advance_map_0, 8x interp_psi_hermite, calc_ed8_3
    P8 = []
    for e in range(8):
        sg = np.array([1 if e & 1 else -1, 1 if e & 2 else -1, 1 if e & 4 else -1], float)
        p = (X + eps * sg.reshape(3, 1, 1, 1)).reshape(3, -1)
        p = orc64.advance_map_0(p, u, dx, -dt)
        p, _ = orc64.compose(HK_FULL, psi, T, dT, d2T, p, None, dx)
        P8.append(p.reshape((3,) + n))
    r = orc64.calc_ed8(np.stack(P8), eps)
    tab = ops.pack_table("full", T_(psi, dev), T_(T, dev), T_(dT, dev), T_(d2T, dev), dx)
    out = [torch.empty((3,) + n, device=dev), torch.empty((3, 3) + n, device=dev),
           torch.empty((3, 3) + n, device=dev), torch.empty((3,) + n, device=dev)]
    _lib.call("edge_march_ed8", *[c.data_ptr() for c in gu], tab.data_ptr(), T_(X, dev).data_ptr(),
              out[0].data_ptr(), *[out[1][c].data_ptr() for c in range(3)], *[out[2][c].data_ptr() for c in range(3)],
              *[out[3][c].data_ptr() for c in range(3)], *n, float(dx), float(dt), float(eps), s)
    assert scale_err(N_(out[0]), r[0]) < 1e-6
    assert np.max(np.abs(N_(out[1]) - r[1])) < 5e-3  # /eps of FP32 positions
    # --- GE: march + compose (full) then FD derivatives
    p = X.reshape(3, -1).copy()
    Tm = np.broadcast_to(np.eye(3)[:, :, None], (3, 3, p.shape[1])).copy()
    p, Tm = orc64.advance_map_1_transpose(p, Tm, u, dx, -dt)
    p, Tm = orc64.compose(HK_FULL, psi, T, dT, d2T, p, Tm, dx)
    rp, rT = p.reshape((3,) + n), Tm.reshape((3, 3) + n)
    rdT, rd2T = orc64.fd3(rT, None, dx, want_dT=True)
    gp, gT = torch.empty((3,) + n, device=dev), torch.empty((3, 3) + n, device=dev)
    _lib.call("edge_march_ge", *[c.data_ptr() for c in gu], tab.data_ptr(), T_(X, dev).data_ptr(), gp.data_ptr(),
              *[gT[c].data_ptr() for c in range(3)], *n, float(dx), float(dt), s)
    gdT, gd2T = ops.calc_fd3(gT, dx)
    assert scale_err(N_(gp), rp) < 1e-6
    assert np.max(np.abs(N_(gT) - rT)) < 2e-5
    assert scale_err(N_(gdT), rdT) < 1e-3
    # --- ED4 (fd variant)
    S4 = [(-1, -1, -1), (1, 1, -1), (-1, 1, 1), (1, -1, 1)]
    P4 = []
    for sg in S4:
        q = (X + eps * np.array(sg, float).reshape(3, 1, 1, 1)).reshape(3, -1)
        q = orc64.advance_map_0(q, u, dx, -dt)
        q, _ = orc64.compose(HK_FD, psi, T, None, None, q, None, dx)
        P4.append(q.reshape((3,) + n))
    r4 = orc64.calc_ed4_1(np.stack(P4), eps)
    tab4 = ops.pack_table("fd", T_(psi, dev), T_(T, dev), None, None, dx)
    gp4, gT4 = torch.empty((3,) + n, device=dev), torch.empty((3, 3) + n, device=dev)
    _lib.call("edge_march_ed4", *[c.data_ptr() for c in gu], tab4.data_ptr(), T_(X, dev).data_ptr(),
              gp4.data_ptr(), *[gT4[c].data_ptr() for c in range(3)], *n, float(dx), float(dt), float(eps), s)
    assert scale_err(N_(gp4), r4[0]) < 1e-6
    assert np.max(np.abs(N_(gT4) - r4[1])) < 5e-3
    # --- split_psi_hermite (full)
    phi, F = ops.split_psi_hermite(tab, n, dx, dev)
    rphi, rF = orc32.split_psi_hermite(HK_FULL, psi, T, dT, d2T, dx)
    for a in range(3):
        assert scale_err(N_(phi[a]), rphi[a]) < 1e-6
        assert np.max(np.abs(N_(F[a]) - rF[a])) < 2e-5


def test_vortex_ring_init(orc32, dev):
    from edge_hip import ops
    n, dx = (32, 16, 16), 1.0 / 16
    X = cell_coords(*n, dx)
    vf, sm = np.zeros((3,) + n, np.float32), np.zeros((4,) + n, np.float32)
    args = (0.02, 0.2, 0.03, (0.6, 0.5, 0.5), (0, 0, -1), (0, 1, 0))
    orc32.vortex_ring(*args, X, vf, sm, (1, 0, 0), 500)
    gv, gs = torch.zeros((3,) + n, device=dev), torch.zeros(n + (4,), device=dev)   # smoke: channels per cell
    ops.vortex_ring(*args, T_(X, dev), gv, gs, (1, 0, 0), 500)
    assert scale_err(N_(gv), vf) < 1e-4
    assert np.mean(cf(N_(gs)) != sm) < 1e-3  # threshold flips only at |s - 0.002| ~ rounding


@pytest.mark.parametrize("nvel", [0, 3, 35])
def test_backtrack_faces_equals_sequential_marches(dev, nvel):
    """run_efm.py:170-185: the one-pass backtrack (faces kept in registers
    across the velocity list, 32 velocities per launch) against a reset and one
    face-march launch per velocity: the same values bit for bit."""
    from edge_hip import ti
    from sim.edge_driver import backtrack_faces, march_phi_grid, reset_to_identity
    n, dx = (20, 12, 12), 1.0 / 12
    fs = [(n[0] + 1, n[1], n[2]), (n[0], n[1] + 1, n[2]), (n[0], n[1], n[2] + 1)]
    vels = []
    for s in range(nvel):
        u = smooth_velocity(*n, dx, seed=40 + s % 5, amp=0.5)
        f = [ti.field(float, shape=sh) for sh in fs]
        for c, a in zip(f, u):
            c.from_numpy(a.astype(np.float32))
        vels.append(tuple(f))
    dts = [-(0.3 + 0.01 * s) * dx for s in range(nvel)]
    A = [ti.Vector.field(3, float, shape=sh) for sh in fs] + [ti.Vector.field(3, float, shape=sh) for sh in fs]
    B = [ti.Vector.field(3, float, shape=sh) for sh in fs] + [ti.Vector.field(3, float, shape=sh) for sh in fs]
    backtrack_faces(*A, vels, dts, dx)
    reset_to_identity(*B, dx)
    for v, d in zip(vels, dts):
        march_phi_grid(*B, *v, dx, d)
    torch.cuda.synchronize()
    for a, b in zip(A, B):
        assert np.array_equal(a.to_numpy(), b.to_numpy())
    if nvel:
        assert not np.array_equal(A[0].to_numpy(), np.moveaxis(face_coords_np(0, n, dx), 0, -1))


def face_coords_np(a, n, dx):
    d = [n[0] + (a == 0), n[1] + (a == 1), n[2] + (a == 2)]
    i, j, k = np.meshgrid(*[np.arange(m) for m in d], indexing="ij")
    o = [0.0 if a == b else 0.5 for b in range(3)]
    return np.stack([(i + o[0]) * dx, (j + o[1]) * dx, (k + o[2]) * dx]).astype(np.float32)


def test_primitives_against_golden_vectors(dev):
    """The HIP B-spline, MAC velocity and full-Hermite evaluations against the
    brute-force golden vectors (tests/golden/make_golden.py), FP32 tolerances."""
    import os
    from edge_hip import ops
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "interp_golden.npz")
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    dx, pts = float(g["dx"]), g["pts"]
    n = g["psi"].shape[1:]
    for name, BL in (("own_x", (0.0, 0.5, 0.5)), ("cell", (0.5, 0.5, 0.5))):
        v, gr = ops.eval_bspline(T_(g["f"], dev), BL, T_(pts, dev), dx, grad=True)
        assert scale_err(N_(v), g[f"bspline_{name}_val"]) < 1e-5
        assert scale_err(N_(gr), g[f"bspline_{name}_grad"]) < 1e-5
    v, J = ops.eval_umac([T_(g[f"u{a}"], dev) for a in range(3)], T_(pts, dev), dx, grad=True)
    assert scale_err(N_(v), g["umac_val"]) < 1e-5
    assert scale_err(N_(J), g["umac_jac"]) < 1e-5
    for kind, key in (("full", "hermite"), ("garm", "hermite_garm"), ("fd", "hermite_fd")):
        tab = ops.pack_table(kind, T_(g["psi"], dev), T_(g["T"], dev), T_(g["dT"], dev), T_(g["d2T"], dev), dx)
        v, G = ops.eval_map(tab, n, T_(pts, dev), dx, grad=True)
        assert scale_err(N_(v), g[f"{key}_val"]) < 1e-5, kind
        assert scale_err(N_(G), g[f"{key}_grad"]) < 1e-5, kind


def test_add_scale_fields_equals_add_then_scale(dev):
    """run_edge.py:675-681's two passes (add_fields then scale_field, in place)
    and the fused one give the same bits, including a length that is not a
    multiple of 4."""
    from edge_hip import ti
    from util.taichi_utils import add_fields, add_scale_fields, scale_field
    rng = np.random.default_rng(17)
    for shape in [(33, 16, 16), (7, 5, 3)]:
        a, b = rng.normal(size=shape).astype(np.float32), rng.normal(size=shape).astype(np.float32)
        f1, f2, g1 = ti.field(float, shape=shape), ti.field(float, shape=shape), ti.field(float, shape=shape)
        f1.from_numpy(a)
        g1.from_numpy(a)
        f2.from_numpy(b)
        add_fields(f1, f2, f1, -1.)
        scale_field(f1, 0.5, f1)
        add_scale_fields(g1, f2, g1, -1., 0.5)
        assert np.array_equal(f1.to_numpy(), g1.to_numpy())
        assert np.array_equal(g1.to_numpy(), ((a - b) * np.float32(0.5)).astype(np.float32))
