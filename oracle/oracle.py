"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY.  This module restates nothing itself: it marshals
numpy arrays into the C++ restatement of the reference algorithm
(edge_oracle*.cpp, which cites the reference line by line).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may import it; the
product path (edge-flow-maps_amd/) never does.

Array conventions (numpy, C order, k fastest), shared with the HIP side:
  scalar cell field        (nx, ny, nz)
  vec3 cell field          (3, nx, ny, nz)          plane c = component c
  map gradient T           (3, 3, nx, ny, nz)       T[c] = grad psi_c (rows)
  map cross derivatives dT (3, 3, nx, ny, nz)       dT[c] = [d2/dydz, d2/dxdz, d2/dxdy] psi_c
  map xyz derivative d2T   (3, nx, ny, nz)
  MAC velocity             (u_x (nx+1,ny,nz), u_y (nx,ny+1,nz), u_z (nx,ny,nz+1))
  vec3 face field          tuple of three (3, n0, n1, n2) arrays
  point sets               (n, 3) for the eval_* hooks, (3, n) planes for marches
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

HK_GARM_FD, HK_FULL, HK_FD = 0, 1, 2


def build(force=False):
    """Compile the oracle with the committed Makefile (g++ only, no GPU)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return LIB_PATH


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        _LIB = ctypes.CDLL(LIB_PATH)
        for name in ("orc_max_speed_f",):
            getattr(_LIB, name).restype = ctypes.c_float
        _LIB.orc_max_speed_d.restype = ctypes.c_double
        _LIB.orc_mgpcg_solve_f.restype = ctypes.c_int
        _LIB.orc_mgpcg_solve_d.restype = ctypes.c_int
    return _LIB


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """Oracle at one precision: np.float32 (the reference's ti.f32) or np.float64."""

    def __init__(self, dtype=np.float32):
        self.dt = np.dtype(dtype)
        self.s = "f" if self.dt == np.float32 else "d"
        self.R = ctypes.c_float if self.s == "f" else ctypes.c_double
        self.L = lib()

    # -- helpers ------------------------------------------------------------
    def a(self, x):
        return np.ascontiguousarray(x, dtype=self.dt)

    def fn(self, name):
        return getattr(self.L, f"orc_{name}_{self.s}")

    def r(self, x):
        return self.R(x)

    @staticmethod
    def cell_dims(u):
        ux = u[0]
        return ux.shape[0] - 1, ux.shape[1], ux.shape[2]

    def mac(self, u):
        return [self.a(c) for c in u]

    # -- point evaluations ---------------------------------------------------
    def eval_bspline(self, f, BL, pts, dx, grad=False):
        f = self.a(f)
        pts = self.a(pts).reshape(-1, 3)
        n = pts.shape[0]
        val = np.zeros(n, self.dt)
        g = np.zeros((n, 3), self.dt) if grad else None
        self.fn("eval_bspline")(_p(f), *f.shape, _p(self.a(BL)), _p(pts), ctypes.c_long(n), self.r(dx),
                                _p(val), _p(g))
        return (val, g) if grad else val

    def eval_umac(self, u, pts, dx, grad=False):
        u = self.mac(u)
        nx, ny, nz = self.cell_dims(u)
        pts = self.a(pts).reshape(-1, 3)
        n = pts.shape[0]
        val = np.zeros((n, 3), self.dt)
        J = np.zeros((n, 3, 3), self.dt) if grad else None
        self.fn("eval_umac")(_p(u[0]), _p(u[1]), _p(u[2]), nx, ny, nz, _p(pts), ctypes.c_long(n),
                             self.r(dx), _p(val), _p(J))
        return (val, J) if grad else val

    def _map_args(self, psi, T, dT, d2T):
        psi = self.a(psi)
        T = self.a(T) if T is not None else None
        dT = self.a(dT) if dT is not None else None
        d2T = self.a(d2T) if d2T is not None else None
        return psi, T, dT, d2T

    def eval_hermite(self, kind, psi, T, dT, d2T, pts, dx, grad=False):
        psi, T, dT, d2T = self._map_args(psi, T, dT, d2T)
        nx, ny, nz = psi.shape[1:]
        pts = self.a(pts).reshape(-1, 3)
        n = pts.shape[0]
        val = np.zeros((n, 3), self.dt)
        G = np.zeros((n, 3, 3), self.dt) if grad else None
        self.fn("eval_hermite")(kind, _p(psi), _p(T), _p(dT), _p(d2T), nx, ny, nz, _p(pts),
                                ctypes.c_long(n), self.r(dx), _p(val), _p(G))
        return (val, G) if grad else val

    def eval_linear(self, f, pts, dx):
        """interp_1 and sample_min_max_1 of every component of f (nc, n0, n1, n2)."""
        f = self.a(f)
        if f.ndim == 3:
            f = f[None]
        nc = f.shape[0]
        pts = self.a(pts).reshape(-1, 3)
        n = pts.shape[0]
        val = np.zeros((n, nc), self.dt)
        mn = np.zeros((n, nc), self.dt)
        mx = np.zeros((n, nc), self.dt)
        self.fn("eval_linear")(_p(f), nc, *f.shape[1:], _p(pts), ctypes.c_long(n), self.r(dx), _p(val),
                               _p(mn), _p(mx))
        return val, mn, mx

    # -- marches ---------------------------------------------------------------
    def advance_map_0(self, pts, u, dx, dt):
        """sim/hermite.py:786-803 on points (3, n); returns the marched copy."""
        pts = self.a(pts).copy()
        u = self.mac(u)
        n = pts.shape[-1] if pts.ndim == 2 else int(np.prod(pts.shape[1:]))
        self.fn("advance_map_0")(_p(pts), ctypes.c_long(n), _p(u[0]), _p(u[1]), _p(u[2]),
                                 *self.cell_dims(u), self.r(dx), self.r(dt))
        return pts

    def advance_map_1_transpose(self, psi, T, u, dx, dt):
        psi = self.a(psi).copy()
        T = self.a(T).copy()
        u = self.mac(u)
        n = int(np.prod(psi.shape[1:]))
        self.fn("advance_map_1_transpose")(_p(psi), _p(T), ctypes.c_long(n), _p(u[0]), _p(u[1]), _p(u[2]),
                                           *self.cell_dims(u), self.r(dx), self.r(dt))
        return psi, T

    def advance_map_col(self, phi, F, u, dx, dt):
        """run_edge.py:336-357 advance_map on one face grid (phi, F: (3, n0, n1, n2))."""
        phi = self.a(phi).copy()
        F = self.a(F).copy()
        u = self.mac(u)
        n = int(np.prod(phi.shape[1:]))
        self.fn("advance_map_col")(_p(phi), _p(F), ctypes.c_long(n), _p(u[0]), _p(u[1]), _p(u[2]),
                                   *self.cell_dims(u), self.r(dx), self.r(dt))
        return phi, F

    def march_edge(self, psi, T, dT, u, X, dx, dt, eps):
        """Fused EDGE march (sim/hermite.py:1426-1521): returns the new (psi, T, dT)."""
        psi, T, dT, _ = self._map_args(psi, T, dT, None)
        u = self.mac(u)
        X = self.a(X)
        nx, ny, nz = psi.shape[1:]
        po, To, dTo = np.zeros_like(psi), np.zeros_like(T), np.zeros_like(dT)
        self.fn("march_edge")(_p(psi), _p(T), _p(dT), _p(po), _p(To), _p(dTo), _p(u[0]), _p(u[1]), _p(u[2]),
                              _p(X), nx, ny, nz, self.r(dx), self.r(dt), self.r(eps))
        return po, To, dTo

    def compose(self, kind, psi, T, dT, d2T, pts, Tp, dx):
        """Compose points pts (3, n...) with rows Tp (3, 3, n...) through the map."""
        psi, T, dT, d2T = self._map_args(psi, T, dT, d2T)
        nx, ny, nz = psi.shape[1:]
        pts = self.a(pts).copy()
        Tp = self.a(Tp).copy() if Tp is not None else None
        n = int(np.prod(pts.shape[1:]))
        self.fn("compose")(kind, _p(psi), _p(T), _p(dT), _p(d2T), nx, ny, nz, _p(pts), _p(Tp),
                           ctypes.c_long(n), self.r(dx))
        return pts, Tp

    def calc_ed8(self, P8, eps):
        P8 = self.a(P8)  # (8, 3, nx, ny, nz)
        shp = P8.shape[2:]
        N = int(np.prod(shp))
        psi = np.zeros((3,) + shp, self.dt)
        T = np.zeros((3, 3) + shp, self.dt)
        dT = np.zeros((3, 3) + shp, self.dt)
        d2T = np.zeros((3,) + shp, self.dt)
        self.fn("calc_ed8")(_p(P8), ctypes.c_long(N), self.r(eps), _p(psi), _p(T), _p(dT), _p(d2T))
        return psi, T, dT, d2T

    def calc_ed4_1(self, P4, eps):
        P4 = self.a(P4)
        shp = P4.shape[2:]
        N = int(np.prod(shp))
        psi = np.zeros((3,) + shp, self.dt)
        T = np.zeros((3, 3) + shp, self.dt)
        self.fn("calc_ed4_1")(_p(P4), ctypes.c_long(N), self.r(eps), _p(psi), _p(T))
        return psi, T

    def fd3(self, T, dT, dx, want_dT):
        """sim/hermite.py:1180-1287; returns (dT, d2T)."""
        T = self.a(T) if T is not None else None
        dT = self.a(dT).copy() if dT is not None else np.zeros_like(T)
        shp = dT.shape[2:]
        d2T = np.zeros((3,) + shp, self.dt)
        self.fn("fd3")(_p(T), _p(dT), _p(d2T), *shp, self.r(dx), int(want_dT))
        return dT, d2T

    # -- driver kernels ----------------------------------------------------------
    def map_u(self, u0, F, phi, dx):
        u0 = self.mac(u0)
        F = [self.a(f) for f in F]
        phi = [self.a(f) for f in phi]
        out = [np.zeros_like(c) for c in u0]
        self.fn("map_u")(*[_p(c) for c in u0], *[_p(c) for c in out], *[_p(c) for c in F],
                         *[_p(c) for c in phi], *self.cell_dims(u0), self.r(dx))
        return out

    def map_u_psi(self, u0, psi, T, dT, dx):
        u0 = self.mac(u0)
        psi, T, dT, _ = self._map_args(psi, T, dT, None)
        out = [np.zeros_like(c) for c in u0]
        self.fn("map_u_psi")(*[_p(c) for c in u0], *[_p(c) for c in out], _p(psi), _p(T), _p(dT),
                             *self.cell_dims(u0), self.r(dx))
        return out

    def map_smoke(self, mode, s0, s1, phi, dx):
        s0 = self.a(s0)
        s1 = self.a(s1).copy() if s1 is not None else np.zeros_like(s0)
        phi = [self.a(f) for f in phi]
        nc, nx, ny, nz = s0.shape
        self.fn("map_smoke")(mode, _p(s0), _p(s1), nc, *[_p(c) for c in phi], nx, ny, nz, self.r(dx))
        return s1

    def map_smoke_psi(self, mode, s0, s1, psi, T, dT, dx):
        s0 = self.a(s0)
        s1 = self.a(s1).copy() if s1 is not None else np.zeros_like(s0)
        psi, T, dT, _ = self._map_args(psi, T, dT, None)
        nc, nx, ny, nz = s0.shape
        self.fn("map_smoke_psi")(mode, _p(s0), _p(s1), nc, _p(psi), _p(T), _p(dT), nx, ny, nz, self.r(dx))
        return s1

    def clamp_u(self, u, u1):
        u, u1 = self.a(u), self.a(u1)
        u2 = np.zeros_like(u)
        self.fn("clamp_u")(_p(u), _p(u1), _p(u2), *u.shape)
        return u2

    def advect_u_sl(self, u0, dx, dt):
        u0 = self.mac(u0)
        out = [np.zeros_like(c) for c in u0]
        self.fn("advect_u_sl")(*[_p(c) for c in u0], *[_p(c) for c in out], *self.cell_dims(u0),
                               self.r(dx), self.r(dt))
        return out

    def central_vector(self, u):
        """util/taichi_utils.py:95-100: (3, nx, ny, nz) face means of the MAC velocity."""
        u = self.mac(u)
        nx, ny, nz = self.cell_dims(u)
        vc = np.zeros((3, nx, ny, nz), self.dt)
        self.fn("central_vector")(*[_p(c) for c in u], _p(vc), nx, ny, nz)
        return vc

    def curl(self, vf, dx):
        """util/taichi_utils.py:70-92: curl of a (3, nx, ny, nz) cell-centred field."""
        vf = self.a(vf)
        cf = np.zeros_like(vf)
        self.fn("curl")(_p(vf), _p(cf), *vf.shape[1:], self.r(dx))
        return cf

    def max_speed(self, u):
        u = self.mac(u)
        return float(self.fn("max_speed")(*[_p(c) for c in u], *self.cell_dims(u)))

    def split_psi_hermite(self, kind, psi, T, dT, d2T, dx):
        psi, T, dT, d2T = self._map_args(psi, T, dT, d2T)
        nx, ny, nz = psi.shape[1:]
        shapes = [(3, nx + 1, ny, nz), (3, nx, ny + 1, nz), (3, nx, ny, nz + 1)]
        phi = [np.zeros(s, self.dt) for s in shapes]
        F = [np.zeros(s, self.dt) for s in shapes]
        self.fn("split_psi_hermite")(kind, _p(psi), _p(T), _p(dT), _p(d2T), *[_p(c) for c in phi],
                                     *[_p(c) for c in F], nx, ny, nz, self.r(dx))
        return phi, F

    def vortex_ring(self, w, rad, delta, c, ex, ey, X, vf, smoke, color, num_samples):
        """sim/init_conditions.py:13-36; vf (3,...) and smoke (4,...) are updated in place."""
        X = self.a(X)
        nx, ny, nz = X.shape[1:]
        assert vf.dtype == self.dt and smoke.dtype == self.dt
        self.fn("vortex_ring")(self.r(w), self.r(rad), self.r(delta), _p(self.a(c)), _p(self.a(ex)),
                               _p(self.a(ey)), _p(X), _p(vf), _p(smoke), _p(self.a(color)),
                               int(num_samples), nx, ny, nz)

    # -- projection ------------------------------------------------------------
    def poisson(self, u, levels, bt=None, bm=None, bvel=None, max_iters=400, stop_mode=0,
                abs_tol=1e-12, rel_tol=1e-12, want_f=False):
        """A whole MGPCG_3.Poisson / MGPCG_Solid_3.Poisson call; returns (u, p, stats[, f])."""
        u = [c.copy() for c in self.mac(u)]
        nx, ny, nz = self.cell_dims(u)
        bt = np.ascontiguousarray(bt if bt is not None else np.full(6, 2), np.int32)
        bm = np.ascontiguousarray(bm, np.int32) if bm is not None else None
        bvel = self.a(bvel) if bvel is not None else None
        p = np.zeros((nx, ny, nz), self.dt)
        f = [np.zeros_like(c) for c in u] if want_f else [None, None, None]
        stats = np.zeros(8, np.float64)
        self.fn("poisson")(*[_p(c) for c in u], nx, ny, nz, levels, _p(bt), _p(bm), _p(bvel),
                           int(max_iters), int(stop_mode), ctypes.c_double(abs_tol), ctypes.c_double(rel_tol),
                           _p(p), *[_p(c) for c in f], _p(stats))
        out = (u, p, stats)
        return out + (f,) if want_f else out

    def vcycle(self, r, levels, bt=None, bm=None):
        r = self.a(r)
        z = np.zeros_like(r)
        bt = np.ascontiguousarray(bt if bt is not None else np.full(6, 2), np.int32)
        bm = np.ascontiguousarray(bm, np.int32) if bm is not None else None
        self.fn("vcycle")(_p(r), _p(z), *r.shape, levels, _p(bt), _p(bm))
        return z

    def apply_A(self, p, bt=None, bm=None):
        p = self.a(p)
        Ap = np.zeros_like(p)
        bt = np.ascontiguousarray(bt if bt is not None else np.full(6, 2), np.int32)
        bm = np.ascontiguousarray(bm, np.int32) if bm is not None else None
        self.fn("apply_A")(_p(p), _p(Ap), *p.shape, _p(bt), _p(bm))
        return Ap

    def mgpcg_solve(self, b, levels, bt=None, bm=None, max_iters=400, stop_mode=1, abs_tol=1e-12,
                    rel_tol=1e-6):
        b = self.a(b)
        x = np.zeros_like(b)
        bt = np.ascontiguousarray(bt if bt is not None else np.full(6, 2), np.int32)
        bm = np.ascontiguousarray(bm, np.int32) if bm is not None else None
        stats = np.zeros(4, np.float64)
        self.fn("mgpcg_solve")(_p(b), _p(x), *b.shape, levels, _p(bt), _p(bm), int(max_iters),
                               int(stop_mode), ctypes.c_double(abs_tol), ctypes.c_double(rel_tol), _p(stats))
        return x, stats

    # -- particles (sim/lp.py) -----------------------------------------------------
    def lp_acc(self, u, vel, pos, num, dx, dt, St, gravity):
        u = self.mac(u)
        vel, pos = self.a(vel), self.a(pos)
        n = vel.shape[1]
        acc = np.zeros_like(vel)
        self.fn("lp_acc")(*[_p(c) for c in u], *self.cell_dims(u), _p(vel), _p(pos), _p(acc), ctypes.c_long(n),
                          ctypes.c_long(num), self.r(dx), self.r(dt), self.r(St), self.r(gravity))
        return acc

    def lp_vel(self, acc, vel, pos, num, dt, dx, L):
        vel = self.a(vel).copy()
        n = vel.shape[1]
        self.fn("lp_vel")(_p(self.a(acc)), _p(vel), _p(self.a(pos)), ctypes.c_long(n), ctypes.c_long(num),
                          self.r(dt), self.r(dx), *[self.r(x) for x in L])
        return vel

    def lp_vel_diff(self, u, vel, pos, num, dx, St, density_ratio, radius_ratio):
        u = self.mac(u)
        vel, pos = self.a(vel), self.a(pos)
        n = vel.shape[1]
        diff = np.zeros_like(vel)
        self.fn("lp_vel_diff")(*[_p(c) for c in u], *self.cell_dims(u), _p(vel), _p(diff), _p(pos),
                               ctypes.c_long(n), ctypes.c_long(num), self.r(dx), self.r(St),
                               self.r(density_ratio), self.r(radius_ratio))
        return diff

    def lp_p2g(self, diff, pos, ref_radius, num, dx, radius, L, dims):
        nx, ny, nz = dims
        out = [np.zeros((nx + 1, ny, nz), self.dt), np.zeros((nx, ny + 1, nz), self.dt),
               np.zeros((nx, ny, nz + 1), self.dt)]
        diff, pos, rr = self.a(diff), self.a(pos), self.a(ref_radius)
        n = diff.shape[1]
        self.fn("lp_p2g")(*[_p(c) for c in out], _p(diff), _p(pos), _p(rr), ctypes.c_long(n),
                          ctypes.c_long(num), self.r(dx), self.r(radius), *[self.r(x) for x in L], nx, ny, nz)
        return out

    def lp_pos(self, vel, pos, num, dt, L):
        pos = self.a(pos).copy()
        n = pos.shape[1]
        self.fn("lp_pos")(_p(self.a(vel)), _p(pos), ctypes.c_long(n), ctypes.c_long(num), self.r(dt),
                          *[self.r(x) for x in L])
        return pos

    def lp_viscous(self, u, dx, Re):
        u = self.a(u)
        vis = np.zeros_like(u)
        self.fn("lp_viscous")(_p(u), _p(vis), *u.shape, self.r(dx), self.r(Re))
        return vis
