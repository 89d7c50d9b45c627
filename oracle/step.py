"""The EDGE step loop of run_edge.py:588-711 composed from oracle kernels.

TEST INFRASTRUCTURE ONLY: the CPU restatement used for trajectory parity and
as the timed CPU baseline.  State is numpy arrays in the shared SoA layout;
every kernel is the oracle's literal restatement (edge_oracle*.cpp).
"""
import numpy as np

from .oracle import HK_FD, HK_FULL, Oracle


def face_coords(a, nx, ny, nz, dx):
    d = [nx + (a == 0), ny + (a == 1), nz + (a == 2)]
    i, j, k = np.meshgrid(np.arange(d[0]), np.arange(d[1]), np.arange(d[2]), indexing="ij")
    o = [0.0 if a == 0 else 0.5, 0.0 if a == 1 else 0.5, 0.0 if a == 2 else 0.5]
    return np.stack([(i + o[0]) * dx, (j + o[1]) * dx, (k + o[2]) * dx])


def dye_advance(o, u, pos, num, dx, dt):
    """Passive dye: each live particle (q < num of the (3, n) positions) takes
    advance_map_0's RK4 step with +dt (the forward face map's point march,
    run_edge.py:336-357) and is clamped into [0, n dx] per axis (the clamp of
    sim/lp.py:158-181)."""
    out = np.array(pos, copy=True)
    if num > 0:
        live = o.advance_map_0(np.ascontiguousarray(out[:, :num]), u, dx, dt)
        nx, ny, nz = u[0].shape[0] - 1, u[0].shape[1], u[0].shape[2]
        L = np.array([nx * dx, ny * dx, nz * dx], out.dtype)[:, None]
        out[:, :num] = np.minimum(np.maximum(live, 0), L)
    return out


class OracleStep:
    """Fields of run_edge.py:174-254 as numpy arrays; `substep()` = one loop iteration."""

    def __init__(self, u, smoke, dx, reinit_every, CFL, eps, levels, BFECC_clamp, dtype=np.float32,
                 stop_mode=1, rel_tol=1e-6, max_iters=400, scheme="edge"):
        self.o = Oracle(dtype)
        self.dt_ = np.dtype(dtype)
        self.u = [np.ascontiguousarray(c, dtype) for c in u]
        self.smoke = np.ascontiguousarray(smoke, dtype)
        self.nx, self.ny, self.nz = self.u[0].shape[0] - 1, self.u[0].shape[1], self.u[0].shape[2]
        self.dx, self.reinit_every, self.CFL, self.eps = dx, reinit_every, CFL, eps
        self.levels, self.BFECC_clamp = levels, BFECC_clamp
        self.stop_mode, self.rel_tol, self.max_iters = stop_mode, rel_tol, max_iters
        n = (self.nx, self.ny, self.nz)
        i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
        self.X = np.stack([(i + .5) * dx, (j + .5) * dx, (k + .5) * dx]).astype(dtype)
        self.i = -1
        self.iters = []
        self.lp = None
        # "edge" (run_edge.py), "ed8" (run_ed8.py), "ed4" (run_ed4.py), "ge" (run_ge.py), "efm" (run_efm.py)
        self.scheme = scheme
        self.u_cache = [None] * reinit_every   # EFM: u after each step's first projection
        self.dts = [0.0] * reinit_every
        self.dye = None

    def set_particles(self, pos, vel, ref_radius, num, params):
        """Laden particles (run_edge.py:235-254): pos/vel (3, n); params as the driver's lp dict."""
        self.lp = dict(params)
        self.pos = np.ascontiguousarray(pos, self.dt_)
        self.vel = np.ascontiguousarray(vel, self.dt_)
        self.rr = np.ascontiguousarray(ref_radius, self.dt_)
        self.num = int(num)
        self.domain = (self.nx * self.dx, self.ny * self.dx, self.nz * self.dx)

    def set_dye(self, pos, num):
        """Passive dye particles: positions (3, n), the first num live."""
        self.dye = np.ascontiguousarray(pos, self.dt_)
        self.dye_num = int(num)

    def march_full_map(self, dt):
        """ED8 (sim/hermite.py:1290-1338), ED4 through the FD-unrolled map
        (1375-1395) or GE (1398-1424), unfused, then the map's values and
        columns at the faces (ED8/GE: split_psi_hermite, run_ed8.py:190-223;
        ED4: the same evaluations inside map_u_psi and map_smoke_psi,
        run_ed4.py:219-262, 285-303)."""
        o, dx, n = self.o, self.dx, (self.nx, self.ny, self.nz)
        X = self.X.reshape(3, -1)
        if self.scheme == "ed4":
            P4 = []
            for sg in ((-1, -1, -1), (1, 1, -1), (-1, 1, 1), (1, -1, 1)):   # sim/hermite.py:737-742
                p = o.advance_map_0((X + self.eps * np.array(sg, float)[:, None]).astype(self.dt_), self.u, dx, -dt)
                p, _ = o.compose(HK_FD, self.psi, self.T, None, None, p, None, dx)
                P4.append(p.reshape((3,) + n))
            self.psi, self.T = o.calc_ed4_1(np.stack(P4), self.eps)
            self.psif, self.Tf = o.split_psi_hermite(HK_FD, self.psi, self.T, None, None, dx)
            return
        if self.scheme == "ed8":
            P8 = []
            for e in range(8):   # bit 0: x sign, bit 1: y, bit 2: z (sim/hermite.py:704-725)
                sg = np.array([1 if e & 1 else -1, 1 if e & 2 else -1, 1 if e & 4 else -1], float)
                p = o.advance_map_0((X + self.eps * sg[:, None]).astype(self.dt_), self.u, dx, -dt)
                p, _ = o.compose(HK_FULL, self.psi, self.T, self.dT, self.d2T, p, None, dx)
                P8.append(p.reshape((3,) + n))
            self.psi, self.T, self.dT, self.d2T = o.calc_ed8(np.stack(P8), self.eps)
        else:
            Tm = np.broadcast_to(np.eye(3, dtype=self.dt_)[:, :, None], (3, 3, X.shape[1])).copy()
            p, Tm = o.advance_map_1_transpose(X.copy(), Tm, self.u, dx, -dt)
            p, Tm = o.compose(HK_FULL, self.psi, self.T, self.dT, self.d2T, p, Tm, dx)
            self.psi, self.T = p.reshape((3,) + n), Tm.reshape((3, 3) + n)
            self.dT, self.d2T = o.fd3(self.T, None, dx, want_dT=True)
        self.psif, self.Tf = o.split_psi_hermite(HK_FULL, self.psi, self.T, self.dT, self.d2T, dx)

    def backtrack_psi_grid(self, j, dt):
        """run_efm.py:170-185, 355-360: cache u, then the face map from the identity
        marched back through the current and the cached velocities."""
        o, dx, n = self.o, self.dx, (self.nx, self.ny, self.nz)
        self.u_cache[j] = [c.copy() for c in self.u]
        self.dts[j] = dt
        psif = [face_coords(a, *n, dx).astype(self.dt_) for a in range(3)]
        Tf = [np.broadcast_to(np.eye(3)[a][:, None, None, None], psif[a].shape).astype(self.dt_) for a in range(3)]
        for a in range(3):
            psif[a], Tf[a] = o.advance_map_col(psif[a], Tf[a], self.u, dx, -dt)
        for step in reversed(range(j)):
            for a in range(3):
                psif[a], Tf[a] = o.advance_map_col(psif[a], Tf[a], self.u_cache[step], dx, -self.dts[step])
        self.psif, self.Tf = psif, Tf

    def particle_forces(self, dt):
        """run_edge.py:627-641 with the oracle's sim/lp.py restatement."""
        o, p, dx, L = self.o, self.lp, self.dx, self.domain
        acc = o.lp_acc(self.u, self.vel, self.pos, self.num, dx, dt, p["St"], p["gravity"])
        self.vel = o.lp_vel(acc, self.vel, self.pos, self.num, dt, dx, L)
        diff = o.lp_vel_diff(self.u, self.vel, self.pos, self.num, dx, p["St"], p["density_ratio"],
                             p["epiral_radius_ratio"])
        drag = o.lp_p2g(diff, self.pos, self.rr, self.num, dx, p["laden_radius"], L, (self.nx, self.ny, self.nz))
        self.pos = o.lp_pos(self.vel, self.pos, self.num, dt, L)
        self.vis = [(o.lp_viscous(c, dx, p["Re"]) + 1.0 * d).astype(self.dt_) for c, d in zip(self.u, drag)]

    def poisson(self):
        self.u, p, st = self.o.poisson(self.u, self.levels, stop_mode=self.stop_mode, rel_tol=self.rel_tol,
                                       abs_tol=1e-12, max_iters=self.max_iters)
        self.iters.append(int(st[0]))
        self.p = p

    def substep(self, dt=None):
        o, dx = self.o, self.dx
        self.i += 1
        j = self.i % self.reinit_every
        if dt is None:
            dt = self.CFL * dx / o.max_speed(self.u)
            if self.lp is not None:
                vmax = max(1e-3, float(np.sqrt((self.vel.astype(np.float64) ** 2).sum(0)).max()))
                dt = min(dt, self.CFL * dx / vmax)
        self.curr_dt = dt
        n = (self.nx, self.ny, self.nz)
        if j == 0:
            self.phi = [face_coords(a, *n, dx).astype(self.dt_) for a in range(3)]
            self.F = [np.broadcast_to(np.eye(3)[a][:, None, None, None], self.phi[a].shape).astype(self.dt_)
                      for a in range(3)]
            self.psi = self.X.copy()
            self.T = np.zeros((3, 3) + n, self.dt_)
            for c in range(3):
                self.T[c, c] = 1
            self.dT = np.zeros((3, 3) + n, self.dt_)
            self.d2T = np.zeros((3,) + n, self.dt_)
            self.init_u = [c.copy() for c in self.u]
            self.init_smoke = self.smoke.copy()
        if self.lp is not None:
            self.particle_forces(dt)
        self.u = o.advect_u_sl(self.u, dx, 0.5 * dt)
        self.poisson()
        if self.scheme == "edge":
            self.psi, self.T, self.dT = o.march_edge(self.psi, self.T, self.dT, self.u, self.X, dx, dt, self.eps)
        elif self.scheme == "efm":
            self.backtrack_psi_grid(j, dt)
        else:
            self.march_full_map(dt)
        for a in range(3):
            self.phi[a], self.F[a] = o.advance_map_col(self.phi[a], self.F[a], self.u, dx, dt)
        if self.dye is not None:
            self.dye = dye_advance(o, self.u, self.dye, self.dye_num, dx, dt)
        if self.lp is not None:
            # run_edge.py:661-667
            tmp = o.map_u(self.vis, self.F, self.phi, dx)
            self.init_u = [(i0 + dt * t).astype(self.dt_) for i0, t in zip(self.init_u, tmp)]
        self.bfecc()
        self.poisson()
        return dt

    def bfecc(self):
        """run_edge.py:669-700: pull init_u / init_smoke back through the
        backward map, push the result forward through (phi, F), and correct by
        half the round-trip error (then the clamps)."""
        o, dx = self.o, self.dx
        if self.scheme == "edge":
            def pull_u(u0):
                return o.map_u_psi(u0, self.psi, self.T, self.dT, dx)

            def pull_s(mode, s0, s1):
                return o.map_smoke_psi(mode, s0, s1, self.psi, self.T, self.dT, dx)
        else:
            def pull_u(u0):   # run_ed8.py:409-427: through the face values of the full map
                return o.map_u(u0, self.Tf, self.psif, dx)

            def pull_s(mode, s0, s1):
                return o.map_smoke(mode, s0, s1, self.psif, dx)
        self.u = pull_u(self.init_u)
        self.smoke = pull_s(0, self.init_smoke, None)
        err_u = o.map_u(self.u, self.F, self.phi, dx)
        err_s = o.map_smoke(0, self.smoke, None, self.phi, dx)
        err_u = [((e + -1. * i0) * 0.5).astype(self.dt_) for e, i0 in zip(err_u, self.init_u)]
        err_s = ((err_s + -1. * self.init_smoke) * 0.5).astype(self.dt_)
        tmp_u = pull_u(err_u)
        tmp_s = pull_s(0, err_s, None)
        err_u = [(u + -1. * t).astype(self.dt_) for u, t in zip(self.u, tmp_u)]
        self.smoke = (self.smoke + -1. * tmp_s).astype(self.dt_)
        self.smoke = pull_s(1, self.init_smoke, self.smoke)
        if self.BFECC_clamp:
            self.u = [o.clamp_u(e, u) for e, u in zip(err_u, self.u)]
        else:
            self.u = err_u
