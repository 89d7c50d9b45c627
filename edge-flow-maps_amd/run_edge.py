"""Drop-in for the reference's run_edge.py: the EDGE step loop on MI355X.

`Simulation.substep()` is one iteration of the reference's `while True` loop
(run_edge.py:588-711): max speed -> dt, reinit every `reinit_every` steps,
semi-Lagrangian half-step + projection, EDGE march of the backward map,
forward face march, impulse/smoke pull-back with BFECC, clamps, projection.
Every kernel is a libedge_hip.so kernel; this file only sequences them.

The case table keeps the reference's cases 0 and 1 (run_edge.py:25-60) and
adds the benchmark configurations ("cfg1": 64x32x32 leapfrogging rings,
"cfg2": 256x128x128 four seeded rings, "cfg3": the reference's delta-wing
case 2 with the solid projection, "cfg3plate": the four rings plus a
triangular plate of span 0.3 in an inflow of speed 1 at 20 degrees (solid
projection), "cfg5": 512x256x256 with 16M passive dye particles that
follow the flow, "cfg5laden": the same with 16M laden particles, whose steps
add the particle forces of run_edge.py:627-641,661-667 and the projection
with a force buffer, run_edge.py:704-708).  Where the reference copies a field
whose source is dead afterwards (u <- tmp_u), the storages are exchanged
instead (Field.swap_storage); the values are identical.

Usage: python run_edge.py -c cfg1 [--steps N] [--solver plain|mixed]
"""
import argparse
import gc
import math
import os
import sys
import time

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

from edge_hip import _lib  # noqa: E402
from edge_hip.field import Field, stream_handle, ti  # noqa: E402
from sim.edge_driver import (calc_max_particle_speed, calc_max_speed, clamp_smoke_psi, map_smoke,  # noqa: E402
                             map_smoke_psi, map_u, map_u_psi, march_phi_grid, prefetch_max_particle_speed,
                             prefetch_max_speed, reset_to_identity)
from sim import lp as lp_kernels  # noqa: E402
from sim.edge_driver import backtrack_faces, clamp_smoke, split_psi_hermite  # noqa: E402
from sim.hermite import (advect_ps_garm, advect_ps_hermite_ed4_fd_unroll, advect_ps_hermite_ed8,  # noqa: E402
                         advect_psi_garm_ed4_fusion_unroll, advect_u_semi_lagrangian, reset_to_identity_1,
                         reset_to_identity_2, reset_to_identity_3)
from sim.init_conditions import (add_inflow, init_delta, init_four_vorts, init_rings,  # noqa: E402
                                 init_vorts_leapfrog, seeded_rings, set_boundary_mask_delta,
                                 set_boundary_mask_plate)
from sim.mgpcg import MGPCG_3  # noqa: E402
from sim.mgpcg_solid import MGPCG_Solid_3  # noqa: E402
from util import io_utils  # noqa: E402
from util.taichi_utils import (add_fields, center_coords_func, curl, get_central_vector,  # noqa: E402
                               split_central_vector, x_coords_func, y_coords_func, z_coords_func)

CASES = {
    # reference case 0 (run_edge.py:25-42)
    0: dict(init="leapfrog", res=(384, 128, 128), visualize_dt=0.1, reinit_every=20, CFL=0.5, BFECC_clamp=False,
            eps_frac=0.01, base_level=3),
    # reference case 1 (run_edge.py:43-60)
    1: dict(init="fourvorts", res=(128, 128, 256), visualize_dt=0.1, reinit_every=20, CFL=0.5, BFECC_clamp=True,
            eps_frac=0.025, base_level=3),
    # 64x32x32, dx = 1/32, two coaxial rings R=0.2, core 0.03, strength 1, gap 0.15, reinit 12, 4 MG levels
    "cfg1": dict(init="rings2", res=(64, 32, 32), visualize_dt=1e9, reinit_every=12, CFL=0.5, BFECC_clamp=False,
                 eps_frac=0.01, levels=4),
    # 256x128x128, dx = 1/128, four seeded rings R=0.15, core 0.025, strength 0.02, reinit 12
    "cfg2": dict(init="rings4", res=(256, 128, 128), visualize_dt=1e9, reinit_every=12, CFL=0.5, BFECC_clamp=True,
                 eps_frac=0.025, base_level=3),
    # reference case 2 (run_edge.py:61-79): the delta-wing plate in an oblique inflow, solid projection
    "cfg3": dict(init="delta", res=(256, 128, 128), visualize_dt=1e9, reinit_every=10, CFL=0.5, BFECC_clamp=True,
                 eps_frac=0.02, base_level=3, solid=True),
    # the 256x128x128 four-ring flow plus a triangular plate (span 0.3, flat, four cells thick) in a
    # uniform inflow of speed 1 at a 20-degree angle of attack, solid projection
    "cfg3plate": dict(init="plate", res=(256, 128, 128), visualize_dt=1e9, reinit_every=12, CFL=0.5,
                      BFECC_clamp=True, eps_frac=0.025, base_level=3, solid=True),
    # 512x256x256, dx = 1/256, one ring (R 0.25, core 0.08) carrying 16,777,216 passive dye particles
    # seeded uniformly in the torus of tube radius 0.08 (numpy default_rng(seed)), moved every step
    # like points of the forward face map (G2P of the velocity at each RK4 stage)
    "cfg5": dict(init="torus_dye", res=(512, 256, 256), visualize_dt=1e9, reinit_every=12, CFL=0.5,
                 BFECC_clamp=True, eps_frac=0.025, base_level=3,
                 dye=dict(max_num=16777216, ring_R=0.25, tube=0.08)),
    # the same ring carrying 16,777,216 laden (inertial, two-way coupled) particles instead: the
    # particle constants are those of the reference's ink cases (run_edge.py:80-140) with Re and St
    # set for this unit domain
    "cfg5laden": dict(init="torus", res=(512, 256, 256), visualize_dt=1e9, reinit_every=12, CFL=0.25,
                      BFECC_clamp=True,
                 eps_frac=0.025, base_level=3,
                 lp=dict(max_num=16777216, Re=1.0e4, St=5.0e-3, gravity=1.0, density_ratio=0.6,
                         epiral_radius_ratio=35.0, laden_radius=0.01, ring_R=0.25, tube=0.08)),
}


def freeze_setup_objects():
    """Before a long step loop: move every object made so far (modules, fields,
    solver state, all living until the end) out of the cyclic collector's
    generations, so that a full collection during the loop scans only the
    loop's own objects instead of pausing the host, and the GPU queue behind
    it, for a pass over every object of the process.  Process-wide: a caller
    that is not the program's own entry point pairs it with gc.unfreeze()."""
    gc.collect()
    gc.freeze()


def cell_order(pos, dx, res, tile=8):
    """Permutation putting particles (3, n) in 8x8x8-cell-tile-major cell order
    (the order the drag scatter groups them in): particles that share cells
    are stored together, so the per-particle gathers and the scatter read
    contiguous memory.  Particle results do not depend on the order."""
    res = np.asarray(res)
    c = np.clip(np.floor(pos.astype(np.float64) / dx), 0, (res - 1)[:, None]).astype(np.int64)
    t = c // tile
    nt = (res + tile - 1) // tile
    key = ((t[0] * nt[1] + t[1]) * nt[2] + t[2]) * tile ** 3 + ((c[0] % tile) * tile + c[1] % tile) * tile + c[2] % tile
    return np.argsort(key, kind="stable")


def torus_particles(n, centre, R, tube, seed=0, dx=None, res=None):
    """n points uniform in the solid torus (axis x) around `centre`: (3, n) float32,
    in cell order (cell_order) when dx and res are given."""
    rng = np.random.default_rng(seed)
    out = np.empty((3, 0), np.float32)
    c = np.asarray(centre, np.float64)
    while out.shape[1] < n:
        m = int((n - out.shape[1]) * 2.4) + 1024
        p = np.stack([rng.uniform(-tube, tube, m), rng.uniform(-R - tube, R + tube, m),
                      rng.uniform(-R - tube, R + tube, m)])
        rho = np.sqrt(p[1] ** 2 + p[2] ** 2)
        keep = p[0] ** 2 + (rho - R) ** 2 < tube ** 2
        out = np.concatenate([out, (p[:, keep] + c[:, None]).astype(np.float32)], axis=1)
    out = out[:, :n]
    if dx is not None:
        out = out[:, cell_order(out, dx, res)]
    return np.ascontiguousarray(out)


class Simulation:
    """The fields of run_edge.py:174-254 and one step of its loop."""

    def __init__(self, case="cfg1", solver_mode="plain", stop="relative", rel_tol=1e-6, res=None,
                 reinit_every=None, check_every=None, num_samples=2000, seed=0, num_particles=None, scheme="edge",
                 reuse_pressure=False, pressure_start=None):
        """scheme: "edge" (run_edge.py), "ed8" (run_ed8.py), "ed4" (run_ed4.py), "ge"
        (run_ge.py) or "efm" (run_efm.py): the backward map's march; ED8 and GE keep a full
        Hermite map, ED4 an FD-unrolled one (psi and T only), and they pull back through
        the map's face values; EFM keeps no cell map: it caches the velocity of every
        step since the last reinitialisation and re-marches the face map back through all
        of them each step.
        pressure_start: where each projection's solve starts: "zero" (the
        reference, sim/mgpcg.py:67-75), "previous" (the same projection's
        pressure of the previous step) or "extrapolated" (the linear
        extrapolation in time of that projection's last two pressures,
        p_n + (dt_new / dt_old) (p_n - p_{n-1})); the answer meets the same stop
        rule either way, and a guess with ||b - A x0|| > ||b|| is dropped for
        zero.  reuse_pressure=True without pressure_start means "previous"."""
        if scheme not in ("edge", "ed8", "ed4", "ge", "efm"):
            raise ValueError(f"scheme must be edge, ed8, ed4, ge or efm, got {scheme!r}")
        self.scheme = scheme
        cfg = dict(CASES[case])
        if res is not None:
            cfg["res"] = tuple(res)
        if reinit_every is not None:
            cfg["reinit_every"] = reinit_every
        self.cfg = cfg
        ti.init()
        self.res_x, self.res_y, self.res_z = rx, ry, rz = cfg["res"]
        self.dx = dx = 1.0 / ry
        self.CFL = cfg["CFL"]
        self.reinit_every = cfg["reinit_every"]
        self.visualize_dt = cfg["visualize_dt"]
        self.BFECC_clamp = cfg["BFECC_clamp"]
        self.ed_eps = cfg["eps_frac"] * dx
        boundary_types = np.full((3, 2), 2, np.int32)
        kw = dict(mode=solver_mode, stop=stop, check_every=check_every)
        if "levels" in cfg:
            kw["levels"] = cfg["levels"]
        self.solid = bool(cfg.get("solid"))
        if self.solid:
            # run_edge.py:168-172
            self.boundary_mask = ti.field(ti.i32, (rx, ry, rz))
            self.boundary_vel = ti.Vector.field(3, float, (rx, ry, rz))
            self.solver = MGPCG_Solid_3(boundary_types=boundary_types, boundary_mask=self.boundary_mask,
                                        boundary_vel=self.boundary_vel, N=[rx, ry, rz],
                                        base_level=cfg.get("base_level", 3),
                                        rel_tol=rel_tol if stop == "relative" else 1e-12, **kw)
            self.poisson_max_iters = 100   # sim/mgpcg_solid.py:415
        else:
            self.solver = MGPCG_3(boundary_types=boundary_types, N=[rx, ry, rz], base_level=cfg.get("base_level", 3),
                                  rel_tol=rel_tol if stop == "relative" else 1e-12, **kw)
            self.poisson_max_iters = 400   # sim/mgpcg.py:436
        self.poisson_stats = []
        if pressure_start is None:
            pressure_start = "previous" if reuse_pressure else "zero"
        if pressure_start not in ("zero", "previous", "extrapolated"):
            raise ValueError(f"pressure_start must be zero, previous or extrapolated, got {pressure_start!r}")
        self.pressure_start = pressure_start
        self.reuse_pressure = pressure_start != "zero"
        self._p_prev = [None, None]     # per projection: the last pressure
        self._p_old = [None, None]      # per projection: the one before (extrapolated start)
        self._p_time = [[], []]         # per projection: simulated times of those two solves
        self._p_kept = [True, True]     # per projection: did the last solve keep its guess
        self.t_sim = 0.0                # simulated time at the start of the current step
        V = ti.Vector.field
        F = ti.field
        cell, fx, fy, fz = (rx, ry, rz), (rx + 1, ry, rz), (rx, ry + 1, rz), (rx, ry, rz + 1)
        self.X = V(3, float, cell)
        self.X_x, self.X_y, self.X_z = V(3, float, fx), V(3, float, fy), V(3, float, fz)
        center_coords_func(self.X, dx)
        x_coords_func(self.X_x, dx)
        y_coords_func(self.X_y, dx)
        z_coords_func(self.X_z, dx)
        self.F_x, self.F_y, self.F_z = V(3, float, fx), V(3, float, fy), V(3, float, fz)
        self.phi_x, self.phi_y, self.phi_z = V(3, float, fx), V(3, float, fy), V(3, float, fz)
        self.u = V(3, float, cell)
        self.w = V(3, float, cell)   # vorticity of the frame output (run_edge.py:719-723)
        self.u_x, self.u_y, self.u_z = F(float, fx), F(float, fy), F(float, fz)
        self.init_u_x, self.init_u_y, self.init_u_z = F(float, fx), F(float, fy), F(float, fz)
        self.err_u_x, self.err_u_y, self.err_u_z = F(float, fx), F(float, fy), F(float, fz)
        self.tmp_u_x, self.tmp_u_y, self.tmp_u_z = F(float, fx), F(float, fy), F(float, fz)
        self.init_smoke = V(4, float, cell)
        self.smoke = V(4, float, cell)
        self.tmp_smoke = V(4, float, cell)
        self.err_smoke = V(4, float, cell)
        self.psi_garm = V(3, float, cell)
        self.T_x_garm, self.T_y_garm, self.T_z_garm = V(3, float, cell), V(3, float, cell), V(3, float, cell)
        self.dT_x_garm, self.dT_y_garm, self.dT_z_garm = V(3, float, cell), V(3, float, cell), V(3, float, cell)
        if scheme in ("ed8", "ge"):
            # run_ed8.py / run_ge.py: the map also holds d2T; psi and T at the faces (split_psi_hermite)
            self.d2T_x, self.d2T_y, self.d2T_z = F(float, cell), F(float, cell), F(float, cell)
        if scheme != "edge":
            self.psi_x, self.psi_y, self.psi_z = V(3, float, fx), V(3, float, fy), V(3, float, fz)
            self.T_x, self.T_y, self.T_z = V(3, float, fx), V(3, float, fy), V(3, float, fz)
        if scheme == "efm":
            # run_efm.py:104-109: the velocity after each step's first projection since the
            # last reinitialisation, and that step's dt
            self.u_cache = [(F(float, fx), F(float, fy), F(float, fz)) for _ in range(self.reinit_every)]
            self.dts = [0.0] * self.reinit_every
            self.efm_fused = True   # False: one reset and j+1 face-march launches, as the reference
        # scratch of the reference's fused march (accepted, unused by the in-place kernel)
        self.psi_tmp = self.T_x_tmp = self.T_y_tmp = self.T_z_tmp = None
        self.dT_x_tmp = self.dT_y_tmp = self.dT_z_tmp = None
        # laden particles (run_edge.py:235-254), when the case has them
        self.lp = dict(cfg["lp"]) if "lp" in cfg else None
        if self.lp is not None:
            if num_particles is not None:
                self.lp["max_num"] = int(num_particles)
            n = self.lp["max_num"]
            self.domain_range = (rx * dx, ry * dx, rz * dx)
            self.laden_particles_pos = V(3, float, n)
            self.laden_particles_vel = V(3, float, n)
            self.laden_particles_vel_diff = V(3, float, n)
            self.laden_particles_acc = V(3, float, n)
            self.laden_particle_ref_radius = F(float, n)
            self.drag_x, self.drag_y, self.drag_z = F(float, fx), F(float, fy), F(float, fz)
            self.vis_x, self.vis_y, self.vis_z = F(float, fx), F(float, fy), F(float, fz)
            self.laden_particle_num = n
        # passive dye particles (the 512x256x256 torus case): positions only
        self.dye = dict(cfg["dye"]) if "dye" in cfg else None
        if self.dye is not None:
            if num_particles is not None:
                self.dye["max_num"] = int(num_particles)
            self.dye_pos = V(3, float, self.dye["max_num"])
            self.dye_num = self.dye["max_num"]
        self.num_samples = num_samples
        self.seed = seed
        self.i = -1
        self.sub_t = 0.0
        self.frame_idx = 0
        self.output_frame = False
        self.timing = None   # dict name -> list of (start, end) torch.cuda.Event pairs, when enabled
        self.rhs_snapshots = None  # list: velocities handed to each projection, when enabled

    def _mark(self, name):
        """Record a HIP event on the current stream at a phase boundary (timing only)."""
        if self.timing is None:
            return
        import torch
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.timing.setdefault("_marks", []).append((name, ev))

    # -- initial condition (run_edge.py:538-557) ---------------------------------
    def initialize(self):
        init = self.cfg["init"]
        self.u.fill(0.0)
        if init == "leapfrog":
            init_vorts_leapfrog(self.X, self.u, self.smoke, self.tmp_smoke)
        elif init == "fourvorts":
            init_four_vorts(self.X, self.u, self.smoke, self.tmp_smoke)
        elif init == "rings2":
            init_vorts_leapfrog(self.X, self.u, self.smoke, self.tmp_smoke, radius=0.2, x_gap=0.15, x_start=0.35,
                                delta=0.03, w=1.0, num_samples=self.num_samples)
        elif init == "rings4":
            dom = (self.res_x * self.dx, self.res_y * self.dx, self.res_z * self.dx)
            init_rings(self.X, self.u, self.smoke, self.tmp_smoke, seeded_rings(4, dom, seed=self.seed),
                       num_samples=self.num_samples)
        elif init == "delta":
            # run_edge.py:545-548 (the per-step dye emission of run_edge.py:643-648 is not part of this build)
            init_delta(self.X, self.u, self.smoke, self.tmp_smoke)
            set_boundary_mask_delta(self.X, self.boundary_mask, self.boundary_vel, 0, self.dx)
        elif init == "plate":
            dom = (self.res_x * self.dx, self.res_y * self.dx, self.res_z * self.dx)
            init_rings(self.X, self.u, self.smoke, self.tmp_smoke, seeded_rings(4, dom, seed=self.seed),
                       num_samples=self.num_samples)
            add_inflow(self.u)
            set_boundary_mask_plate(self.boundary_mask, self.boundary_vel, self.dx)
        elif init == "torus_dye":
            dom = np.array([self.res_x, self.res_y, self.res_z]) * self.dx
            c = dom / 2
            R, tube = self.dye["ring_R"], self.dye["tube"]
            ring = dict(w=1.0, rad=R, delta=tube, c=tuple(c), ex=(0.0, 0.0, -1.0), ey=(0.0, 1.0, 0.0))
            init_rings(self.X, self.u, self.smoke, self.tmp_smoke, [ring], num_samples=self.num_samples)
            pos = torus_particles(self.dye_num, c, R, tube, seed=self.seed, dx=self.dx,
                                  res=(self.res_x, self.res_y, self.res_z))
            self.dye_pos.from_numpy(pos.T)
        elif init == "torus":
            dom = np.array([self.res_x, self.res_y, self.res_z]) * self.dx
            c = dom / 2
            R, tube = self.lp["ring_R"], self.lp["tube"]
            ring = dict(w=1.0, rad=R, delta=tube, c=tuple(c), ex=(0.0, 0.0, -1.0), ey=(0.0, 1.0, 0.0))
            init_rings(self.X, self.u, self.smoke, self.tmp_smoke, [ring], num_samples=self.num_samples)
            n = self.lp["max_num"]
            pos = torus_particles(n, c, R, tube, seed=self.seed, dx=self.dx,
                                  res=(self.res_x, self.res_y, self.res_z))
            self.laden_particles_pos.from_numpy(pos.T)
            self.laden_particles_vel.fill(0.0)
            self.laden_particle_ref_radius.fill(self.lp["laden_radius"])
        else:
            raise ValueError(init)
        split_central_vector(self.u, self.u_x, self.u_y, self.u_z)
        self.solver.Poisson(self.u_x, self.u_y, self.u_z)

    @property
    def garm_map(self):
        return (self.psi_garm, self.T_x_garm, self.T_y_garm, self.T_z_garm, self.dT_x_garm, self.dT_y_garm,
                self.dT_z_garm)

    def poisson(self, buffers=None, slot=0):
        """The projection.  slot: which of the step's two projections (0 after
        the semi-Lagrangian half step, 1 at the end of the step).  Unless
        pressure_start is "zero" the solve starts from a guess built from that
        projection's earlier pressures (kept only when ||b - A x0|| <= ||b||,
        same stop rule)."""
        if self.rhs_snapshots is not None:
            self.rhs_snapshots.append([self.u_x.data.clone(), self.u_y.data.clone(), self.u_z.data.clone()])
        x0 = None
        if self.reuse_pressure:
            # each projection keeps its own pressure buffers: the solver's p is pointed
            # at the guess, so the solve reads it (x0) at its start and writes the new
            # pressure into the same buffer at its end, with no copy in between
            t_now = self.t_sim + (0.5 if slot == 0 else 1.0) * self.curr_dt
            prev, old, times = self._p_prev[slot], self._p_old[slot], self._p_time[slot]
            if prev is None:
                prev = self._p_prev[slot] = torch.zeros_like(self.solver.p.data)
            elif self.pressure_start == "extrapolated":
                if old is None:
                    old = torch.empty_like(prev)
                if len(times) == 2 and self._p_kept[slot]:
                    # old <- prev + r (prev - old) = (1 + r) (prev - r / (1 + r) old)
                    r = (t_now - times[1]) / (times[1] - times[0])
                    _lib.call("edge_axpy_scale", prev.data_ptr(), old.data_ptr(), old.data_ptr(), -r / (1.0 + r),
                              1.0 + r, old.numel(), stream_handle())
                else:
                    # one pressure so far, or the last solve dropped its guess (a regime
                    # change, such as the second projection at a reinitialisation): start
                    # from the last pressure itself
                    old.copy_(prev)
                # the guess's buffer receives the new pressure; prev becomes the older one
                self._p_prev[slot], self._p_old[slot] = old, prev
                x0 = old
            else:
                x0 = prev
            self.solver.p.data = self._p_prev[slot]
            times.append(t_now)
            del times[:-2]
        else:
            self._p_time[slot].clear()   # steps solved from zero break the time series
        if buffers is None:
            self.solver.Poisson(self.u_x, self.u_y, self.u_z, max_iters=self.poisson_max_iters, x0=x0)
        else:
            self.solver.Poisson_with_buffer(self.u_x, self.u_y, self.u_z, *buffers, max_iters=self.poisson_max_iters,
                                            x0=x0)
        st = self.solver.stats
        self._p_kept[slot] = x0 is None or bool(st.get("x0_used", 1))
        self.poisson_stats.append(dict(st, time=self.solver.last_time))

    # -- one step (run_edge.py:588-711) --------------------------------------------
    def substep(self):
        self.i += 1
        j = self.i % self.reinit_every
        dx = self.dx
        max_speed = calc_max_speed(self.u_x, self.u_y, self.u_z)
        if self.lp is not None:
            max_speed2 = calc_max_particle_speed(self.laden_particles_vel)
            curr_dt = min(self.CFL * dx / max_speed, self.CFL * dx / max_speed2)
        else:
            curr_dt = self.CFL * dx / max_speed
        if self.sub_t + curr_dt >= self.visualize_dt:
            curr_dt = self.visualize_dt - self.sub_t
            self.sub_t = 0.
            self.frame_idx += 1
            self.output_frame = True
        elif self.sub_t + curr_dt * 2 >= self.visualize_dt:
            curr_dt = (self.visualize_dt - self.sub_t) / 2
            self.sub_t += curr_dt
            self.output_frame = False
        else:
            self.sub_t += curr_dt
            self.output_frame = False
        self.curr_dt = curr_dt

        if j == 0:
            reset_to_identity(self.phi_x, self.phi_y, self.phi_z, self.F_x, self.F_y, self.F_z, dx)
            if self.scheme == "edge":
                reset_to_identity_2(self.psi_garm, self.T_x_garm, self.T_y_garm, self.T_z_garm, self.dT_x_garm,
                                    self.dT_y_garm, self.dT_z_garm, self.X)
            elif self.scheme == "ed4":
                reset_to_identity_1(self.psi_garm, self.T_x_garm, self.T_y_garm, self.T_z_garm, self.X)
            elif self.scheme != "efm":
                reset_to_identity_3(self.psi_garm, self.T_x_garm, self.T_y_garm, self.T_z_garm, self.dT_x_garm,
                                    self.dT_y_garm, self.dT_z_garm, self.d2T_x, self.d2T_y, self.d2T_z, self.X)
            self.init_u_x.copy_from(self.u_x)
            self.init_u_y.copy_from(self.u_y)
            self.init_u_z.copy_from(self.u_z)
            self.init_smoke.copy_from(self.smoke)

        if self.lp is not None:
            self._mark("particles")
            self.particle_forces(curr_dt)

        self._mark("sl")
        advect_u_semi_lagrangian(self.u_x, self.u_y, self.u_z, self.tmp_u_x, self.tmp_u_y, self.tmp_u_z, dx,
                                 0.5 * curr_dt, self.X_x, self.X_y, self.X_z)
        self.u_x.swap_storage(self.tmp_u_x)   # u.copy_from(tmp_u): tmp_u is rewritten before it is read
        self.u_y.swap_storage(self.tmp_u_y)
        self.u_z.swap_storage(self.tmp_u_z)
        self._mark("poisson1")
        self.poisson(slot=0)

        self._mark("march")
        if self.scheme == "edge":
            advect_psi_garm_ed4_fusion_unroll(self.psi_tmp, self.T_x_tmp, self.T_y_tmp, self.T_z_tmp, self.dT_x_tmp,
                                              self.dT_y_tmp, self.dT_z_tmp, self.psi_garm, self.T_x_garm,
                                              self.T_y_garm, self.T_z_garm, self.dT_x_garm, self.dT_y_garm,
                                              self.dT_z_garm, self.u_x, self.u_y, self.u_z, self.X, dx, curr_dt,
                                              self.ed_eps)
        elif self.scheme == "efm":
            self.backtrack_psi_grid(j, curr_dt)
        else:
            self.march_full_map(curr_dt)
        self._mark("phi_march")
        march_phi_grid(self.phi_x, self.phi_y, self.phi_z, self.F_x, self.F_y, self.F_z, self.u_x, self.u_y,
                       self.u_z, dx, curr_dt)
        if self.dye is not None:
            # the dye moves with the same velocity and RK4 as the face map's points
            self._mark("dye")
            lp_kernels.advect_dye_particles(self.u_x, self.u_y, self.u_z, self.dye_pos, dx, curr_dt, self.dye_num)

        self._mark("maps")
        if self.lp is not None:
            # run_edge.py:661-667: the force pulled back through the face map, into init_u
            map_u(self.vis_x, self.vis_y, self.vis_z, self.tmp_u_x, self.tmp_u_y, self.tmp_u_z, self.F_x, self.F_y,
                  self.F_z, self.phi_x, self.phi_y, self.phi_z, dx)
            add_fields(self.init_u_x, self.tmp_u_x, self.init_u_x, curr_dt)
            add_fields(self.init_u_y, self.tmp_u_y, self.init_u_y, curr_dt)
            add_fields(self.init_u_z, self.tmp_u_z, self.init_u_z, curr_dt)
        if self.scheme == "edge":
            gm = self.garm_map

            def pull_u(u0, u1, sub_clamp=None):
                map_u_psi(*u0, *u1, dx, gm, sub_clamp=sub_clamp)

            def pull_s(s0, s1, clamp=False, sub_map=None):
                if clamp:
                    clamp_smoke_psi(s0, s1, dx, gm, sub_map=sub_map)
                else:
                    map_smoke_psi(s0, s1, dx, gm)
        else:
            # run_ed8.py:409-433 (run_ed4.py:447-467): through the map's face values
            def pull_u(u0, u1, sub_clamp=None):
                map_u(*u0, *u1, self.T_x, self.T_y, self.T_z, self.psi_x, self.psi_y, self.psi_z, dx,
                      sub_clamp=sub_clamp)

            def pull_s(s0, s1, clamp=False, sub_map=None):
                if clamp:
                    clamp_smoke(s0, s1, self.psi_x, self.psi_y, self.psi_z, dx, sub_map=sub_map)
                else:
                    map_smoke(s0, s1, self.psi_x, self.psi_y, self.psi_z, dx)
        U, IU = (self.u_x, self.u_y, self.u_z), (self.init_u_x, self.init_u_y, self.init_u_z)
        EU, TU = (self.err_u_x, self.err_u_y, self.err_u_z), (self.tmp_u_x, self.tmp_u_y, self.tmp_u_z)
        pull_u(IU, U)
        pull_s(self.init_smoke, self.smoke)
        # run_edge.py:669-682: err = (map(u) - init) * 0.5 in the map's own pass
        map_u(*U, *EU, self.F_x, self.F_y, self.F_z, self.phi_x, self.phi_y, self.phi_z, dx,
              add_scale=(*IU, -1., 0.5))
        map_smoke(self.smoke, self.err_smoke, self.phi_x, self.phi_y, self.phi_z, dx,
                  add_scale=(self.init_smoke, -1., 0.5))
        if self.BFECC_clamp:
            # run_edge.py:683-693: tmp_u = map(err_u), err_u = u - tmp_u, tmp_u =
            # clamp_u(err_u) in one pass (neither intermediate is stored)
            pull_u(EU, TU, sub_clamp=U)
        else:
            pull_u(EU, TU)
            for u_, t_, e_ in zip(U, TU, EU):
                add_fields(u_, t_, e_, -1.)
        # run_edge.py:684-689: tmp_smoke = map(err_smoke), smoke = clamp(smoke -
        # tmp_smoke), one pass (the pulled-back error stays in registers)
        pull_s(self.init_smoke, self.smoke, clamp=True, sub_map=self.err_smoke)
        if self.BFECC_clamp:
            self.u_x.swap_storage(self.tmp_u_x)
            self.u_y.swap_storage(self.tmp_u_y)
            self.u_z.swap_storage(self.tmp_u_z)
        else:
            self.u_x.swap_storage(self.err_u_x)
            self.u_y.swap_storage(self.err_u_y)
            self.u_z.swap_storage(self.err_u_z)
        self._mark("poisson2")
        if self.lp is not None:
            # run_edge.py:704-708
            self.vis_x.fill(0.0)
            self.vis_y.fill(0.0)
            self.vis_z.fill(0.0)
            self.poisson(buffers=(self.vis_x, self.vis_y, self.vis_z), slot=1)
        else:
            self.poisson(slot=1)
        # the next step's CFL input, launched now so that it is ready when that step starts
        prefetch_max_speed(self.u_x, self.u_y, self.u_z)
        if self.lp is not None:
            prefetch_max_particle_speed(self.laden_particles_vel)
        self._mark("end")
        self.t_sim += curr_dt
        return curr_dt

    def backtrack_psi_grid(self, j, curr_dt):
        """run_efm.py:170-185 after caching u (run_efm.py:355-357): the face map psi_*, T_*
        from the identity, marched back (-dt) through the current velocity and then through
        the cached velocities of steps j-1 .. 0 (the reference first copies each cached
        field into tmp_u; the march reads the cache directly)."""
        dx = self.dx
        self.dts[j] = curr_dt
        for c, f in zip(self.u_cache[j], (self.u_x, self.u_y, self.u_z)):
            c.copy_from(f)
        faces = (self.psi_x, self.psi_y, self.psi_z, self.T_x, self.T_y, self.T_z)
        steps = list(reversed(range(j)))
        if self.efm_fused:
            # one kernel: each face marched through every velocity in registers
            backtrack_faces(*faces, [(self.u_x, self.u_y, self.u_z)] + [self.u_cache[s] for s in steps],
                            [-curr_dt] + [-self.dts[s] for s in steps], dx)
            return
        reset_to_identity(*faces, dx)
        march_phi_grid(*faces, self.u_x, self.u_y, self.u_z, dx, -curr_dt)
        for step in steps:
            march_phi_grid(*faces, *self.u_cache[step], dx, -self.dts[step])

    def march_full_map(self, curr_dt):
        """run_ed8.py:226-230 / run_ge.py:220-231: the ED8 or GE march of the full
        Hermite map, then psi and T at the faces; run_ed4.py:167-169: the ED4 march
        of the FD-unrolled map, whose face values run_ed4.py's map_u_psi and
        map_smoke_psi evaluate (here once per step, shared by all pull-backs)."""
        dx = self.dx
        if self.scheme == "ed4":
            fd = (self.psi_garm, self.T_x_garm, self.T_y_garm, self.T_z_garm)
            advect_ps_hermite_ed4_fd_unroll(None, None, None, None, *fd, self.u_x, self.u_y, self.u_z, self.X, dx,
                                            curr_dt, self.ed_eps)
            split_psi_hermite(self.psi_x, self.psi_y, self.psi_z, self.T_x, self.T_y, self.T_z, fd, dx)
            return
        full = (self.psi_garm, self.T_x_garm, self.T_y_garm, self.T_z_garm, self.dT_x_garm, self.dT_y_garm,
                self.dT_z_garm, self.d2T_x, self.d2T_y, self.d2T_z)
        if self.scheme == "ed8":
            advect_ps_hermite_ed8(*([None] * 8), *full, self.u_x, self.u_y, self.u_z, self.X, dx, curr_dt,
                                  self.ed_eps)
        else:
            advect_ps_garm(None, None, None, None, *full, self.u_x, self.u_y, self.u_z, self.X, dx, curr_dt)
        split_psi_hermite(self.psi_x, self.psi_y, self.psi_z, self.T_x, self.T_y, self.T_z, full, dx)

    def particle_forces(self, curr_dt):
        """run_edge.py:627-641: particle drag and viscous force into vis_*."""
        p, dx = self.lp, self.dx
        n = self.laden_particle_num
        Lx, Ly, Lz = self.domain_range
        lp_kernels.update_particles_acc(self.u_x, self.u_y, self.u_z, self.laden_particles_vel,
                                        self.laden_particles_pos, self.laden_particles_acc, dx, curr_dt, n, p["St"],
                                        p["gravity"])
        lp_kernels.update_particles_velocity(self.laden_particles_acc, self.laden_particles_vel,
                                             self.laden_particles_pos, curr_dt, dx, n, Lx, Ly, Lz)
        lp_kernels.calculate_velocity_diff(self.u_x, self.u_y, self.u_z, self.laden_particles_vel,
                                           self.laden_particles_vel_diff, self.laden_particles_pos, dx, n, p["St"],
                                           p["density_ratio"], p["epiral_radius_ratio"])
        lp_kernels.laden_particles_back_with_computing_particle(
            self.drag_x, self.drag_y, self.drag_z, self.u_x, self.u_y, self.u_z, self.laden_particles_vel,
            self.laden_particles_vel_diff, self.laden_particles_pos, self.laden_particle_ref_radius, dx, n,
            p["laden_radius"], Lx, Ly, Lz, self.res_x, self.res_y, self.res_z)
        lp_kernels.update_particles_position(self.laden_particles_vel, self.laden_particles_pos, curr_dt, n, Lx, Ly,
                                             Lz)
        lp_kernels.calculate_viscous_force(self.u_x, self.u_y, self.u_z, self.vis_x, self.vis_y, self.vis_z, dx,
                                           p["Re"])
        add_fields(self.vis_x, self.drag_x, self.vis_x, 1.0)
        add_fields(self.vis_y, self.drag_y, self.vis_y, 1.0)
        add_fields(self.vis_z, self.drag_z, self.vis_z, 1.0)

    def velocity_numpy(self):
        return [self.u_x.to_numpy(), self.u_y.to_numpy(), self.u_z.to_numpy()]

    # -- frame output and checkpoints (run_edge.py:557-581, 719-737) ----------------
    def vorticity_norm(self):
        """|curl u| at the cells (run_edge.py:720-723): face means, curl on the GPU, norm on the host."""
        get_central_vector(self.u_x, self.u_y, self.u_z, self.u)
        curl(self.u, self.w, self.dx)
        return np.linalg.norm(self.w.to_numpy(), axis=-1)

    def _frame_buffers(self):
        if getattr(self, "_fbuf", None) is None:
            n = self.res_x * self.res_y * self.res_z
            dev = self.u_x.data.device
            self._fbuf = (torch.empty(n, device=dev), torch.empty(4 * n, device=dev),
                          torch.empty(n, pin_memory=True), torch.empty(4 * n, pin_memory=True))
        return self._fbuf

    def write_frame(self, vtkdir, particle_vtkdir=None, frame=None):
        """The visualization block of run_edge.py:719-727: field_<frame>.vti with
        "vorticity" (|curl u|) and the 4-channel "smoke", and the particles'
        field_<frame>.vtu.  The norm and the VTK point order are made on the GPU
        (one kernel), then one copy per array to pinned host memory."""
        frame = self.frame_idx if frame is None else frame
        wn, sm, hw, hs = self._frame_buffers()
        _lib.call("edge_vtk_frame", self.u_x.ptr, self.u_y.ptr, self.u_z.ptr, self.smoke.ptr, 4, wn.data_ptr(),
                  sm.data_ptr(), self.res_x, self.res_y, self.res_z, float(self.dx), stream_handle())
        hw.copy_(wn, non_blocking=True)
        hs.copy_(sm, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        io_utils.write_vti_points(os.path.join(vtkdir, "field_{:04d}.vti".format(frame)),
                                  (self.res_x, self.res_y, self.res_z),
                                  [("vorticity", hw.numpy()), ("smoke", hs.numpy().reshape(-1, 4))], "vorticity")
        if self.lp is not None and particle_vtkdir is not None:
            pos = self.laden_particles_pos.to_numpy()[:self.laden_particle_num, :]
            io_utils.write_to_vtks_lp(pos, {}, {}, os.path.join(particle_vtkdir, "field_{:03d}".format(frame)))
        if self.dye is not None and particle_vtkdir is not None:
            pos = self.dye_pos.to_numpy()[:self.dye_num, :]
            io_utils.write_to_vtks_lp(pos, {}, {}, os.path.join(particle_vtkdir, "field_{:03d}".format(frame)))

    def _state_fields(self):
        """Fields a step reads before it writes them, besides u and smoke: the two maps,
        the fields at the last reinitialisation and the particles' state."""
        names = ["init_u_x", "init_u_y", "init_u_z", "init_smoke", "phi_x", "phi_y", "phi_z", "F_x", "F_y", "F_z"]
        if self.scheme != "efm":
            names += ["psi_garm", "T_x_garm", "T_y_garm", "T_z_garm"]
        if self.scheme not in ("efm", "ed4"):
            names += ["dT_x_garm", "dT_y_garm", "dT_z_garm"]
        if self.scheme in ("ed8", "ge"):
            names += ["d2T_x", "d2T_y", "d2T_z"]
        if self.scheme == "efm":
            names += [f"u_{a}_cache_{s}" for s in range(self.reinit_every) for a in "xyz"]
        if self.dye is not None:
            names += ["dye_pos"]
        if self.lp is not None:
            names += ["laden_particles_pos", "laden_particles_vel", "laden_particles_acc",
                      "laden_particles_vel_diff", "laden_particle_ref_radius", "vis_x", "vis_y", "vis_z"]
        return names

    def _field(self, name):
        if name.startswith("u_") and "_cache_" in name:   # u_{x,y,z}_cache_<step> (EFM)
            a, step = name[2], int(name.rsplit("_", 1)[1])
            return self.u_cache[step]["xyz".index(a)]
        return getattr(self, name)

    def save_checkpoint(self, ckptdir, frame=None, state=True):
        """vel_{x,y,z}_numpy_<frame>.npy (run_edge.py:576-581, 731-737), smoke_numpy_<frame>.npy,
        and with state=True edge_state_<frame>.npz: the maps, the fields at the last
        reinitialisation, the particles and the step counters, so that a run resumes
        mid-reinit and continues bit for bit.  Returns the files written."""
        frame = self.frame_idx if frame is None else frame
        os.makedirs(ckptdir, exist_ok=True)
        paths = io_utils.velocity_paths(ckptdir, frame)
        for p, f in zip(paths, (self.u_x, self.u_y, self.u_z)):
            np.save(p, f.to_numpy())
        np.save(io_utils.smoke_path(ckptdir, frame), self.smoke.to_numpy())
        paths.append(io_utils.smoke_path(ckptdir, frame))
        if state:
            sp = io_utils.state_path(ckptdir, frame)
            arrays = {k: self._field(k).data.detach().cpu().numpy() for k in self._state_fields()}
            if self.scheme == "efm":
                arrays["dts"] = np.asarray(self.dts, np.float64)
            io_utils.save_state(sp, arrays, dict(i=self.i, sub_t=float(self.sub_t), frame_idx=self.frame_idx,
                                                 reinit_every=self.reinit_every))
            paths.append(sp)
        return paths

    def load_checkpoint(self, ckptdir, frame):
        """Resume from frame `frame` after initialize() (run_edge.py:557-566).  Velocity from
        the .npy files; smoke when its file exists (the reference loads it, but never writes
        it); with edge_state_<frame>.npz the whole map state and the step counters, else,
        as the reference, the next step reinitialises the maps.  Returns True when the
        map state was restored."""
        for p, f in zip(io_utils.velocity_paths(ckptdir, frame), (self.u_x, self.u_y, self.u_z)):
            f.from_numpy(np.load(p, allow_pickle=False))
        sm = io_utils.smoke_path(ckptdir, frame)
        if os.path.exists(sm):
            self.smoke.from_numpy(np.load(sm, allow_pickle=False))
        self.frame_idx = int(frame)
        self.sub_t = 0.0
        self.i = -1
        # the pressures of the steps before the load are no guess for the steps after it
        self._p_prev, self._p_old, self._p_time, self._p_kept = [None, None], [None, None], [[], []], [True, True]
        sp = io_utils.state_path(ckptdir, frame)
        if not os.path.exists(sp):
            return False
        arrays, counters = io_utils.load_state(sp)
        if counters.get("reinit_every", self.reinit_every) != self.reinit_every:
            raise ValueError(f"{sp}: saved with reinit_every {counters['reinit_every']}, "
                             f"this run uses {self.reinit_every}")
        names = self._state_fields()
        missing = [k for k in names if k not in arrays]
        if missing:
            raise ValueError(f"{sp}: missing fields {missing}")
        for k in names:
            f = self._field(k)
            a = arrays[k]
            if tuple(a.shape) != tuple(f.data.shape):
                raise ValueError(f"{sp}: field {k} has shape {a.shape}, expected {tuple(f.data.shape)}")
            f.data.copy_(torch.from_numpy(a).to(f.data.dtype))
            f.touch()
        if self.scheme == "efm":
            self.dts = [float(v) for v in arrays["dts"]]
        self.i = int(counters["i"])
        self.sub_t = float(counters["sub_t"])
        self.frame_idx = int(counters["frame_idx"])
        return True

    def run(self, total_frames, logsdir, from_frame=0, ckpt_every=1, state=True, log=print):
        """The frame loop of run_edge.py:522-745: a first frame (written at from_frame), then
        steps until frame total_frames, writing vtks/field_<frame>.vti (and particle/ files)
        on every output frame and ckpts/ every ckpt_every frames."""
        vtkdir, ckptdir = os.path.join(logsdir, "vtks"), os.path.join(logsdir, "ckpts")
        pdir = os.path.join(logsdir, "particle") if (self.lp is not None or self.dye is not None) else None
        for d in (vtkdir, ckptdir) + ((pdir,) if pdir else ()):
            os.makedirs(d, exist_ok=True)
        self.initialize()
        if from_frame > 0:
            self.load_checkpoint(ckptdir, from_frame)
        self.write_frame(vtkdir, pdir, self.frame_idx)
        self.save_checkpoint(ckptdir, self.frame_idx, state=state)
        freeze_setup_objects()
        try:
            while self.frame_idx < total_frames:
                t0 = time.perf_counter()
                dt = self.substep()
                log(f"[Simulate] Done step {self.i} dt {dt:.5g} poisson iters {self.poisson_stats[-1]['iters']}")
                if self.output_frame:
                    self.write_frame(vtkdir, pdir)
                    if self.frame_idx % ckpt_every == 0:
                        self.save_checkpoint(ckptdir, state=state)
                    ti.sync()
                    log(f"[Simulate] Done Frame {self.frame_idx} time {time.perf_counter() - t0:.2f}s")
        finally:
            # hand the frozen objects back to the collector: the freeze is
            # process-wide and must not outlive this loop
            gc.unfreeze()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-c", "--case", default="cfg1")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--solver", default="plain", choices=["plain", "mixed", "dense"])
    ap.add_argument("--res", type=int, nargs=3, default=None)
    ap.add_argument("--particles", type=int, default=None, help="particle count of a particle case")
    ap.add_argument("--scheme", default="edge", choices=["edge", "ed8", "ed4", "ge", "efm"])
    a = ap.parse_args(argv)
    case = int(a.case) if a.case.isdigit() else a.case
    sim = Simulation(case, solver_mode=a.solver, res=a.res, num_particles=a.particles, scheme=a.scheme)
    sim.initialize()
    ti.sync()
    t0 = time.perf_counter()
    for s in range(a.steps):
        dt = sim.substep()
        st = sim.poisson_stats[-1]
        print(f"[Simulate] step {s} dt {dt:.5g} poisson iters {st['iters']}", flush=True)
    ti.sync()
    print(f"{a.steps} steps in {time.perf_counter() - t0:.2f}s")


if __name__ == "__main__":
    main()
