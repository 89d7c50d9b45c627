"""The measured workload at its real shape, 256x128x128 (four seeded rings):
one step against the oracle, and properties that do not depend on size."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RES = (256, 128, 128)


def scale_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.timeout(600)
def test_full_size_step_against_the_oracle(dev):
    """Two EDGE steps at 256x128x128, both projections converged to
    ||r|| <= 1e-6 ||b||, against the oracle's composition of the same step in
    FP64 (about 40 s of CPU time), with the FP32 oracle's own distance from it
    as the yardstick: at this resolution the FP32 rounding of the map's
    Jacobian near the walls (where the interpolation clamp binds for some
    RK stages and not others) alone moves T and u by ~1e-4 of scale in either
    implementation.  The bound: max(3 x the FP32 oracle's error, 1e-4) for
    velocity and smoke, max(3 x, 1e-5) for the map position."""
    from test_gpu_step import make
    from oracle.step import OracleStep
    sim, ref32 = make(RES, case="cfg2")
    ref64 = OracleStep([np.asarray(c, np.float64) for c in ref32.u], ref32.smoke.astype(np.float64), sim.dx,
                       sim.reinit_every, sim.CFL, sim.ed_eps, levels=sim.solver.n_mg_levels,
                       BFECC_clamp=sim.BFECC_clamp, dtype=np.float64, stop_mode=1, rel_tol=1e-6)
    # two steps: the second composes through a map that is no longer the identity
    for step in range(2):
        dt = sim.substep()
        torch.cuda.synchronize()
        print(f"full-size step {step} dt {dt:.4g}; FP32 oracle step running", flush=True)
        ref32.substep(dt=dt)
        print("FP64 oracle step running", flush=True)
        ref64.substep(dt=dt)
    pairs = [(g, r32, r64, 1e-4) for g, r32, r64 in zip(sim.velocity_numpy(), ref32.u, ref64.u)]
    pairs.append((np.moveaxis(sim.smoke.to_numpy(), -1, 0), ref32.smoke, ref64.smoke, 1e-4))
    pairs.append((np.moveaxis(sim.psi_garm.to_numpy(), -1, 0), ref32.psi, ref64.psi, 1e-5))
    for name, (g, r32, r64, floor) in zip(("u_x", "u_y", "u_z", "smoke", "psi"), pairs):
        e_gpu, e_32 = scale_err(g, r64), scale_err(r32, r64)
        print(f"EDGE {name}: GPU vs FP64 oracle {e_gpu:.3g}, FP32 oracle vs FP64 {e_32:.3g}")
        assert e_gpu <= max(3.0 * e_32, floor), (e_gpu, e_32)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("scheme", ["ed8", "ed4", "ge"])
def test_full_size_ed8_ge_step_against_the_oracle(dev, scheme):
    """Two ED8 (run_ed8.py:364-482), ED4 (run_ed4.py:401-515) or GE (run_ge.py:363-481) steps at
    256x128x128 with converged projections, against the FP64 oracle's step,
    with the FP32 oracle's own distance from it as the yardstick (as for EDGE
    above)."""
    from test_gpu_step import make
    from oracle.step import OracleStep
    sim, ref32 = make(RES, case="cfg2", scheme=scheme)
    ref64 = OracleStep([np.asarray(c, np.float64) for c in ref32.u], ref32.smoke.astype(np.float64), sim.dx,
                       sim.reinit_every, sim.CFL, sim.ed_eps, levels=sim.solver.n_mg_levels,
                       BFECC_clamp=sim.BFECC_clamp, dtype=np.float64, stop_mode=1, rel_tol=1e-6, scheme=scheme)
    # two steps: on the first the face values of the map depend on psi and T only (at a face
    # centre the cross terms' weights vanish), the second composes through the first map
    for step in range(2):
        dt = sim.substep()
        torch.cuda.synchronize()
        print(f"full-size {scheme} step {step} dt {dt:.4g}; FP32 oracle step running", flush=True)
        ref32.substep(dt=dt)
        print("FP64 oracle step running", flush=True)
        ref64.substep(dt=dt)
    pairs = [(g, r32, r64, 1e-4) for g, r32, r64 in zip(sim.velocity_numpy(), ref32.u, ref64.u)]
    pairs.append((np.moveaxis(sim.smoke.to_numpy(), -1, 0), ref32.smoke, ref64.smoke, 1e-4))
    pairs.append((np.moveaxis(sim.psi_garm.to_numpy(), -1, 0), ref32.psi, ref64.psi, 1e-5))
    for name, (g, r32, r64, floor) in zip(("u_x", "u_y", "u_z", "smoke", "psi"), pairs):
        e_gpu, e_32 = scale_err(g, r64), scale_err(r32, r64)
        print(f"{scheme} {name}: GPU vs FP64 oracle {e_gpu:.3g}, FP32 oracle vs FP64 {e_32:.3g}")
        assert e_gpu <= max(3.0 * e_32, floor), (e_gpu, e_32)


def test_full_size_mixed_steps_are_deterministic_and_converged(dev):
    """Two steps with the measured solver ("dense") twice from the same initial state: identical
    bits (FP64 reductions in a fixed order, no atomics in this case), every
    projection at ||b - A p|| <= 1e-6 ||b|| (FP64 on the device), the map inside
    the domain, and the kinetic energy within 5%."""
    from run_edge import Simulation
    out = []
    for _ in range(2):
        sim = Simulation("cfg2", solver_mode="dense", res=RES, num_samples=400, check_every=0)
        sim.initialize()
        e0 = sum(float((f.data.double() ** 2).sum()) for f in (sim.u_x, sim.u_y, sim.u_z))
        for _ in range(2):
            sim.substep()
        for st in sim.poisson_stats[-4:]:
            assert st["converged"] == 1 and st["rTr"] <= 1e-12 * st["b2"], st
        e1 = sum(float((f.data.double() ** 2).sum()) for f in (sim.u_x, sim.u_y, sim.u_z))
        assert abs(e1 / e0 - 1) < 0.05
        psi = sim.psi_garm.data
        L = torch.tensor([RES[0], RES[1], RES[2]], device=psi.device, dtype=psi.dtype) * sim.dx
        assert bool(((psi >= 0) & (psi <= L.view(3, 1, 1, 1))).all())
        out.append([f.to_numpy() for f in (sim.u_x, sim.u_y, sim.u_z, sim.smoke)])
        del sim
        torch.cuda.empty_cache()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.timeout(300)
def test_full_size_thirty_steps_invariants(dev):
    """The bench's workload (256x128x128, four rings, mixed projection started
    from the extrapolation of the last two steps' pressures) for 30 steps, across two map reinitialisations
    (every 12 steps): every projection converged (||r|| <= 1e-6 ||b||), the
    velocity divergence after the step's last projection at the solver's
    level (at most 1e-5 ||b||), the kinetic energy within 10% of the start (the reference scheme
    is energy-preserving up to its BFECC clamp and the projections), the
    smoke inside the initial range (the BFECC clamp), the map inside the
    domain, and a dt within the CFL bound every step."""
    from run_edge import Simulation
    from test_gpu_step import _fluid_div
    sim = Simulation("cfg2", solver_mode="mixed", res=RES, num_samples=400, check_every=0,
                     pressure_start="extrapolated")
    sim.initialize()
    e0 = sum(float((f.data.double() ** 2).sum()) for f in (sim.u_x, sim.u_y, sim.u_z))
    s_lo, s_hi = float(sim.smoke.data.min()), float(sim.smoke.data.max())
    L = torch.tensor([RES[0], RES[1], RES[2]], device=sim.psi_garm.data.device) * sim.dx
    dts = []
    for step in range(30):
        dts.append(sim.substep())
        st = sim.poisson_stats[-1]
        assert st["converged"] == 1 and st["rTr"] <= 1e-12 * st["b2"] * (1 + 1e-9), (step, st)
    torch.cuda.synchronize()
    assert sim.i == 29 and sim.i // sim.reinit_every == 2   # two reinitialisations passed
    for st in sim.poisson_stats[-60:]:
        assert st["converged"] == 1, st
    e1 = sum(float((f.data.double() ** 2).sum()) for f in (sim.u_x, sim.u_y, sim.u_z))
    print(f"30 steps: kinetic energy ratio {e1 / e0:.4f}, dt {min(dts):.4g} .. {max(dts):.4g}, "
          f"iterations {[s['iters'] for s in sim.poisson_stats[-6:]]}")
    assert abs(e1 / e0 - 1) < 0.10
    # div u = -(b - A p) after the subtraction: its largest cell is at most the final
    # residual's L2 norm, <= 1e-6 ||b|| (10x for the FP32 rounding of u and p)
    div = _fluid_div(sim.velocity_numpy(), np.zeros(RES, np.int32))
    b_norm = sim.poisson_stats[-1]["b2"] ** 0.5
    print(f"largest divergence {div:.3g}, ||b|| of the last projection {b_norm:.3g}")
    assert div <= 1e-5 * b_norm, (div, b_norm)
    sm = sim.smoke.data
    assert float(sm.min()) >= s_lo - 1e-6 and float(sm.max()) <= s_hi + 1e-6
    psi = sim.psi_garm.data
    assert bool(((psi >= 0) & (psi <= L.view(3, 1, 1, 1))).all())
    assert all(0 < d <= sim.CFL * sim.dx / 1e-3 for d in dts)


def test_full_size_plate_steps_converge_divergence_free(dev):
    """The triangular plate workload at 256x128x128 (four rings, inflow 1 at
    20 degrees, span 0.3), mixed projection: two steps, every projection at
    ||r|| <= 1e-6 ||b||, the fluid cells' divergence after the last one
    within 1e-4 of the largest divergence it removed, the plate in the mask."""
    from run_edge import Simulation
    from test_gpu_step import _fluid_div
    sim = Simulation("cfg3plate", solver_mode="mixed", res=RES, num_samples=400, check_every=0)
    sim.initialize()
    bm = sim.boundary_mask.to_numpy()
    assert bm[1:-1, 1:-1, 1:-1].sum() > 1000
    sim.rhs_snapshots = []
    for _ in range(2):
        sim.substep()
    for st in sim.poisson_stats[-4:]:
        assert st["converged"] == 1 and st["rTr"] <= 1e-12 * st["b2"] * 1.0001, st
    before = _fluid_div([t.cpu().numpy() for t in sim.rhs_snapshots[-1]], bm)
    after = _fluid_div(sim.velocity_numpy(), bm)
    print(f"plate 256x128x128: iterations {[s['iters'] for s in sim.poisson_stats]}, fluid divergence "
          f"{before:.3g} -> {after:.3g}")
    assert after <= 1e-4 * before


def test_full_size_dye_particles_follow_the_ring(dev):
    """The 512x256x256 torus with 16,777,216 passive dye particles, two steps
    (mixed projection): every particle stays a finite point inside the domain
    (none is lost), the projections converge, and the dye's centroid moves
    with the ring along its axis (+-x) while staying centred in y and z (the
    ring is symmetric about its axis)."""
    from run_edge import Simulation
    sim = Simulation("cfg5", solver_mode="mixed", num_samples=400, check_every=0)
    sim.initialize()
    n = sim.dye_num
    assert n == 16777216 and sim.res_x == 512
    c0 = sim.dye_pos.data.double().mean(dim=1).cpu().numpy()   # (3, n): component planes
    for _ in range(2):
        sim.substep()
    for st in sim.poisson_stats[-4:]:
        assert st["converged"] == 1, st
    p = sim.dye_pos.data
    L = torch.tensor([sim.res_x, sim.res_y, sim.res_z], device=p.device, dtype=p.dtype) * sim.dx
    assert bool(torch.isfinite(p).all())
    assert bool(((p >= 0) & (p <= L.view(3, 1))).all())
    c1 = p.double().mean(dim=1).cpu().numpy()
    d = c1 - c0
    print(f"dye centroid moved {d / sim.dx} cells in two steps")
    assert abs(d[0]) > 1e-3 * sim.dx
    assert abs(d[1]) < 0.05 * abs(d[0]) and abs(d[2]) < 0.05 * abs(d[0])


def test_full_size_learned_warm_start_gain(dev):
    """What the committed trained network (sim/unet_warmstart.npz) gains at
    256x128x128: the two projections of an EDGE step of the four-ring flow,
    solved by the mixed solver from zero and from the network's p0 (inference
    included), on the same right-hand sides, to ||r|| <= 1e-6 ||b||.  The
    iterations, milliseconds and the guess's relative residual are printed;
    the test holds the invariants: both starts converge to the same velocity
    (within the tolerance), and the guess is kept exactly when
    ||b - A p0|| <= ||b||."""
    from run_edge import Simulation
    from sim.learned import load_default
    from sim.mgpcg import MGPCG_3
    net = load_default()
    assert net is not None, "sim/unet_warmstart.npz is missing"
    sim = Simulation("cfg2", solver_mode="mixed", res=RES, num_samples=400, check_every=0)
    sim.initialize()
    sim.rhs_snapshots = []
    sim.substep()
    rhs = sim.rhs_snapshots
    bt = np.full((3, 2), 2)
    solvers = {"zero": MGPCG_3(bt, list(RES), mode="mixed", stop="relative", rel_tol=1e-6),
               "learned": MGPCG_3(bt, list(RES), mode="mixed", stop="relative", rel_tol=1e-6, warm_start=net),
               "zero_dense": MGPCG_3(bt, list(RES), mode="dense", stop="relative", rel_tol=1e-6),
               "learned_dense": MGPCG_3(bt, list(RES), mode="dense", stop="relative", rel_tol=1e-6, warm_start=net)}
    out = {}
    for which, base in enumerate(rhs):
        res = {}
        for name, s in solvers.items():
            times = []
            for rep in range(3):
                for f, b in zip((sim.tmp_u_x, sim.tmp_u_y, sim.tmp_u_z), base):
                    f.data.copy_(b)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                s.Poisson(sim.tmp_u_x, sim.tmp_u_y, sim.tmp_u_z)
                e1.record()
                torch.cuda.synchronize()
                times.append(e0.elapsed_time(e1))
            st = s.stats
            assert st["converged"] == 1.0 and st["rTr"] <= 1e-12 * st["b2"] * 1.0001
            res[name] = dict(ms=min(times[1:]), iters=st["iters"], u=[f.to_numpy() for f in
                                                                     (sim.tmp_u_x, sim.tmp_u_y, sim.tmp_u_z)])
            if name.startswith("learned"):
                r0 = (st["r02"] / st["b2"]) ** 0.5
                assert st["x0_used"] == (1.0 if st["r02"] <= st["b2"] else 0.0)
                res[name]["r0"] = r0
        for a in range(3):
            assert scale_err(res["learned"]["u"][a], res["zero"]["u"][a]) < 1e-4
            assert scale_err(res["learned_dense"]["u"][a], res["zero"]["u"][a]) < 1e-4
        out[which] = res
        for sfx, label in (("", "mixed"), ("_dense", "dense")):
            z, l = res["zero" + sfx], res["learned" + sfx]
            print(f"projection {which} ({label}): zero start {z['iters']} iterations {z['ms']:.3f} ms; "
                  f"learned start {l['iters']} iterations {l['ms']:.3f} ms (||b - A p0|| / ||b|| = {l['r0']:.3g})")
    for sfx, label in (("", "mixed"), ("_dense", "dense")):
        gain_ms = sum(out[w]["zero" + sfx]["ms"] - out[w]["learned" + sfx]["ms"] for w in out)
        gain_it = sum(out[w]["zero" + sfx]["iters"] - out[w]["learned" + sfx]["iters"] for w in out)
        print(f"learned warm start gain over both projections ({label}): {gain_it} iterations, {gain_ms:.3f} ms "
              f"(negative: a loss)")
