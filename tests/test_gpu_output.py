"""Frame output and checkpoints on the GPU (run_edge.py:557-581, 719-737):
central vector and curl against the oracle, the .vti a frame writes, and a
resume from a mid-reinit checkpoint that continues bit for bit."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def scale_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_central_vector_and_curl_parity(orc32, dev):
    from edge_hip import _lib
    from edge_hip.field import stream_handle
    rng = np.random.default_rng(5)
    n, dx = (24, 20, 12), 1.0 / 20
    u = [rng.normal(size=(n[0] + 1, n[1], n[2])).astype(np.float32),
         rng.normal(size=(n[0], n[1] + 1, n[2])).astype(np.float32),
         rng.normal(size=(n[0], n[1], n[2] + 1)).astype(np.float32)]
    gu = [torch.from_numpy(c).to(dev) for c in u]
    vc = torch.empty((3,) + n, device=dev)
    cf = torch.empty((3,) + n, device=dev)
    _lib.call("edge_get_central_vector", *[c.data_ptr() for c in gu], vc.data_ptr(), *n, stream_handle())
    _lib.call("edge_curl", vc.data_ptr(), cf.data_ptr(), *n, float(dx), stream_handle())
    rvc = orc32.central_vector(u)
    assert np.array_equal(vc.cpu().numpy(), rvc)          # same two-term mean
    assert scale_err(cf.cpu().numpy(), orc32.curl(rvc, dx)) < 1e-6


def _sim(case="cfg1", res=(64, 32, 32), **kw):
    from run_edge import Simulation
    return Simulation(case, solver_mode="mixed", res=res, num_samples=400, **kw)


def _state(sim):
    names = ["u_x", "u_y", "u_z", "smoke"] + sim._state_fields()
    return {k: sim._field(k).data.detach().cpu().numpy().copy() for k in names}


def test_frame_vti_holds_vorticity_and_smoke(orc32, dev, tmp_path):
    from util import io_utils
    sim = _sim()
    sim.initialize()
    sim.substep()
    sim.write_frame(str(tmp_path), frame=3)
    d = io_utils.read_vti(os.path.join(str(tmp_path), "field_0003.vti"))
    u = sim.velocity_numpy()
    w = orc32.curl(orc32.central_vector(u), sim.dx)
    assert d["_dims"] == (64, 32, 32)
    assert scale_err(d["vorticity"], np.linalg.norm(w, axis=0)) < 1e-5
    assert np.array_equal(d["smoke"], sim.smoke.to_numpy())
    # the fused frame kernel against the reference's sequence on the GPU
    # (get_central_vector, curl, numpy's float32 norm): the same values up to
    # the contraction of the curl's products, i.e. within a few ulp
    ref = sim.vorticity_norm()
    assert ref.dtype == np.float32
    assert np.max(np.abs(d["vorticity"] - ref) / np.maximum(ref, 1e-30)) < 5e-7


def test_vtk_frame_one_channel_and_no_smoke(dev):
    from edge_hip import _lib
    from edge_hip.field import stream_handle
    rng = np.random.default_rng(8)
    n, dx = (12, 10, 8), 0.1
    u = [torch.from_numpy(rng.normal(size=s).astype(np.float32)).to(dev)
         for s in ((n[0] + 1, n[1], n[2]), (n[0], n[1] + 1, n[2]), (n[0], n[1], n[2] + 1))]
    sm0 = torch.from_numpy(rng.random(n).astype(np.float32)).to(dev)
    wn = torch.empty(int(np.prod(n)), device=dev)
    sm = torch.empty(int(np.prod(n)), device=dev)
    _lib.call("edge_vtk_frame", *[c.data_ptr() for c in u], sm0.data_ptr(), 1, wn.data_ptr(), sm.data_ptr(), *n,
              float(dx), stream_handle())
    got_w = wn.cpu().numpy().reshape(n, order="F")
    assert np.array_equal(sm.cpu().numpy().reshape(n, order="F"), sm0.cpu().numpy())
    wn2 = torch.zeros_like(wn)
    _lib.call("edge_vtk_frame", *[c.data_ptr() for c in u], None, 0, wn2.data_ptr(), None, *n, float(dx),
              stream_handle())
    assert torch.equal(wn, wn2)
    from oracle.oracle import Oracle
    o = Oracle(np.float64)
    w = o.curl(o.central_vector([c.cpu().numpy().astype(np.float64) for c in u]), dx)
    assert scale_err(got_w, np.linalg.norm(w, axis=0)) < 1e-6


@pytest.mark.parametrize("case", ["cfg1", "cfg5laden", "efm"])
def test_resume_mid_reinit_continues_bit_for_bit(dev, tmp_path, case):
    kw = {"cfg5laden": dict(case="cfg5laden", res=(32, 16, 16), num_particles=3000),
          "efm": dict(scheme="efm", reinit_every=4)}.get(case, {})
    a = _sim(**kw)
    a.initialize()
    for _ in range(3):
        a.substep()
    files = a.save_checkpoint(str(tmp_path), frame=5)
    assert all(os.path.exists(p) for p in files) and len(files) == 5
    for _ in range(2):
        a.substep()
    want = _state(a)
    b = _sim(**kw)
    b.initialize()
    assert b.load_checkpoint(str(tmp_path), 5)
    assert b.i == 2 and b.frame_idx == 0
    for _ in range(2):
        b.substep()
    got = _state(b)
    for k in want:
        if case != "cfg5laden":
            assert np.array_equal(got[k], want[k]), k
        else:
            # the drag scatter adds tile borders with FP32 atomics, in an order that
            # varies from run to run: the same state up to that rounding
            assert scale_err(got[k], want[k]) < 1e-4, k
    del a, b


def test_velocity_only_checkpoint_reinitialises_like_the_reference(dev, tmp_path):
    a = _sim()
    a.initialize()
    a.substep()
    a.save_checkpoint(str(tmp_path), frame=2, state=False)
    assert not os.path.exists(os.path.join(str(tmp_path), "edge_state_2.npz"))
    b = _sim()
    b.initialize()
    assert not b.load_checkpoint(str(tmp_path), 2)
    assert b.i == -1 and b.frame_idx == 2
    assert np.array_equal(b.u_x.to_numpy(), a.u_x.to_numpy())
    assert np.array_equal(b.smoke.to_numpy(), a.smoke.to_numpy())
    b.substep()   # i = 0: the maps restart from the identity (run_edge.py:587, 616)
    assert np.isfinite(b.u_x.to_numpy()).all()


def test_run_writes_frames_and_resumes(dev, tmp_path):
    from util import io_utils
    logs = str(tmp_path)
    a = _sim()
    a.visualize_dt = 0.05   # a frame every few steps (the case's own value never ends a frame)
    a.run(2, logs, log=lambda *_: None)
    for f in range(3):
        assert os.path.exists(os.path.join(logs, "vtks", "field_{:04d}.vti".format(f)))
        assert all(os.path.exists(p) for p in io_utils.velocity_paths(os.path.join(logs, "ckpts"), f))
    v2 = io_utils.read_vti(os.path.join(logs, "vtks", "field_0002.vti"))
    os.remove(os.path.join(logs, "vtks", "field_0002.vti"))
    b = _sim()
    b.visualize_dt = 0.05
    b.run(2, logs, from_frame=1, log=lambda *_: None)
    r2 = io_utils.read_vti(os.path.join(logs, "vtks", "field_0002.vti"))
    assert np.array_equal(r2["vorticity"], v2["vorticity"]) and np.array_equal(r2["smoke"], v2["smoke"])
