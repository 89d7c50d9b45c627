"""Frame and checkpoint files (util/io_utils.py, run_edge.py:557-581, 719-737):
VTK XML round trips, the point ordering and component layout the reference's
vtk/pyevtk calls produce, and the checkpoint names and state archive."""
import os

import numpy as np

from util import io_utils as io


def _raw_blocks(path):
    """The appended blocks of a file, in order, as raw bytes (independent of read_vti)."""
    raw = open(path, "rb").read()
    p = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    out = []
    while raw[p:p + 1] != b"\n":
        n = int(np.frombuffer(raw[p:p + 8], np.uint64)[0])
        out.append(raw[p + 8:p + 8 + n])
        p += 8 + n
    return out


def test_vti_four_channel_round_trip_and_point_order(tmp_path):
    rng = np.random.default_rng(0)
    n = (5, 4, 3)
    w = rng.random(n).astype(np.float32)
    smoke = rng.random(n + (4,)).astype(np.float32)
    io.write_vtks_4channel_smoke(w, smoke, str(tmp_path), 7)
    path = os.path.join(str(tmp_path), "field_0007.vti")
    d = io.read_vti(path)
    assert d["_dims"] == n and d["_origin"] == (0, 0, 0) and d["_spacing"] == (1, 1, 1)
    assert d["_scalars"] == "vorticity"
    assert np.array_equal(d["vorticity"], w) and np.array_equal(d["smoke"], smoke)
    # VTK point order: x fastest; smoke components interleaved per point
    bw, bs = _raw_blocks(path)
    vw = np.frombuffer(bw, np.float32)
    vs = np.frombuffer(bs, np.float32)
    for (i, j, k) in [(0, 0, 0), (4, 0, 0), (1, 2, 0), (3, 1, 2)]:
        p = i + n[0] * (j + n[1] * k)
        assert vw[p] == w[i, j, k]
        assert np.array_equal(vs[4 * p:4 * p + 4], smoke[i, j, k])


def test_vti_one_channel(tmp_path):
    n = (3, 3, 2)
    w = np.arange(np.prod(n), dtype=np.float32).reshape(n)
    io.write_vtks(w[None], (2 * w)[..., None], str(tmp_path), 12)   # squeezed like the reference
    d = io.read_vti(os.path.join(str(tmp_path), "field_0012.vti"))
    assert np.array_equal(d["vorticity"], w) and np.array_equal(d["smoke"], 2 * w)


def test_vtu_particles(tmp_path):
    rng = np.random.default_rng(1)
    pos = rng.random((9, 3)).astype(np.float32)
    io.write_to_vtks_lp(pos, {"r": np.arange(9, dtype=np.float32)}, {"v": -pos}, os.path.join(str(tmp_path), "field_003"))
    d = io.read_vtu(os.path.join(str(tmp_path), "field_003.vtu"))
    assert np.array_equal(d["points"], pos)
    assert np.array_equal(d["connectivity"], np.arange(9)) and np.array_equal(d["offsets"], np.arange(1, 10))
    assert np.all(d["types"] == 1)
    assert np.array_equal(d["r"], np.arange(9)) and np.array_equal(d["v"], -pos)
    io.write_to_vtks_lp(pos, {}, {}, os.path.join(str(tmp_path), "field_004"))   # as write_w_and_laden_particles
    assert np.array_equal(io.read_vtu(os.path.join(str(tmp_path), "field_004.vtu"))["points"], pos)


def test_checkpoint_names_and_state_archive(tmp_path):
    d = str(tmp_path)
    assert [os.path.basename(p) for p in io.velocity_paths(d, 40)] == [
        "vel_x_numpy_40.npy", "vel_y_numpy_40.npy", "vel_z_numpy_40.npy"]
    assert os.path.basename(io.smoke_path(d, 40)) == "smoke_numpy_40.npy"
    arrays = {"psi_garm": np.random.rand(3, 4, 5, 6).astype(np.float32), "mask": np.arange(6, dtype=np.int32)}
    io.save_state(io.state_path(d, 40), arrays, {"i": 17, "sub_t": 0.125, "frame_idx": 3})
    a, c = io.load_state(io.state_path(d, 40))
    assert c == {"i": 17, "sub_t": 0.125, "frame_idx": 3}
    assert set(a) == set(arrays) and all(np.array_equal(a[k], arrays[k]) and a[k].dtype == arrays[k].dtype
                                         for k in arrays)
