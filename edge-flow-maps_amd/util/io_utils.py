"""Frame and checkpoint files of the reference's run loop (util/io_utils.py,
run_edge.py:557-581, 719-737), written with numpy only.

VTK: the reference writes ImageData (.vti) through the `vtk` package and
particle UnstructuredGrid (.vtu) files through `pyevtk`; neither is a
dependency here.  These writers emit the same VTK XML datasets (same file
names, array names, dimensions, origin, spacing, point ordering and
component layout) as uncompressed raw appended data with UInt64 block
headers, which ParaView and VTK read; `read_vti` / `read_vtu` read them back.

Checkpoints: `vel_{x,y,z}_numpy_<frame>.npy` as the reference saves them,
plus `smoke_numpy_<frame>.npy`, which the reference loads on resume but never
writes, and `edge_state_<frame>.npz`, the whole flow-map state (maps, initial
fields, step counters), so a run resumes mid-reinit.  Everything is plain
arrays: no pickles are written or read.
"""
import os
import re
import shutil
import xml.etree.ElementTree as ET

import numpy as np

_VTK_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32",
              np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def remove_everything_in(folder):
    """util/io_utils.py:16-25."""
    for name in os.listdir(folder):
        path = os.path.join(folder, name)
        if os.path.isfile(path) or os.path.islink(path):
            os.unlink(path)
        elif os.path.isdir(path):
            shutil.rmtree(path)


# ---------------------------------------------------------------------------
# VTK XML writers (raw appended data)
# ---------------------------------------------------------------------------
class _Appended:
    """Collects DataArray blocks; offsets count the UInt64 size headers."""

    def __init__(self):
        self.blocks = []
        self.offset = 0

    def add(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.dtype.byteorder == ">":
            arr = arr.astype(arr.dtype.newbyteorder("<"))
        off = self.offset
        self.blocks.append(arr)
        self.offset += 8 + arr.nbytes
        return off

    def write(self, f):
        f.write(b"  <AppendedData encoding=\"raw\">\n   _")
        for a in self.blocks:
            f.write(np.uint64(a.nbytes).tobytes())
            f.write(a.tobytes())
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def _data_array(name, arr, ncomp, app):
    t = _VTK_TYPES[np.dtype(arr.dtype)]
    off = app.add(arr)
    comp = f" NumberOfComponents=\"{ncomp}\"" if ncomp > 1 else ""
    return f"        <DataArray type=\"{t}\" Name=\"{name}\"{comp} format=\"appended\" offset=\"{off}\"/>\n"


def write_vti_points(path, dims, arrays, scalars):
    """An ImageData of dims (origin 0, spacing 1) with point arrays already in VTK
    point order: arrays = [(name, (n_points,) or (n_points, ncomp) array)]."""
    _write_image(path, dims, arrays, scalars)


def _write_image(path, dims, arrays, scalars):
    """arrays: list of (name, (n_points, ncomp) or (n_points,) array in x-fastest order)."""
    ext = f"0 {dims[0] - 1} 0 {dims[1] - 1} 0 {dims[2] - 1}"
    app = _Appended()
    body = ""
    for name, a in arrays:
        body += _data_array(name, a, a.shape[1] if a.ndim == 2 else 1, app)
    head = ("<?xml version=\"1.0\"?>\n"
            "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
            f"  <ImageData WholeExtent=\"{ext}\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
            f"    <Piece Extent=\"{ext}\">\n"
            f"      <PointData Scalars=\"{scalars}\">\n{body}      </PointData>\n"
            "      <CellData>\n      </CellData>\n"
            "    </Piece>\n  </ImageData>\n")
    with open(path, "wb") as f:
        f.write(head.encode())
        app.write(f)


def write_vtks(w_numpy, smoke_numpy, outdir, i):
    """util/io_utils.py:28-50: point data "vorticity" (the scalars) and a
    one-channel "smoke" on an ImageData of w's shape, origin 0, spacing 1."""
    data = np.squeeze(w_numpy)
    smoke = np.squeeze(smoke_numpy)
    arrays = [("vorticity", data.ravel(order="F")), ("smoke", smoke.ravel(order="F"))]
    _write_image(os.path.join(outdir, "field_{:04d}.vti".format(i)), data.shape, arrays, "vorticity")


def write_vtks_4channel_smoke(w_numpy, smoke_numpy, outdir, i):
    """util/io_utils.py:53-75: as write_vtks with a 4-component "smoke" (the
    (nx, ny, nz, 4) array reshaped to (-1, 4) in Fortran order)."""
    data = np.squeeze(w_numpy)
    smoke = np.asarray(smoke_numpy).reshape(-1, 4, order="F")
    arrays = [("vorticity", data.ravel(order="F")), ("smoke", smoke)]
    _write_image(os.path.join(outdir, "field_{:04d}.vti".format(i)), data.shape, arrays, "vorticity")


def write_to_vtks_lp(pos, scalar_data, vector_data, file_path):
    """util/io_utils.py:83-150: particles as an UnstructuredGrid of vertices
    (file_path + ".vtu"), point data from scalar_data / vector_data (name ->
    (n,) / (n, 3) arrays), the first sorted key of each marked active."""
    pos = np.asarray(pos)
    n = pos.shape[0]
    app = _Appended()
    pts = _data_array("points", np.ascontiguousarray(pos[:, :3]), 3, app)
    conn = _data_array("connectivity", np.arange(n, dtype=np.int32), 1, app)
    offs = _data_array("offsets", np.arange(1, n + 1, dtype=np.int32), 1, app)
    types = _data_array("types", np.full(n, 1, np.uint8), 1, app)   # VTK_VERTEX
    skeys, vkeys = sorted(scalar_data), sorted(vector_data)
    attrs = (f" Scalars=\"{skeys[0]}\"" if skeys else "") + (f" Vectors=\"{vkeys[0]}\"" if vkeys else "")
    if not skeys and not vkeys:
        attrs = " Scalars=\"scalar\""
    pdata = "".join(_data_array(k, np.asarray(scalar_data[k]), 1, app) for k in skeys)
    pdata += "".join(_data_array(k, np.asarray(vector_data[k]), 3, app) for k in vkeys)
    head = ("<?xml version=\"1.0\"?>\n"
            "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
            "header_type=\"UInt64\">\n  <UnstructuredGrid>\n"
            f"    <Piece NumberOfPoints=\"{n}\" NumberOfCells=\"{n}\">\n"
            f"      <Points>\n{pts}      </Points>\n"
            f"      <Cells>\n{conn}{offs}{types}      </Cells>\n"
            f"      <PointData{attrs}>\n{pdata}      </PointData>\n"
            "    </Piece>\n  </UnstructuredGrid>\n")
    with open(file_path + ".vtu", "wb") as f:
        f.write(head.encode())
        app.write(f)


# ---------------------------------------------------------------------------
# Readers of the files above
# ---------------------------------------------------------------------------
def _read_xml(path):
    raw = open(path, "rb").read()
    cut = raw.index(b"<AppendedData")
    root = ET.fromstring(raw[:cut] + b"</VTKFile>")
    start = raw.index(b"_", raw.index(b">", cut)) + 1
    if root.get("header_type") != "UInt64" or root.get("byte_order") != "LittleEndian":
        raise ValueError(f"{path}: expected UInt64 headers, little endian")

    def array(el):
        off = start + int(el.get("offset"))
        nbytes = int(np.frombuffer(raw, np.uint64, 1, off)[0])
        a = np.frombuffer(raw, _NP_TYPES[el.get("type")], nbytes // _NP_TYPES[el.get("type")].itemsize, off + 8)
        nc = int(el.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    return root, array


def read_vti(path):
    """Point arrays of a .vti as (nx, ny, nz) / (nx, ny, nz, nc) arrays, plus
    "_dims", "_origin", "_spacing" and "_scalars" (the active scalars' name)."""
    root, array = _read_xml(path)
    img = root.find("ImageData")
    e = [int(v) for v in img.get("WholeExtent").split()]
    dims = (e[1] - e[0] + 1, e[3] - e[2] + 1, e[5] - e[4] + 1)
    pd = img.find("Piece/PointData")
    out = {"_dims": dims, "_origin": tuple(float(v) for v in img.get("Origin").split()),
           "_spacing": tuple(float(v) for v in img.get("Spacing").split()), "_scalars": pd.get("Scalars")}
    for el in pd.findall("DataArray"):
        a = array(el)
        if a.ndim == 2:
            out[el.get("Name")] = a.reshape(dims + (a.shape[1],), order="F")
        else:
            out[el.get("Name")] = a.reshape(dims, order="F")
    return out


def read_vtu(path):
    """Points (n, 3), cells (connectivity, offsets, types) and point arrays of a .vtu."""
    root, array = _read_xml(path)
    piece = root.find("UnstructuredGrid/Piece")
    out = {"points": array(piece.find("Points/DataArray"))}
    for el in piece.findall("Cells/DataArray"):
        out[el.get("Name")] = array(el)
    for el in piece.findall("PointData/DataArray"):
        out[el.get("Name")] = array(el)
    return out


# ---------------------------------------------------------------------------
# Checkpoints
# ---------------------------------------------------------------------------
def velocity_paths(ckptdir, frame):
    """run_edge.py:557-562, 576-581: vel_{x,y,z}_numpy_<frame>.npy."""
    return [os.path.join(ckptdir, f"vel_{a}_numpy_{frame}.npy") for a in "xyz"]


def smoke_path(ckptdir, frame):
    """run_edge.py:565-566 (loaded on resume by the reference, never written there)."""
    return os.path.join(ckptdir, f"smoke_numpy_{frame}.npy")


def state_path(ckptdir, frame):
    """The whole flow-map state of a run (no counterpart in the reference)."""
    return os.path.join(ckptdir, f"edge_state_{frame}.npz")


_STATE_KEY = re.compile(r"^[A-Za-z_][A-Za-z0-9_]*$")


def save_state(path, arrays, counters):
    """arrays: name -> numpy array; counters: name -> int/float.  One .npz, no pickles."""
    out = {}
    for k, v in arrays.items():
        assert _STATE_KEY.match(k), k
        out["f_" + k] = np.ascontiguousarray(v)
    for k, v in counters.items():
        assert _STATE_KEY.match(k), k
        out["c_" + k] = np.asarray(v, np.float64 if isinstance(v, float) else np.int64)
    np.savez(path, **out)


def load_state(path):
    """(arrays, counters) written by save_state."""
    arrays, counters = {}, {}
    with np.load(path, allow_pickle=False) as z:
        for k in z.files:
            v = z[k]
            if k.startswith("f_"):
                arrays[k[2:]] = v
            elif k.startswith("c_"):
                counters[k[2:]] = v.item()
    return arrays, counters
